// Host unit test of the captured-graph keys (metriko_amd/csrc/graph_key.h): a hierarchy or a
// context rebuilt at the same host address must not match a graph captured through the old one.
#include <cstdio>
#include <cstdlib>
#include <set>
#include "graph_key.h"

struct FakeHierarchy {  // stands in for mk::Mg: an id taken at construction
    const uint64_t uid = mk::next_uid();
    char payload[512];
};

static int fails = 0;
#define EXPECT(c)                                                        \
    do {                                                                 \
        if (!(c)) { std::fprintf(stderr, "FAILED: %s (line %d)\n", #c, __LINE__); fails++; } \
    } while (0)

static std::vector<uintptr_t> key_for(uint64_t ctx_uid, const FakeHierarchy* g, uintptr_t x) {
    mk::GraphKeyParts k;
    k.ctx_uid = ctx_uid;
    k.mg_uid = g ? g->uid : 0;
    k.state = 0x1000;
    k.partials = 0x2000;
    k.scal = 0x3000;
    k.buffers = {4, x, 0x5000, 0x6000};
    k.batch = 16;
    return mk::make_graph_key(k);
}

int main() {
    // ids are unique and never 0
    std::set<uint64_t> seen;
    for (int i = 0; i < 1000; i++) {
        const uint64_t u = mk::next_uid();
        EXPECT(u != 0);
        EXPECT(seen.insert(u).second);
    }
    // rebuild a hierarchy until the allocator returns the freed address again
    const uint64_t ctx = mk::next_uid();
    FakeHierarchy* a = new FakeHierarchy();
    const uintptr_t addr_a = uintptr_t(a);
    const std::vector<uintptr_t> ka = key_for(ctx, a, 0x4000);
    EXPECT(ka == key_for(ctx, a, 0x4000));  // same hierarchy, same buffers: the graph is reused
    EXPECT(ka != key_for(ctx, a, 0x4100));  // another vector: a new capture
    delete a;
    bool reused = false;
    for (int i = 0; i < 64 && !reused; i++) {
        FakeHierarchy* b = new FakeHierarchy();
        if (uintptr_t(b) == addr_a) {
            reused = true;
            EXPECT(key_for(ctx, b, 0x4000) != ka);  // same address, new hierarchy: no match
        }
        delete b;
    }
    // another context with the same buffers: no match
    EXPECT(key_for(mk::next_uid(), nullptr, 0x4000) != key_for(ctx, nullptr, 0x4000));
    std::printf("graph key test: %s (allocator reused the address: %s)\n", fails ? "FAILED" : "ok",
                reused ? "yes" : "no");
    return fails ? 1 : 0;
}
