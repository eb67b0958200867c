// Keys of the captured PCG iteration graphs (host-only; no HIP types).
//
// A captured graph holds every device address its kernels were launched with: the solve's
// vectors, the operator's arrays, the multigrid hierarchy's level operators and buffers, and
// the context's reduction partials, CG state and scalars.  A replay is only valid while every
// one of them still names the same live allocation.  Raw host addresses of a hierarchy or a
// context are not enough for that: once one is freed, the allocator can hand the same address
// to the next one, whose buffers are elsewhere.  So each context and each hierarchy gets a
// process-unique id at creation, never reused, and the key carries the ids next to the raw
// vector addresses (which stay valid for the graph's owner, the operator, as long as the key
// matches: a reallocated vector changes its address or is at least as large as before).
#pragma once
#include <atomic>
#include <cstdint>
#include <vector>

namespace mk {

// process-unique, monotonically increasing; 0 is never returned
inline uint64_t next_uid() {
    static std::atomic<uint64_t> counter{1};
    return counter.fetch_add(1, std::memory_order_relaxed);
}

struct GraphKeyParts {
    uint64_t ctx_uid = 0;       // context whose partials / state / scalars the kernels use
    uint64_t mg_uid = 0;        // multigrid hierarchy the V-cycles run through (0: none)
    uintptr_t state = 0, partials = 0, scal = 0;
    std::vector<uintptr_t> buffers;  // solve vectors and value arrays, in launch-argument order
    int batch = 0;                   // iterations per replay
};

inline std::vector<uintptr_t> make_graph_key(const GraphKeyParts& k) {
    std::vector<uintptr_t> v;
    v.reserve(k.buffers.size() + 6);
    v.push_back(uintptr_t(k.ctx_uid));
    v.push_back(uintptr_t(k.mg_uid));
    v.push_back(k.state);
    v.push_back(k.partials);
    v.push_back(k.scal);
    v.insert(v.end(), k.buffers.begin(), k.buffers.end());
    v.push_back(uintptr_t(k.batch));
    return v;
}

}  // namespace mk
