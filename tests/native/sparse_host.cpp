// Host sparse kernels of include/metriko/core/common/sparse.h against plain restatements:
//  * transpose with 32-bit and 64-bit offset counters: the same CSR, and the same as a
//    one-thread counting transpose (rows ascending inside each column);
//  * spgemm(At, A) for complex values, rows both below and above the sort-path threshold:
//    every entry bit-identical to the sum of std::complex products taken in the order of the
//    inner index (first product, then +=), and mulv equal to std::complex's product bit for bit.
// Prints "ok <checks>" or the first mismatch; exit status 1 on a mismatch.
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <random>

#include "metriko/core/common/sparse.h"

using cd = std::complex<double>;
using metriko::Csr;

static bool same(const cd& a, const cd& b) { return std::memcmp(&a, &b, sizeof(cd)) == 0; }

template <typename T>
static bool same_csr(const Csr<T>& a, const Csr<T>& b) {
    if (a.nr != b.nr || a.nc != b.nc || a.ptr.size() != b.ptr.size() || a.col.size() != b.col.size()) return false;
    for (size_t i = 0; i < a.ptr.size(); i++) if (a.ptr[i] != b.ptr[i]) return false;
    for (size_t k = 0; k < a.col.size(); k++)
        if (a.col[k] != b.col[k] || std::memcmp(&a.val[k], &b.val[k], sizeof(T)) != 0) return false;
    return true;
}

int main(int argc, char** argv) {
    const int seed = argc > 1 ? std::atoi(argv[1]) : 1;
    std::mt19937_64 g{uint64_t(seed)};
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    long checks = 0;
    // a tall matrix with three entries in most rows and a few rows that couple many columns, so
    // that At * A has rows on both sides of the sort-path threshold
    const int nr = 120000, nc = 40000;
    Csr<cd> A(nr, nc);
    std::vector<std::vector<int>> rows(nr);
    std::uniform_int_distribution<int> C(0, nc - 1);
    for (int r = 0; r < nr; r++) {
        const int m = (r % 997 == 0) ? 40 : 3;
        std::vector<int> cs;
        const int base = C(g);
        for (int q = 0; q < m; q++) cs.push_back((base + q * 7 + (q ? C(g) % 50 : 0)) % nc);
        std::sort(cs.begin(), cs.end());
        cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
        rows[r] = cs;
        A.ptr[r + 1] = A.ptr[r] + int64_t(cs.size());
    }
    A.col.resize(size_t(A.ptr[nr]));
    A.val.resize(size_t(A.ptr[nr]));
    for (int r = 0; r < nr; r++)
        for (size_t q = 0; q < rows[r].size(); q++) {
            A.col[size_t(A.ptr[r]) + q] = rows[r][q];
            A.val[size_t(A.ptr[r]) + q] = cd(U(g), U(g));
        }
    // transposes
    const Csr<cd> T32 = A.transpose_impl<int>(), T64 = A.transpose_impl<int64_t>(), Tdef = A.transpose();
    Csr<cd> Tref(nc, nr);
    {
        std::vector<int64_t> cnt(size_t(nc) + 1, 0);
        for (int64_t k = 0; k < A.nonZeros(); k++) cnt[size_t(A.col[k]) + 1]++;
        for (int j = 0; j < nc; j++) cnt[size_t(j) + 1] += cnt[size_t(j)];
        Tref.ptr.assign(cnt.begin(), cnt.end());
        Tref.col.resize(size_t(cnt[nc]));
        Tref.val.resize(size_t(cnt[nc]));
        for (int r = 0; r < nr; r++)
            for (int64_t k = A.ptr[r]; k < A.ptr[r + 1]; k++) {
                const int64_t p = cnt[size_t(A.col[k])]++;
                Tref.col[size_t(p)] = r;
                Tref.val[size_t(p)] = A.val[k];
            }
    }
    checks += 3;
    if (!same_csr(T32, Tref) || !same_csr(T64, Tref) || !same_csr(Tdef, Tref)) {
        std::printf("transpose mismatch\n");
        return 1;
    }
    // product At * A against ordered sums of std::complex products
    const Csr<cd> P = metriko::spgemm(Tref, A);
    int64_t longrows = 0;
    for (int i = 0; i < nc; i++) {
        std::map<int, cd> ref;
        int64_t work = 0;
        for (int64_t ka = Tref.ptr[i]; ka < Tref.ptr[i + 1]; ka++) {
            const int r = Tref.col[ka];
            const cd a = Tref.val[ka];
            for (int64_t kb = A.ptr[r]; kb < A.ptr[r + 1]; kb++, work++) {
                const cd p = a * A.val[kb];
                auto it = ref.find(A.col[kb]);
                if (it == ref.end()) ref.emplace(A.col[kb], p);
                else it->second += p;
            }
        }
        longrows += work > 32 ? 1 : 0;
        if (int64_t(ref.size()) != P.ptr[i + 1] - P.ptr[i]) {
            std::printf("row %d: %zu entries, spgemm %lld\n", i, ref.size(), (long long)(P.ptr[i + 1] - P.ptr[i]));
            return 1;
        }
        int64_t k = P.ptr[i];
        for (const auto& e : ref) {
            checks++;
            if (P.col[k] != e.first || !same(P.val[k], e.second)) {
                std::printf("row %d col %d: spgemm (%.17g, %.17g), ordered sum (%.17g, %.17g)\n", i, e.first,
                            P.val[k].real(), P.val[k].imag(), e.second.real(), e.second.imag());
                return 1;
            }
            k++;
        }
    }
    if (longrows == 0) {
        std::printf("no row above the sort-path threshold\n");
        return 1;
    }
    // mulv against the library product on random finite operands of mixed magnitudes
    std::uniform_int_distribution<int> E(-300, 300);
    for (int t = 0; t < 1000000; t++) {
        const cd a(std::ldexp(U(g), E(g)), std::ldexp(U(g), E(g))), b(std::ldexp(U(g), E(g)), std::ldexp(U(g), E(g)));
        const cd p = a * b, q = metriko::mulv(a, b);
        if (!(std::isfinite(p.real()) && std::isfinite(p.imag()))) continue;
        checks++;
        if (!same(p, q)) {
            std::printf("mulv (%.17g, %.17g) * (%.17g, %.17g): (%.17g, %.17g) vs (%.17g, %.17g)\n", a.real(), a.imag(),
                        b.real(), b.imag(), q.real(), q.imag(), p.real(), p.imag());
            return 1;
        }
    }
    std::printf("ok %ld\n", checks);
    return 0;
}
