// metriko (MI355X build) -- host-side parallel helpers (OpenMP).
#ifndef METRIKO_PARALLEL_H
#define METRIKO_PARALLEL_H
#include <algorithm>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace metriko {
    // Sorts v ascending: chunks sorted on their own threads, then merged pairwise in log2(chunks)
    // rounds.  A sorted sequence is unique up to the order of equal elements, which are
    // indistinguishable, so the result is what std::sort gives.
    template<typename T>
    void parallel_sort(std::vector<T>& v) {
        const size_t n = v.size();
        int nch = 1;
#ifdef _OPENMP
        if (n > 200000) nch = std::max(1, std::min(omp_get_max_threads(), int(n / 50000)));
#endif
        if (nch <= 1) { std::sort(v.begin(), v.end()); return; }
        std::vector<size_t> b(size_t(nch) + 1);
        for (int q = 0; q <= nch; q++) b[size_t(q)] = n * size_t(q) / size_t(nch);
#pragma omp parallel for schedule(static, 1)
        for (int q = 0; q < nch; q++) std::sort(v.begin() + long(b[size_t(q)]), v.begin() + long(b[size_t(q) + 1]));
        std::vector<T> tmp(n);
        std::vector<T>* src = &v;
        std::vector<T>* dst = &tmp;
        for (size_t w = 1; w < size_t(nch); w *= 2) {
#pragma omp parallel for schedule(dynamic, 1)
            for (int q = 0; q < nch; q += int(2 * w)) {
                const size_t lo = b[size_t(q)], mid = b[std::min(size_t(q) + w, size_t(nch))],
                             hi = b[std::min(size_t(q) + 2 * w, size_t(nch))];
                std::merge(src->begin() + long(lo), src->begin() + long(mid), src->begin() + long(mid),
                           src->begin() + long(hi), dst->begin() + long(lo));
            }
            std::swap(src, dst);
        }
        if (src != &v) v.swap(*src);
    }
}

#endif
