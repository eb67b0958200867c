// Shared pieces of the sparse kernels: deterministic grid reductions, the hybrid
// SELL-C-sigma device view and its row kernel, launch helpers and the SpMV timer.
#pragma once
#include <algorithm>
#include <string>
#include <hip/hip_ext.h>
#include "mk_internal.h"

namespace mk {

static constexpr int kMaxWaves = 16;        // 1024 threads / 64
static constexpr size_t kMaxParts = 1 << 17;  // partial-sum slots per reduced quantity
static constexpr int kMaxK = 4;
// Longest row kept in the SELL part; longer rows go to the medium (lane group per row) or
// long (wavefront per row) parts at upload, so SELL slice widths never exceed it.
static constexpr int kSellMaxLen = 12;
// Entries per segment of a long row in the multi-column kernels (a 1,330-entry row: 11
// wavefronts instead of one serial chain)
static constexpr int kLongSeg = 128;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-reduce K values; thread 0 gets the block totals.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K]) {
    __shared__ double sh[K][kMaxWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0) sh[k][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double s = 0.0;
            for (int w = 0; w < nw; w++) s += sh[k][w];
            v[k] = s;
        }
    }
    __syncthreads();
}

// Fence-free grid reduction in two launches: the working kernel leaves one partial sum per
// block (partials[k * kMaxParts + block]) and a one-block finishing kernel adds them in a fixed
// order after the kernel boundary.  No device-scope fence or ticket per block: on a chip whose
// L2 is private to each XCD a device-scope release writes back the block's L2, which costs far
// more than the extra launch in kernels that stream large vectors.
template <int K>
__device__ __forceinline__ void partials_write(double (&v)[K], double* partials) {
    block_sum<K>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) partials[k * kMaxParts + blockIdx.x] = v[k];
    }
}
// Sum of nparts partials per value, fixed order; every thread of the (single) block gets tot.
template <int K>
__device__ __forceinline__ void partials_sum(const double* __restrict__ partials, int nparts, double (&tot)[K]) {
    __shared__ double shtot[K];
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        acc[k] = 0.0;
        for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc[k] += partials[k * kMaxParts + i];
    }
    block_sum<K>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) shtot[k] = acc[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) tot[k] = shtot[k];
}

// The blocks of a launch go round-robin over the 8 XCDs, each with its own L2.  xcd_block maps
// the hardware block id b in [0, nb) to a logical block id so that every XCD works through one
// contiguous range of logical blocks: neighbouring rows then find the x entries they gather
// in their own XCD's L2.  A bijection for any nb.
static constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b % kXcd, q = nb / kXcd, r = nb % kXcd;
    return x * q + (x < r ? x : r) + b / kXcd;
}

template <typename T>
struct SellDev {
    int n_short, C, nb_sell, nlong;
    int nvb;  // virtual blocks: nb_sell SELL blocks, nb_mid medium-row blocks, then long-row blocks
    int nb_mid, nmid, mid_lanes;
    const int* mid_rows;
    const int64_t* mid_ptr;
    const int* mid_col;
    const T* mid_val;
    const int64_t* slice_ptr;
    const int* slice_w;
    const int* col;
    const T* val;
    const int* perm;
    const int* long_rows;
    const int64_t* long_ptr;
    const int* long_col;
    const T* long_val;
    int nseg;                // long-row segments (multi-column kernels)
    const int64_t* seg_beg;
    const int64_t* seg_end;
    const int* lseg_ptr;
    T* seg_part;
};

template <typename T>
__device__ __forceinline__ T wave_sum_t(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Vector access of the row kernels: a gather g(c, xr, xi) loads entry c of the vector the
// operator multiplies (xi unused for real operators).
template <typename T, bool CPLX>
struct PlainGather {
    const T* __restrict__ x;
    __device__ __forceinline__ void operator()(int c, T& xr, T& xi) const {
        if constexpr (CPLX) { xr = x[2 * c]; xi = x[2 * c + 1]; }
        else { xr = x[c]; xi = T(0); }
    }
};

// No row-local operands to fetch early.
struct NoPrefetch {
    __device__ __forceinline__ void operator()(int) {}
};

// Computes one row of A x for virtual block vb (uniform in the block), x given by a gather.
// Returns true in the thread that owns the result (row, ar, ai).  pre(row) is called as soon
// as the row is known -- for the SELL rows before their column loads -- so that the caller's
// row-local operands (right-hand side, diagonal, vectors updated in place) are in flight
// together with the gathers instead of after them.
template <typename T, bool CPLX, typename G, typename P>
__device__ __forceinline__ bool spmv_row_gp(const SellDev<T>& A, const G& g, int vb, int& row, T& ar, T& ai, P& pre) {
    ar = 0;
    ai = 0;
    // the long rows' blocks come first: their wavefronts run the longest chains, so they start
    // with the launch and overlap the SELL rows instead of trailing them
    {
        const int nlb = A.nvb - A.nb_sell - A.nb_mid;
        vb = vb < nlb ? A.nb_sell + A.nb_mid + vb : vb - nlb;
    }
    if (vb < A.nb_sell) {
        const int r = vb * blockDim.x + threadIdx.x;
        if (r >= A.n_short) return false;
        const int s = r / A.C, lane = r - s * A.C;
        const int64_t base = A.slice_ptr[s] + lane;
        const int w = A.slice_w[s];
        row = A.perm ? A.perm[r] : r;
        pre(row);
        // every column index first (w <= kSellMaxLen), then the gathers: two memory latencies
        // per row instead of two per entry
        int ci[kSellMaxLen];
#pragma unroll
        for (int j = 0; j < kSellMaxLen; j++)
            if (j < w) ci[j] = A.col[base + int64_t(j) * A.C];
#pragma unroll
        for (int j = 0; j < kSellMaxLen; j++) {
            if (j < w) {
                const int64_t e = base + int64_t(j) * A.C;
                T xr, xi;
                g(ci[j], xr, xi);
                if constexpr (CPLX) {
                    const T vr = A.val[2 * e], vi = A.val[2 * e + 1];
                    ar += vr * xr - vi * xi;
                    ai += vr * xi + vi * xr;
                } else {
                    ar += A.val[e] * xr;
                }
            }
        }
        return true;
    }
    if (vb < A.nb_sell + A.nb_mid) {  // medium rows: a group of mid_lanes lanes per row
        const int L = A.mid_lanes;
        const int t = (vb - A.nb_sell) * int(blockDim.x) + int(threadIdx.x);
        const int m = t / L, sub = t - m * L;
        if (m < A.nmid) {
            for (int64_t k = A.mid_ptr[m] + sub; k < A.mid_ptr[m + 1]; k += L) {
                const int c = A.mid_col[k];
                T xr, xi;
                g(c, xr, xi);
                if constexpr (CPLX) {
                    const T vr = A.mid_val[2 * k], vi = A.mid_val[2 * k + 1];
                    ar += vr * xr - vi * xi;
                    ai += vr * xi + vi * xr;
                } else {
                    ar += A.mid_val[k] * xr;
                }
            }
        }
        // reduce inside the group (every lane takes part; L divides 64)
        for (int off = L >> 1; off > 0; off >>= 1) {
            ar += __shfl_xor(ar, off, 64);
            if constexpr (CPLX) ai += __shfl_xor(ai, off, 64);
        }
        if (m >= A.nmid || sub != 0) return false;
        row = A.mid_rows[m];
        pre(row);
        return true;
    }
    const int w = (vb - A.nb_sell - A.nb_mid) * int(blockDim.x >> 6) + int(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w < A.nlong) {  // uniform inside the wavefront
        // a lane's entries k = lane, lane + 64, ...: U gathers in flight, summed in order (two
        // for complex values, whose registers would otherwise lower the occupancy of the whole
        // kernel, long rows or not)
        constexpr int U = CPLX ? 2 : 8;
        const int64_t kend = A.long_ptr[w + 1];
        int64_t k = A.long_ptr[w] + lane;
        auto add = [&](int64_t q, T xr, T xi) {
            if constexpr (CPLX) {
                const T vr = A.long_val[2 * q], vi = A.long_val[2 * q + 1];
                ar += vr * xr - vi * xi;
                ai += vr * xi + vi * xr;
            } else {
                ar += A.long_val[q] * xr;
            }
        };
        for (; k + (U - 1) * 64 < kend; k += U * 64) {
            T xr[U], xi[U];
#pragma unroll
            for (int u = 0; u < U; u++) g(A.long_col[k + 64 * u], xr[u], xi[u]);
#pragma unroll
            for (int u = 0; u < U; u++) add(k + 64 * u, xr[u], xi[u]);
        }
        for (; k < kend; k += 64) {
            T xr, xi;
            g(A.long_col[k], xr, xi);
            add(k, xr, xi);
        }
        ar = wave_sum_t(ar);
        if constexpr (CPLX) ai = wave_sum_t(ai);
        row = A.long_rows[w];
        if (lane == 0) pre(row);
        return lane == 0;
    }
    return false;
}
template <typename T, bool CPLX, typename G>
__device__ __forceinline__ bool spmv_row_g(const SellDev<T>& A, const G& g, int vb, int& row, T& ar, T& ai) {
    NoPrefetch none;
    return spmv_row_gp<T, CPLX>(A, g, vb, row, ar, ai, none);
}


template <typename T, bool CPLX>
__device__ __forceinline__ bool spmv_row(const SellDev<T>& A, const T* __restrict__ x, int vb, int& row, T& ar,
                                         T& ai) {
    return spmv_row_g<T, CPLX>(A, PlainGather<T, CPLX>{x}, vb, row, ar, ai);
}

static inline unsigned blocks_for(size_t len, int bs = 256) { return unsigned((len + bs - 1) / bs); }
// grid of the grid-stride reduction kernels: few enough blocks that the per-block arrival
// (partial store, device fence, ticket) stays a small part of the kernel
// Grid cap of the reducing kernels (grid-stride loops; partial sums per block): large enough
// to keep every CU's memory pipeline busy, small enough for a short partial-sum pass.
static inline unsigned red_grid() {
    static const unsigned g = [] {
        const char* s = std::getenv("METRIKO_RED_GRID");
        const long v = s ? std::atol(s) : 0;
        return v >= 64 && v <= 16384 ? unsigned(v) : 512u;
    }();
    return g;
}
static inline unsigned rgrid(size_t len) { return std::max(1u, std::min(blocks_for(len), red_grid())); }
// Grid cap of the fence-free two-launch reductions (partials_write / partials_sum).
static inline unsigned fin_grid() {
    static const unsigned g = [] {
        const char* s = std::getenv("METRIKO_FIN_GRID");
        const long v = s ? std::atol(s) : 0;
        return v >= 64 && v <= long(kMaxParts) ? unsigned(v) : 4096u;
    }();
    return g;
}
static inline unsigned fgrid(size_t len) { return std::max(1u, std::min(blocks_for(len), fin_grid())); }

hipEvent_t next_event(mk_ctx* c);
void timing_flush(mk_ctx* c);

// Counts an SpMV launch and, when timing is on, brackets it with events attributed to
// "<operator label>:<kernel>:<fp32|fp64>".
struct SpmvTimer {
    mk_ctx* c;
    // with timing on: start / end stamps for each of up to three launches of one timed
    // quantity (MK_TLAUNCH part 0, 1, 2); the durations are added (the intervals between the
    // launches are not counted)
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int parts = 1;
    mutable unsigned used = 0;
    SpmvTimer(mk_ctx* c_, double bytes, const mk_op* A = nullptr, const char* kernel = "spmv", bool fp32 = false,
              int nparts = 1)
        : c(c_), parts(nparts) {
        c->spmv_calls++;
        c->spmv_bytes += bytes;
        if (c->time_spmv) {
            if (c->ev_used + 2 * size_t(parts) > 4096) timing_flush(c);
            std::string key = std::string(A ? A->label : "op") + ":" + kernel + ":" + (fp32 ? "fp32" : "fp64");
            auto it = std::find_if(c->kstats.begin(), c->kstats.end(), [&](const auto& p) { return p.first == key; });
            if (it == c->kstats.end()) {
                mk_ctx::KStat s;
                s.bytes_per_call = bytes;
                s.rows = A ? A->n : 0;
                c->kstats.emplace_back(key, s);
                it = c->kstats.end() - 1;
            }
            it->second.bytes += bytes;
            for (int q = 0; q < parts; q++) {
                c->ev_key.push_back(q ? "+" + key : key);  // "+": time added to the call, no extra call
                ev[2 * q] = next_event(c);
                ev[2 * q + 1] = next_event(c);
            }
        }
    }
    ~SpmvTimer() {
        for (int q = 0; q < parts; q++)  // a part that was not launched: an empty interval
            if (ev[2 * q] && !(used & (1u << q))) {
                (void)hipEventRecord(ev[2 * q], c->stream);
                (void)hipEventRecord(ev[2 * q + 1], c->stream);
            }
    }
};

// Launches kernel K on STREAM.  With timing on, the launch carries the timer's start and end
// stamps of part PART through hipExtLaunchKernelGGL: they are taken at the kernel's own start
// and end, so the timed interval holds no event packet.
#define MK_TLAUNCH(tm, PART, K, G, B, SH, STREAM, ...)                                                  \
    do {                                                                                                \
        if ((tm).ev[0]) {                                                                               \
            (tm).used |= 1u << (PART);                                                                  \
            hipExtLaunchKernelGGL(K, G, B, SH, STREAM, (tm).ev[2 * (PART)], (tm).ev[2 * (PART) + 1], 0, \
                                  __VA_ARGS__);                                                         \
        } else {                                                                                        \
            hipLaunchKernelGGL(K, G, B, SH, STREAM, __VA_ARGS__);                                       \
        }                                                                                               \
    } while (0)

template <typename T>
static SellDev<T> sell_dev(const mk_op* A, const T* val, int block, unsigned* grid) {
    SellDev<T> d;
    d.n_short = A->n_short;
    d.C = A->C;
    d.nb_sell = int(blocks_for(size_t(A->n_short), block));
    d.nlong = A->nlong;
    d.slice_ptr = A->slice_ptr.p;
    d.slice_w = A->slice_w.p;
    d.col = A->col.p;
    d.val = val;
    d.perm = A->perm.p;
    d.long_rows = A->long_rows.p;
    d.long_ptr = A->long_ptr.p;
    d.long_col = A->long_col.p;
    d.nseg = A->nseg;
    d.seg_beg = A->seg_beg.p;
    d.seg_end = A->seg_end.p;
    d.lseg_ptr = A->lseg_ptr.p;
    d.seg_part = reinterpret_cast<T*>(A->seg_part.p);
    if constexpr (std::is_same<T, double>::value) d.long_val = A->long_val64.p;
    else d.long_val = A->long_val32.p;
    d.nmid = A->nmid;
    d.mid_lanes = A->mid_lanes;
    d.nb_mid = int(blocks_for(size_t(A->nmid) * size_t(A->mid_lanes), block));
    d.mid_rows = A->mid_rows.p;
    d.mid_ptr = A->mid_ptr.p;
    d.mid_col = A->mid_col.p;
    if constexpr (std::is_same<T, double>::value) d.mid_val = A->mid_val64.p;
    else d.mid_val = A->mid_val32.p;
    const unsigned nb_long = blocks_for(size_t(A->nlong), block / 64);
    *grid = std::max(1u, unsigned(d.nb_sell) + unsigned(d.nb_mid) + nb_long);
    d.nvb = int(*grid);
    return d;
}

}  // namespace mk
