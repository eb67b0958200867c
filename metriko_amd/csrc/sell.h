// Shared pieces of the sparse kernels: deterministic grid reductions, the hybrid
// SELL-C-sigma device view and its row kernel, launch helpers and the SpMV timer.
#pragma once
#include <algorithm>
#include <string>
#include "mk_internal.h"

namespace mk {

static constexpr int kMaxWaves = 16;        // 1024 threads / 64
static constexpr size_t kMaxParts = 1 << 17;  // partial-sum slots per reduced quantity
static constexpr int kMaxK = 4;

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

// Writes the block partials and returns true in the last block to arrive, with the
// grand totals in tot (valid in all threads of that block).
template <int K>
__device__ __forceinline__ bool partials_last(double (&v)[K], double* partials, unsigned* counter, double (&tot)[K]) {
    __shared__ int is_last;
    __shared__ double shtot[K];
    block_sum<K>(v);
    const unsigned G = gridDim.x;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) partials[k * kMaxParts + blockIdx.x] = v[k];
        __threadfence();
        unsigned t = atomicAdd(counter, 1u);
        is_last = (t == G - 1);
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        acc[k] = 0.0;
        const volatile double* pk = partials + k * kMaxParts;
        for (unsigned i = threadIdx.x; i < G; i += blockDim.x) acc[k] += pk[i];
    }
    block_sum<K>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) shtot[k] = acc[k];
        *counter = 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) tot[k] = shtot[k];
    return true;
}

// ---------------------------------------------------------------------------------------
// SELL SpMV.  T = float/double; CPLX: values/vectors interleaved (re, im).
// MODE 0: y = A x; 1: y = A x, out = Re(x^H y); 2: y = b - A x, out = |y|^2;
// 3: y = b - A x; 4: y += A x.  skip (optional): device flag, nonzero = do nothing.
//
// Hybrid layout: the first nb_sell blocks run the SELL part (short rows, thread per row);
// the remaining blocks give each long row a whole wavefront (lanes stride the row, then a
// shuffle reduction), so a few very long rows never serialise one lane.
template <typename T>
struct SellDev {
    int n_short, C, nb_sell, nlong;
    const int64_t* slice_ptr;
    const int* slice_w;
    const int* col;
    const T* val;
    const int* perm;
    const int* long_rows;
    const int64_t* long_ptr;
    const int* long_col;
    const T* long_val;
};

template <typename T>
__device__ __forceinline__ T wave_sum_t(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Computes one row of A x.  Returns true in the thread that owns the result (row, ar, ai).
template <typename T, bool CPLX>
__device__ __forceinline__ bool spmv_row(const SellDev<T>& A, const T* __restrict__ x, int& row, T& ar, T& ai) {
    ar = 0;
    ai = 0;
    if (int(blockIdx.x) < A.nb_sell) {
        const int r = blockIdx.x * blockDim.x + threadIdx.x;
        if (r >= A.n_short) return false;
        const int s = r / A.C, lane = r - s * A.C;
        const int64_t base = A.slice_ptr[s] + lane;
        const int w = A.slice_w[s];
        row = A.perm ? A.perm[r] : r;
        for (int j = 0; j < w; j++) {
            const int64_t e = base + int64_t(j) * A.C;
            const int c = A.col[e];
            if constexpr (CPLX) {
                const T vr = A.val[2 * e], vi = A.val[2 * e + 1];
                const T xr = x[2 * c], xi = x[2 * c + 1];
                ar += vr * xr - vi * xi;
                ai += vr * xi + vi * xr;
            } else {
                ar += A.val[e] * x[c];
            }
        }
        return true;
    }
    const int w = (int(blockIdx.x) - A.nb_sell) * int(blockDim.x >> 6) + int(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w < A.nlong) {  // uniform inside the wavefront
        for (int64_t k = A.long_ptr[w] + lane; k < A.long_ptr[w + 1]; k += 64) {
            const int c = A.long_col[k];
            if constexpr (CPLX) {
                const T vr = A.long_val[2 * k], vi = A.long_val[2 * k + 1];
                const T xr = x[2 * c], xi = x[2 * c + 1];
                ar += vr * xr - vi * xi;
                ai += vr * xi + vi * xr;
            } else {
                ar += A.long_val[k] * x[c];
            }
        }
        ar = wave_sum_t(ar);
        if constexpr (CPLX) ai = wave_sum_t(ai);
        row = A.long_rows[w];
        return lane == 0;
    }
    return false;
}


static inline unsigned blocks_for(size_t len, int bs = 256) { return unsigned((len + bs - 1) / bs); }

hipEvent_t next_event(mk_ctx* c);
void timing_flush(mk_ctx* c);

// Counts an SpMV launch and, when timing is on, brackets it with events attributed to
// "<operator label>:<kernel>:<fp32|fp64>".
struct SpmvTimer {
    mk_ctx* c;
    SpmvTimer(mk_ctx* c_, double bytes, const mk_op* A = nullptr, const char* kernel = "spmv", bool fp32 = false)
        : c(c_) {
        c->spmv_calls++;
        c->spmv_bytes += bytes;
        if (c->time_spmv) {
            if (c->ev_used + 2 > 4096) timing_flush(c);
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
            c->ev_key.push_back(key);
            MK_HIP(hipEventRecord(next_event(c), c->stream));
        }
    }
    ~SpmvTimer() {
        if (c->time_spmv) (void)hipEventRecord(next_event(c), c->stream);
    }
};

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
    if constexpr (std::is_same<T, double>::value) d.long_val = A->long_val64.p;
    else d.long_val = A->long_val32.p;
    const unsigned nb_long = blocks_for(size_t(A->nlong), block / 64);
    *grid = std::max(1u, unsigned(d.nb_sell) + nb_long);
    return d;
}

}  // namespace mk
