// Multi-right-hand-side kernels: SpMM on the hybrid SELL layout and a batched PCG that
// solves A X = B for up to 64 right-hand sides at once (one matrix read serves every
// column).  The IGM rounding loop needs A0^{-1} u_j for every candidate constraint row u_j;
// those solves are independent of the rounding decisions, so they run as one batch.
//
// Block layout: MW right-hand sides (MW in {8, 16, 32, 64}) stored row-major, X[i*MW + j].
// SpMM: a wavefront covers 64/MW rows at a time, lane = sub*MW + j (sub = row of the group,
// j = column): each matrix entry is loaded once per row and shared by its MW lanes, and every
// access to X is one contiguous MW-wide segment.
//
// Batched PCG: independent CG recurrences per column (own alpha, beta, convergence flag),
// FP64 Krylov vectors, flexible beta, the operator's multigrid V-cycle applied to the whole
// block in FP32.  Column reductions are deterministic: fixed grid, per-block partials, the last
// block to arrive sums them in a fixed order.
#include "sell.h"

namespace mk {

static constexpr int kBlkParts = 2048;  // grid of the column-reduction kernels

// 16-byte vectors of T (4 floats or 2 doubles) carried by one lane
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    static constexpr int N = 4;
    using type = float4;
};
template <>
struct Vec16<double> {
    static constexpr int N = 2;
    using type = double2;
};
template <typename V>
__device__ __forceinline__ void vfma(V& acc, float a, const V& x) {
    acc.x += a * x.x; acc.y += a * x.y; acc.z += a * x.z; acc.w += a * x.w;
}
__device__ __forceinline__ void vfma(double2& acc, double a, const double2& x) {
    acc.x += a * x.x; acc.y += a * x.y;
}
__device__ __forceinline__ void vzero(float4& a) { a.x = a.y = a.z = a.w = 0.f; }
__device__ __forceinline__ void vzero(double2& a) { a.x = a.y = 0.0; }
__device__ __forceinline__ void vshfl_add(float4& a, int off) {
    a.x += __shfl_xor(a.x, off, 64); a.y += __shfl_xor(a.y, off, 64);
    a.z += __shfl_xor(a.z, off, 64); a.w += __shfl_xor(a.w, off, 64);
}
__device__ __forceinline__ void vshfl_add(double2& a, int off) {
    a.x += __shfl_xor(a.x, off, 64); a.y += __shfl_xor(a.y, off, 64);
}
__device__ __forceinline__ float4 vsub(const float4& b, const float4& a) {
    return make_float4(b.x - a.x, b.y - a.y, b.z - a.z, b.w - a.w);
}
__device__ __forceinline__ double2 vsub(const double2& b, const double2& a) { return make_double2(b.x - a.x, b.y - a.y); }
__device__ __forceinline__ float4 vadd(const float4& b, const float4& a) {
    return make_float4(b.x + a.x, b.y + a.y, b.z + a.z, b.w + a.w);
}
__device__ __forceinline__ double2 vadd(const double2& b, const double2& a) { return make_double2(b.x + a.x, b.y + a.y); }

__device__ __forceinline__ float4 vaxpy(float a, const float4& x, float b, const float4& y) {  // a x + b y
    return make_float4(a * x.x + b * y.x, a * x.y + b * y.y, a * x.z + b * y.z, a * x.w + b * y.w);
}
__device__ __forceinline__ double2 vaxpy(double a, const double2& x, double b, const double2& y) {
    return make_double2(a * x.x + b * y.x, a * x.y + b * y.y);
}
__device__ __forceinline__ float4 vscale(float a, const float4& x) {
    return make_float4(a * x.x, a * x.y, a * x.z, a * x.w);
}
__device__ __forceinline__ double2 vscale(double a, const double2& x) { return make_double2(a * x.x, a * x.y); }

// Y = A X (MODE 0), Y = B - A X (3), Y += A X (4) for MW right-hand sides.  MODES 5-7: one
// Chebyshev smoothing step fused with the residual it needs, in the three-term form (the
// step d_{k-1} of the two-term form is x_k - x_{k-1}), row-local, X = x_k (gathered) and
// Y != X:  r = B - A X;  5: Y = X + cs D^-1 r (first step from a given x);
// 6: Y = X + ca (X - Y) + cs D^-1 r (Y holds x_{k-1} on entry and x_{k+1} on exit);
// 7: Y = (1 + ca) X + cs D^-1 r (the step after x_1 from a zero start, x_0 = 0).
// 8: mode 7 with x_1 = c0 D^-1 B formed inside the gathers (X unused, x_1 never stored).
// 9: mode 6 with x_{k-1} = x_1 = c0 D^-1 B formed from the row's own B (Y is not read).
// A lane carries one 16-byte vector of columns (LPR = MW / N lanes per row), a wavefront
// covers 64 / LPR rows of the SELL part at a time; medium and long rows take a whole
// wavefront whose row groups split the entries.  PART 1 / 2: only the SELL rows / only the
// medium and long rows (launched with nw_sell = 0), so the SELL launch does not carry the
// registers of the eight-gather chains and runs at full occupancy.
// Stores row `row` of the result (lane j of its LPR lanes): Y = acc (MODE 0), B - acc (3),
// Y + acc (4), the Chebyshev updates (5-7, x_k = xd when the row's own entry was gathered).
template <typename T, int MW, int MODE, typename V>
__device__ __forceinline__ void spmm_store(const V& acc, int row, int j, const V* __restrict__ Xv, T* __restrict__ Y,
                                           const T* __restrict__ B, const T* __restrict__ dinv, T ca, T cs, T c0,
                                           bool have_xd, const V& xd) {
    constexpr int LPR = MW / Vec16<T>::N;
    V* Yv = reinterpret_cast<V*>(Y);
    const size_t o = size_t(row) * LPR + j;
    if constexpr (MODE == 0) Yv[o] = acc;
    if constexpr (MODE == 3) Yv[o] = vsub(reinterpret_cast<const V*>(B)[o], acc);
    if constexpr (MODE == 4) Yv[o] = vadd(Yv[o], acc);
    if constexpr (MODE >= 5) {
        const V bo = reinterpret_cast<const V*>(B)[o];
        const V xo = have_xd ? xd : MODE == 8 ? vscale(c0 * dinv[row], bo) : Xv[o];
        const V r = vsub(bo, acc);
        const T sd = cs * dinv[row];
        if constexpr (MODE == 5) Yv[o] = vaxpy(T(1), xo, sd, r);
        if constexpr (MODE == 7 || MODE == 8) Yv[o] = vaxpy(T(1) + ca, xo, sd, r);
        if constexpr (MODE == 6) Yv[o] = vaxpy(T(1), vaxpy(T(1) + ca, xo, -ca, Yv[o]), sd, r);
        if constexpr (MODE == 9)
            Yv[o] = vaxpy(T(1), vaxpy(T(1) + ca, xo, -ca, vscale(c0 * dinv[row], bo)), sd, r);
    }
}

// Gathered row c of the iterate: X's (lane vector j), or in mode 8 x_1 = c0 D^-1 B formed on
// the fly (the same product the separate first step would have stored).
template <typename T, int MODE, int LPR, typename V>
__device__ __forceinline__ V gather_row(const V* __restrict__ Gv, const T* __restrict__ dinv, T c0, int c, int j) {
    const V g = Gv[size_t(c) * LPR + j];
    if constexpr (MODE == 8) return vscale(c0 * dinv[c], g);
    else return g;
}

// Sum over entries [k0, k1) of a CSR row (stride R between a lane group's entries, starting
// at k0 + sub): eight independent gathers in flight per lane.
template <typename T, int R, int LPR, int MODE, typename V>
__device__ __forceinline__ V chain_sum(int64_t k0, int64_t k1, int sub, int j, const int* __restrict__ cols,
                                       const T* __restrict__ vals, const V* __restrict__ Xv,
                                       const T* __restrict__ dinv, T s0) {
    V acc, a1, a2, a3, a4, a5, a6, a7;
    vzero(acc); vzero(a1); vzero(a2); vzero(a3); vzero(a4); vzero(a5); vzero(a6); vzero(a7);
    int64_t k = k0 + sub;
    for (; k + 7 * R < k1; k += 8 * R) {
        const int c0 = cols[k], c1 = cols[k + R], c2 = cols[k + 2 * R], c3 = cols[k + 3 * R];
        const int c4 = cols[k + 4 * R], c5 = cols[k + 5 * R], c6 = cols[k + 6 * R], c7 = cols[k + 7 * R];
        const V x0 = gather_row<T, MODE, LPR>(Xv, dinv, s0, c0, j), x1 = gather_row<T, MODE, LPR>(Xv, dinv, s0, c1, j);
        const V x2 = gather_row<T, MODE, LPR>(Xv, dinv, s0, c2, j), x3 = gather_row<T, MODE, LPR>(Xv, dinv, s0, c3, j);
        const V x4 = gather_row<T, MODE, LPR>(Xv, dinv, s0, c4, j), x5 = gather_row<T, MODE, LPR>(Xv, dinv, s0, c5, j);
        const V x6 = gather_row<T, MODE, LPR>(Xv, dinv, s0, c6, j), x7 = gather_row<T, MODE, LPR>(Xv, dinv, s0, c7, j);
        vfma(acc, vals[k], x0);
        vfma(a1, vals[k + R], x1);
        vfma(a2, vals[k + 2 * R], x2);
        vfma(a3, vals[k + 3 * R], x3);
        vfma(a4, vals[k + 4 * R], x4);
        vfma(a5, vals[k + 5 * R], x5);
        vfma(a6, vals[k + 6 * R], x6);
        vfma(a7, vals[k + 7 * R], x7);
    }
    for (; k < k1; k += R) vfma(acc, vals[k], gather_row<T, MODE, LPR>(Xv, dinv, s0, cols[k], j));
    return vadd(vadd(vadd(acc, a1), vadd(a2, a3)), vadd(vadd(a4, a5), vadd(a6, a7)));
}

template <typename T, int MW, int MODE, int PART = 0>
__global__ void __launch_bounds__(256) sell_spmm(SellDev<T> A, int nw_sell, const T* __restrict__ X,
                                               T* __restrict__ Y, const T* __restrict__ B, const int* skip,
                                               const T* __restrict__ dinv = nullptr, T ca = T(0), T cs = T(0),
                                               T c0 = T(0)) {
    if (skip && *skip) return;
    using V = typename Vec16<T>::type;
    constexpr int N = Vec16<T>::N;
    constexpr int LPR = MW / N;   // lanes per row
    constexpr int R = 64 / LPR;   // rows per wavefront
    // the gathered block: X, or B in mode 8 (scaled by c0 D^-1 as it is gathered)
    const V* __restrict__ Xv = reinterpret_cast<const V*>(MODE == 8 ? B : X);
    const int lane = threadIdx.x & 63;
    const int sub = lane / LPR, j = lane - sub * LPR;
    const int gw = xcd_block(int(blockIdx.x), int(gridDim.x)) * int(blockDim.x >> 6) + int(threadIdx.x >> 6);
    V acc;
    vzero(acc);
    // modes 5-7 need the row's own x_k: taken from the gather of the diagonal entry, so the
    // SELL rows read X once (medium and long rows, a few hundred, read it again at the end)
    V xd;
    vzero(xd);
    bool have_xd = false;
    int row = -1;
    if (PART != 2 && gw < nw_sell) {
        const int r = gw * R + sub;
        if (r >= A.n_short) return;
        const int s = r / A.C, lr = r - s * A.C;
        const int64_t base = A.slice_ptr[s] + lr;
        const int w = A.slice_w[s];
        row = A.perm ? A.perm[r] : r;
        // the row's column indices and values first (w <= kSellMaxLen, uniform in the
        // wavefront), then its gathers kG at a time in flight, summed in entry order
        int ci[kSellMaxLen];
        T cv[kSellMaxLen];
#pragma unroll
        for (int k = 0; k < kSellMaxLen; k++)
            if (k < w) {
                ci[k] = A.col[base + int64_t(k) * A.C];
                cv[k] = A.val[base + int64_t(k) * A.C];
            }
        constexpr int kG = 8;
#pragma unroll
        for (int k0 = 0; k0 < kSellMaxLen; k0 += kG) {
            if (k0 >= w) break;
            V xg[kG];
#pragma unroll
            for (int u = 0; u < kG; u++)
                if (k0 + u < kSellMaxLen && k0 + u < w) xg[u] = gather_row<T, MODE, LPR>(Xv, dinv, c0, ci[k0 + u], j);
#pragma unroll
            for (int u = 0; u < kG; u++)
                if (k0 + u < kSellMaxLen && k0 + u < w) {
                    vfma(acc, cv[k0 + u], xg[u]);
                    if constexpr (MODE >= 5) {
                        if (ci[k0 + u] == row) { xd = xg[u]; have_xd = true; }
                    }
                }
        }
    } else {
        if constexpr (PART == 1) return;
        int lw = gw - nw_sell;
        if (lw < A.nmid) {  // a medium row: one wavefront, its R lane groups split the entries
            acc = chain_sum<T, R, LPR, MODE>(A.mid_ptr[lw], A.mid_ptr[lw + 1], sub, j, A.mid_col, A.mid_val, Xv, dinv,
                                             c0);
#pragma unroll
            for (int off = LPR; off < 64; off <<= 1) vshfl_add(acc, off);
            if (sub != 0) return;
            row = A.mid_rows[lw];
        } else {  // a segment of a long row: its partial sum, combined by sell_spmm_long
            lw -= A.nmid;
            if (lw >= A.nseg) return;  // uniform inside the wavefront
            acc = chain_sum<T, R, LPR, MODE>(A.seg_beg[lw], A.seg_end[lw], sub, j, A.long_col, A.long_val, Xv, dinv,
                                             c0);
#pragma unroll
            for (int off = LPR; off < 64; off <<= 1) vshfl_add(acc, off);
            if (sub == 0) reinterpret_cast<V*>(A.seg_part)[size_t(lw) * LPR + j] = acc;
            return;
        }
    }
    spmm_store<T, MW, MODE>(acc, row, j, Xv, Y, B, dinv, ca, cs, c0, have_xd, xd);
}

// Long rows of the multi-column kernels: one wavefront per row adds its segments' partial
// sums in segment order (lanes j < LPR) and stores the row as sell_spmm does.
template <typename T, int MW, int MODE>
__global__ void __launch_bounds__(256) sell_spmm_long(SellDev<T> A, const T* __restrict__ X, T* __restrict__ Y,
                                                    const T* __restrict__ B, const int* skip,
                                                    const T* __restrict__ dinv, T ca, T cs, T c0 = T(0)) {
    if (skip && *skip) return;
    using V = typename Vec16<T>::type;
    constexpr int LPR = MW / Vec16<T>::N;
    const int w = int(blockIdx.x) * int(blockDim.x >> 6) + int(threadIdx.x >> 6);
    const int j = int(threadIdx.x & 63);
    if (w >= A.nlong || j >= LPR) return;
    const V* __restrict__ P = reinterpret_cast<const V*>(A.seg_part);
    V acc;
    vzero(acc);
    for (int q = A.lseg_ptr[w]; q < A.lseg_ptr[w + 1]; q++) acc = vadd(acc, P[size_t(q) * LPR + j]);
    V xd;
    vzero(xd);
    spmm_store<T, MW, MODE>(acc, A.long_rows[w], j, reinterpret_cast<const V*>(X), Y, B, dinv, ca, cs, c0, false, xd);
}

template <typename T, int MW>
static void spmm_mw(mk_ctx* c, const mk_op* A, const T* val, const T* X, T* Y, const T* B, SpmvMode mode,
                    const int* skip, const SpmvTimer& tm) {
    unsigned g0 = 0;
    const SellDev<T> D = sell_dev<T>(A, val, 256, &g0);
    constexpr int R = 64 / (MW / Vec16<T>::N);
    const int nw_sell = int((int64_t(A->n_short) + R - 1) / R);
    const int64_t wl = int64_t(A->nmid) + A->nseg;
    const unsigned g1 = unsigned(std::max<int64_t>(1, (int64_t(nw_sell) + 3) / 4));
    const unsigned g2 = unsigned((wl + 3) / 4);
    const unsigned g3 = unsigned((int64_t(A->nlong) + 3) / 4);
    // SELL rows, then (when there are any) the medium rows and long-row segments, then the
    // long rows' sums
#define MK_SPMM_PARTS(M)                                                                                        \
    MK_TLAUNCH(tm, 0, (sell_spmm<T, MW, M, 1>), dim3(g1), dim3(256), 0, c->stream, D, nw_sell, X, Y, B, skip, \
               (const T*)nullptr, T(0), T(0), T(0));                                                                    \
    if (g2)                                                                                                             \
        MK_TLAUNCH(tm, 1, (sell_spmm<T, MW, M, 2>), dim3(g2), dim3(256), 0, c->stream, D, 0, X, Y, B, skip,          \
                   (const T*)nullptr, T(0), T(0), T(0));                                                                \
    if (g3)                                                                                                             \
        MK_TLAUNCH(tm, 2, (sell_spmm_long<T, MW, M>), dim3(g3), dim3(256), 0, c->stream, D, X, Y, B, skip,            \
                   (const T*)nullptr, T(0), T(0), T(0));
    switch (mode) {
        case SPMV_PLAIN: MK_SPMM_PARTS(0) break;
        case SPMV_RESID_NORED: MK_SPMM_PARTS(3) break;
        case SPMV_ACCUM: MK_SPMM_PARTS(4) break;
        default: throw Error(MK_ERR_ARG, "spmm: unsupported mode");
    }
#undef MK_SPMM_PARTS
}

template <typename T, int MW>
static void spmm_cheb_mw(mk_ctx* c, const mk_op* A, const T* val, const T* X, T* Y, const T* B, const T* dinv, T ca,
                         T cs, T c0, int mode, const int* skip, const SpmvTimer& tm) {
    unsigned g0 = 0;
    const SellDev<T> Dd = sell_dev<T>(A, val, 256, &g0);
    constexpr int R = 64 / (MW / Vec16<T>::N);
    const int nw_sell = int((int64_t(A->n_short) + R - 1) / R);
    const int64_t wl = int64_t(A->nmid) + A->nseg;
    const unsigned g1 = unsigned(std::max<int64_t>(1, (int64_t(nw_sell) + 3) / 4));
    const unsigned g2 = unsigned((wl + 3) / 4);
    const unsigned g3 = unsigned((int64_t(A->nlong) + 3) / 4);
#define MK_CHEB_PARTS(M)                                                                                            \
    MK_TLAUNCH(tm, 0, (sell_spmm<T, MW, M, 1>), dim3(g1), dim3(256), 0, c->stream, Dd, nw_sell, X, Y, B,   \
               skip, dinv, ca, cs, c0);                                                                             \
    if (g2)                                                                                                         \
        MK_TLAUNCH(tm, 1, (sell_spmm<T, MW, M, 2>), dim3(g2), dim3(256), 0, c->stream, Dd, 0, X, Y, B, skip,     \
                   dinv, ca, cs, c0);                                                                               \
    if (g3)                                                                                                         \
        MK_TLAUNCH(tm, 2, (sell_spmm_long<T, MW, M>), dim3(g3), dim3(256), 0, c->stream, Dd, X, Y, B, skip, dinv, \
                   ca, cs, c0);
    switch (mode) {
        case 5: MK_CHEB_PARTS(5) break;
        case 6: MK_CHEB_PARTS(6) break;
        case 7: MK_CHEB_PARTS(7) break;
        case 8: MK_CHEB_PARTS(8) break;
        case 9: MK_CHEB_PARTS(9) break;
        default: throw Error(MK_ERR_ARG, "block Chebyshev: mode 5 to 9");
    }
#undef MK_CHEB_PARTS
}

// One fused three-term Chebyshev step on a block of mw right-hand sides (modes 5-9 of
// sell_spmm; x_out != x_in, x_out holds x_{k-1} on entry in mode 6; modes 8 and 9 take
// x_1 = c0 D^-1 b, mode 8 in place of x_in, which it does not read).
template <typename T>
void spmm_cheb_launch(mk_ctx* c, const mk_op* A, int mw, const T* x_in, T* x_out, const T* b, T ca, T cs, int mode,
                      const int* skip, T c0) {
    MK_CHECK(!A->cplx && x_in != x_out, MK_ERR_ARG, "fused block Chebyshev: real operator, x_out != x_in");
    MK_CHECK(mode >= 5 && mode <= 9, MK_ERR_ARG, "block Chebyshev: mode 5 to 9");
    // algorithmic bytes: the matrix (values, columns, slice / row metadata), the gathered block
    // once (A->m rows: every column is gathered, the row's own entry among them; B itself in
    // mode 8), B read, x_out written (and read first in mode 6), the inverse diagonal
    const double s = double(sizeof(T)), vb = s * mw;
    const double bytes = (sizeof(T) == 8 ? A->bytes64 : A->bytes32) - double(A->n) * 2.0 * s +
                         (mode == 8 ? 0.0 : double(A->m) * vb) + double(A->n) * vb * (mode == 6 ? 3.0 : 2.0) +
                         double(A->n) * s;
    SpmvTimer tm(c, bytes, A, mw == 64 ? "spmm64_cheb" : mw == 32 ? "spmm32_cheb" : mw == 16 ? "spmm16_cheb" :
                 "spmm8_cheb", sizeof(T) == 4, 3);
    const T* val;
    const T* dinv;
    if constexpr (std::is_same<T, double>::value) { val = A->val64.p; dinv = A->dinv64.p; }
    else { val = A->val32.p; dinv = A->dinv32.p; }
    switch (mw) {
        case 8: spmm_cheb_mw<T, 8>(c, A, val, x_in, x_out, b, dinv, ca, cs, c0, mode, skip, tm); break;
        case 16: spmm_cheb_mw<T, 16>(c, A, val, x_in, x_out, b, dinv, ca, cs, c0, mode, skip, tm); break;
        case 32: spmm_cheb_mw<T, 32>(c, A, val, x_in, x_out, b, dinv, ca, cs, c0, mode, skip, tm); break;
        case 64: spmm_cheb_mw<T, 64>(c, A, val, x_in, x_out, b, dinv, ca, cs, c0, mode, skip, tm); break;
        default: throw Error(MK_ERR_ARG, "spmm: block width must be 8, 16, 32 or 64");
    }
    MK_HIP(hipGetLastError());
}
template void spmm_cheb_launch<float>(mk_ctx*, const mk_op*, int, const float*, float*, const float*, float, float,
                                      int, const int*, float);
template void spmm_cheb_launch<double>(mk_ctx*, const mk_op*, int, const double*, double*, const double*, double,
                                       double, int, const int*, double);

template <typename T>
void spmm_launch(mk_ctx* c, const mk_op* A, int mw, const T* X, T* Y, const T* B, SpmvMode mode, const int* skip) {
    MK_CHECK(!A->cplx, MK_ERR_UNSUPPORTED, "spmm: real operators only");
    const T* val;
    if constexpr (std::is_same<T, double>::value) val = A->val64.p;
    else val = A->val32.p;
    const double vb = double(sizeof(T)) * mw;
    const double bytes = (sizeof(T) == 8 ? A->bytes64 : A->bytes32) - double(A->n) * 2.0 * sizeof(T) +
                         double(A->m) * vb + double(A->n) * vb * (mode == SPMV_PLAIN ? 1.0 : 2.0);
    SpmvTimer tm(c, bytes, A, mw == 64 ? "spmm64" : mw == 32 ? "spmm32" : mw == 16 ? "spmm16" : "spmm8",
                 sizeof(T) == 4, 3);
    switch (mw) {
        case 8: spmm_mw<T, 8>(c, A, val, X, Y, B, mode, skip, tm); break;
        case 16: spmm_mw<T, 16>(c, A, val, X, Y, B, mode, skip, tm); break;
        case 32: spmm_mw<T, 32>(c, A, val, X, Y, B, mode, skip, tm); break;
        case 64: spmm_mw<T, 64>(c, A, val, X, Y, B, mode, skip, tm); break;
        default: throw Error(MK_ERR_ARG, "spmm: block width must be 8, 16, 32 or 64");
    }
    MK_HIP(hipGetLastError());
}
template void spmm_launch<double>(mk_ctx*, const mk_op*, int, const double*, double*, const double*, SpmvMode,
                                  const int*);
template void spmm_launch<float>(mk_ctx*, const mk_op*, int, const float*, float*, const float*, SpmvMode,
                                 const int*);

// ---- batched PCG --------------------------------------------------------------------------

// Per-column partial sums of K quantities of this block (the thread's column is
// threadIdx.x % MW, blockDim 256, grid <= kBlkParts), stored at
// partials[(k * 64 + j) * kBlkParts + block].  No fence or ticket: a one-block finishing kernel
// adds them (colsum_total) after the kernel boundary, in a fixed order.
template <int K, int MW>
__device__ __forceinline__ void colsum_write(double (&v)[K], double* partials) {
    __shared__ double sh[K][4][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double a = v[k];
#pragma unroll
        for (int off = MW; off < 64; off <<= 1) a += __shfl_xor(a, off, 64);
        if (lane < MW) sh[k][wid][lane] = a;
    }
    __syncthreads();
    if (threadIdx.x < MW) {
        const int t = threadIdx.x;
#pragma unroll
        for (int k = 0; k < K; k++)
            partials[(size_t(k) * 64 + t) * kBlkParts + blockIdx.x] = sh[k][0][t] + sh[k][1][t] + sh[k][2][t] + sh[k][3][t];
    }
}
// Column totals over nparts blocks (one block of 256 threads): tot[k][j], j < MW, valid in every
// thread after the call.
template <int K, int MW>
__device__ __forceinline__ void colsum_total(const double* __restrict__ partials, int nparts, double (*tot)[64]) {
    __shared__ double sh[K][4][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, j = threadIdx.x % MW;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double a = 0.0;
        const double* pk = partials + (size_t(k) * 64 + j) * kBlkParts;
        for (int b = threadIdx.x / MW; b < nparts; b += 256 / MW) a += pk[b];
#pragma unroll
        for (int off = MW; off < 64; off <<= 1) a += __shfl_xor(a, off, 64);
        if (lane < MW) sh[k][wid][lane] = a;
    }
    __syncthreads();
    if (threadIdx.x < MW) {
        const int t = threadIdx.x;
#pragma unroll
        for (int k = 0; k < K; k++) tot[k][t] = sh[k][0][t] + sh[k][1][t] + sh[k][2][t] + sh[k][3][t];
    }
    __syncthreads();
}

__device__ __forceinline__ void bstate_all_done(BState* st, int mw) {
    int all = 1;
    for (int j = 0; j < mw; j++) all &= st->done[j];
    st->all_done = all;
}

// X = 0 ; R = B ; Rz = ZT(B) (the V-cycle's input copy; none for an FP64 V-cycle) ; bb_j = |B_j|^2
template <int MW, typename ZT>
__global__ void __launch_bounds__(256) bk_init0(size_t len, const double* __restrict__ B, double* __restrict__ X,
                                              double* __restrict__ R, ZT* __restrict__ Rz, double* partials) {
    double v[1] = {0.0};
    for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < len; e += size_t(gridDim.x) * 256) {
        const double b = B[e];
        X[e] = 0.0;
        R[e] = b;
        if (Rz) Rz[e] = ZT(b);
        v[0] += b * b;
    }
    colsum_write<1, MW>(v, partials);
}
template <int MW>
__global__ void __launch_bounds__(256) bk_init0_fin(BState* st, const double* __restrict__ partials, int nparts) {
    __shared__ double tot[1][64];
    colsum_total<1, MW>(partials, nparts, tot);
    if (threadIdx.x < MW) {
        const int j = threadIdx.x;
        st->bb[j] = tot[0][j];
        st->it[j] = 0;
        st->status[j] = MK_OK;
    }
}

// P = Z ; rz_j = R.Z ; rr_j = R.R ; done flags
template <int MW, typename ZT>
__global__ void __launch_bounds__(256) bk_init1(size_t len, const double* __restrict__ R, const ZT* __restrict__ Z,
                                              double* __restrict__ P, double* partials) {
    double v[2] = {0.0, 0.0};
    for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < len; e += size_t(gridDim.x) * 256) {
        const double z = double(Z[e]), r = R[e];
        P[e] = z;
        v[0] += r * z;
        v[1] += r * r;
    }
    colsum_write<2, MW>(v, partials);
}
template <int MW>
__global__ void __launch_bounds__(256) bk_init1_fin(BState* st, const double* __restrict__ partials, int nparts) {
    __shared__ double tot[2][64];
    colsum_total<2, MW>(partials, nparts, tot);
    {
        if (threadIdx.x < MW) {
            const int j = threadIdx.x;
            const double rz = tot[0][j], rr = tot[1][j];
            st->rz[j] = rz;
            st->rr[j] = rr;
            int d = (rr <= st->tol2 * st->bb[j]) ? 1 : 0;
            if (!(rz == rz) || !(rr == rr)) { st->status[j] = MK_ERR_NAN; d = 1; }
            else if (!d && !(rz > 0.0)) { st->status[j] = MK_ERR_NOT_SPD; d = 1; }
            st->done[j] = d;
        }
        __syncthreads();
        if (threadIdx.x == 0) bstate_all_done(st, MW);
    }
}

// pq_j = P.Q ; alpha_j = rz_j / pq_j (0 for finished columns)
template <int MW>
__global__ void __launch_bounds__(256) bk_alpha(size_t len, const double* __restrict__ P, const double* __restrict__ Q,
                                              const BState* st, double* partials) {
    if (st->all_done) return;
    double v[1] = {0.0};
    for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < len; e += size_t(gridDim.x) * 256) v[0] += P[e] * Q[e];
    colsum_write<1, MW>(v, partials);
}
template <int MW>
__global__ void __launch_bounds__(256) bk_alpha_fin(BState* st, const double* __restrict__ partials, int nparts) {
    if (st->all_done) return;
    __shared__ double tot[1][64];
    colsum_total<1, MW>(partials, nparts, tot);
    if (threadIdx.x < MW) {
        const int j = threadIdx.x;
        const double pq = tot[0][j];
        if (st->done[j]) { st->alpha[j] = 0.0; }
        else if (!(pq > 0.0)) {
            st->status[j] = (pq == pq) ? MK_ERR_NOT_SPD : MK_ERR_NAN;
            st->done[j] = 1;
            st->alpha[j] = 0.0;
        } else {
            st->alpha[j] = st->rz[j] / pq;
        }
    }
}

// X += alpha P ; R -= alpha Q ; Rz = ZT(R) (none for an FP64 V-cycle) ; rr_j ; convergence
template <int MW, typename ZT>
__global__ void __launch_bounds__(256) bk_update(size_t len, double* __restrict__ X, double* __restrict__ R,
                                               ZT* __restrict__ Rz, const double* __restrict__ P,
                                               const double* __restrict__ Q, const BState* st, double* partials) {
    if (st->all_done) return;
    const double alpha = st->alpha[threadIdx.x % MW];
    double v[1] = {0.0};
    for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < len; e += size_t(gridDim.x) * 256) {
        double r = R[e];
        if (alpha != 0.0) {
            X[e] += alpha * P[e];
            r -= alpha * Q[e];
            R[e] = r;
            if (Rz) Rz[e] = ZT(r);
        }
        v[0] += r * r;
    }
    colsum_write<1, MW>(v, partials);
}
template <int MW>
__global__ void __launch_bounds__(256) bk_update_fin(BState* st, const double* __restrict__ partials, int nparts) {
    if (st->all_done) return;
    __shared__ double tot[1][64];
    colsum_total<1, MW>(partials, nparts, tot);
    {
        if (threadIdx.x < MW) {
            const int j = threadIdx.x;
            const double rr = tot[0][j];
            if (!st->done[j]) {
                st->rr[j] = rr;
                st->it[j] += 1;
                if (!(rr == rr)) { st->status[j] = MK_ERR_NAN; st->done[j] = 1; }
                else if (rr <= st->tol2 * st->bb[j]) st->done[j] = 1;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) bstate_all_done(st, MW);
    }
}

// flexible beta_j = (R.Z - R.Zold) / rz_j ; rz_j = R.Z
template <int MW, typename ZT>
__global__ void __launch_bounds__(256) bk_beta(size_t len, const double* __restrict__ R, const ZT* __restrict__ Z,
                                             const ZT* __restrict__ Zold, const BState* st, double* partials) {
    if (st->all_done) return;
    double v[2] = {0.0, 0.0};
    for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < len; e += size_t(gridDim.x) * 256) {
        const double r = R[e];
        v[0] += r * double(Z[e]);
        v[1] += r * double(Zold[e]);
    }
    colsum_write<2, MW>(v, partials);
}
template <int MW>
__global__ void __launch_bounds__(256) bk_beta_fin(BState* st, const double* __restrict__ partials, int nparts) {
    if (st->all_done) return;
    __shared__ double tot[2][64];
    colsum_total<2, MW>(partials, nparts, tot);
    {
        if (threadIdx.x < MW) {
            const int j = threadIdx.x;
            if (!st->done[j]) {
                const double a = tot[0][j];
                if (!(a > 0.0)) {
                    st->status[j] = (a == a) ? MK_ERR_NOT_SPD : MK_ERR_NAN;
                    st->done[j] = 1;
                } else {
                    st->beta[j] = (a - tot[1][j]) / st->rz[j];
                    st->rz[j] = a;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) bstate_all_done(st, MW);
    }
}

// P = Z + beta P (running columns only)
template <int MW, typename ZT>
__global__ void __launch_bounds__(256) bk_pupdate(size_t len, double* __restrict__ P, const ZT* __restrict__ Z,
                                                const BState* st) {
    if (st->all_done) return;
    const int j = threadIdx.x % MW;
    if (st->done[j]) return;
    const double beta = st->beta[j];
    for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < len; e += size_t(gridDim.x) * 256)
        P[e] = double(Z[e]) + beta * P[e];
}

__global__ void k_set_bstate(BState* st, double tol2) {
    st->tol2 = tol2;
    st->all_done = 0;
}

// ZT: the V-cycle's precision (float: FP32 multigrid inside the FP64 recurrences; double: an
// FP64 V-cycle on R itself, the plain FP64 form of the same algorithm)
template <int MW, typename ZT>
static BlockResult block_pcg_mw(mk_ctx* c, const mk_op* A, Mg* mg, const double* B, double* X, double rel_tol,
                                int maxiter, int check_every) {
    constexpr bool fp64 = std::is_same<ZT, double>::value;
    const size_t len = size_t(A->n) * MW;
    DevBuf<double> R(len), P(len), Q(len), part(size_t(2) * 64 * kBlkParts);
    // two preconditioned residual blocks, swapped every iteration: the flexible beta needs the
    // previous one (kept in the V-cycle's own precision, so no copy is made)
    DevBuf<ZT> Rz(fp64 ? 0 : len), Za(len), Zb(len);
    ZT* zc = Za.p;  // the V-cycle's current output
    ZT* zo = Zb.p;  // the previous one
    ZT* rz = fp64 ? nullptr : Rz.p;                 // the V-cycle's input copy of R
    const ZT* vin = fp64 ? reinterpret_cast<const ZT*>(R.p) : Rz.p;
    DevBuf<BState> st(1);
    MK_HIP(hipMemsetAsync(st.p, 0, sizeof(BState), c->stream));
    const unsigned g = unsigned(std::min<size_t>(std::min<size_t>(size_t(fin_grid()), size_t(kBlkParts)),
                                                 std::max<size_t>(1, (len + 255) / 256)));
    const int ng = int(g);
    // the recurrence residual is driven to half the target so the true residual meets it
    hipLaunchKernelGGL(k_set_bstate, dim3(1), dim3(1), 0, c->stream, st.p, 0.25 * rel_tol * rel_tol);
    hipLaunchKernelGGL((bk_init0<MW, ZT>), dim3(g), dim3(256), 0, c->stream, len, B, X, R.p, rz, part.p);
    hipLaunchKernelGGL((bk_init0_fin<MW>), dim3(1), dim3(256), 0, c->stream, st.p, (const double*)part.p, ng);
    mg_apply_block<ZT>(c, mg, MW, vin, zc, nullptr);
    hipLaunchKernelGGL((bk_init1<MW, ZT>), dim3(g), dim3(256), 0, c->stream, len, (const double*)R.p,
                       (const ZT*)zc, P.p, part.p);
    std::swap(zc, zo);
    hipLaunchKernelGGL((bk_init1_fin<MW>), dim3(1), dim3(256), 0, c->stream, st.p, (const double*)part.p, ng);
    MK_HIP(hipGetLastError());
    BState h{};
    int iters = 0;
    // the batch ends where the slowest running column is predicted to converge
    const int interval = c->time_spmv ? 1 : std::max(1, check_every);
    int prev_iters = -1;
    double prev_rr[MW] = {};
    for (;;) {
        MK_HIP(hipMemcpyAsync(&h, st.p, sizeof(BState), hipMemcpyDeviceToHost, c->stream));
        MK_HIP(hipStreamSynchronize(c->stream));
        timing_flush(c);
        if (h.all_done || iters >= maxiter) break;
        int plan = 0;
        for (int j = 0; j < MW && plan < interval; j++) {
            if (h.done[j]) continue;
            const double nd = BatchPlanner::need(iters, h.rr[j], prev_iters, prev_rr[j], h.tol2 * h.bb[j]);
            plan = std::max(plan, BatchPlanner::clamp_need(nd, interval));
        }
        if (plan == 0) plan = interval;
        for (int j = 0; j < MW; j++) prev_rr[j] = h.rr[j];
        prev_iters = iters;
        const int batch = std::min(plan, maxiter - iters);
        for (int k = 0; k < batch; k++) {
            spmm_launch<double>(c, A, MW, P.p, Q.p, nullptr, SPMV_PLAIN, &st.p->all_done);
            hipLaunchKernelGGL((bk_alpha<MW>), dim3(g), dim3(256), 0, c->stream, len, (const double*)P.p,
                               (const double*)Q.p, (const BState*)st.p, part.p);
            hipLaunchKernelGGL((bk_alpha_fin<MW>), dim3(1), dim3(256), 0, c->stream, st.p, (const double*)part.p, ng);
            hipLaunchKernelGGL((bk_update<MW, ZT>), dim3(g), dim3(256), 0, c->stream, len, X, R.p, rz,
                               (const double*)P.p, (const double*)Q.p, (const BState*)st.p, part.p);
            hipLaunchKernelGGL((bk_update_fin<MW>), dim3(1), dim3(256), 0, c->stream, st.p, (const double*)part.p, ng);
            mg_apply_block<ZT>(c, mg, MW, vin, zc, &st.p->all_done);
            hipLaunchKernelGGL((bk_beta<MW, ZT>), dim3(g), dim3(256), 0, c->stream, len, (const double*)R.p,
                               (const ZT*)zc, (const ZT*)zo, (const BState*)st.p, part.p);
            hipLaunchKernelGGL((bk_beta_fin<MW>), dim3(1), dim3(256), 0, c->stream, st.p, (const double*)part.p, ng);
            hipLaunchKernelGGL((bk_pupdate<MW, ZT>), dim3(g), dim3(256), 0, c->stream, len, P.p, (const ZT*)zc,
                               (const BState*)st.p);
            std::swap(zc, zo);
        }
        iters += batch;
        MK_HIP(hipGetLastError());
    }
    BlockResult res;
    res.status = MK_OK;
    res.iters = 0;
    res.max_rel = 0.0;
    for (int j = 0; j < MW; j++) {
        res.iters = std::max(res.iters, h.it[j]);
        if (h.status[j] != MK_OK && res.status == MK_OK) res.status = h.status[j];
        if (h.bb[j] > 0.0) res.max_rel = std::max(res.max_rel, std::sqrt(h.rr[j] / h.bb[j]));
    }
    if (res.status == MK_OK && !h.all_done) res.status = MK_ERR_NO_CONVERGENCE;
    return res;
}

BlockResult block_pcg(mk_ctx* c, const mk_op* A, Mg* mg, int mw, const double* B, double* X, double rel_tol,
                      int maxiter, int check_every, bool fp64_vcycle) {
    MK_CHECK(mg, MK_ERR_ARG, "block PCG needs the operator's multigrid hierarchy");
    MK_CHECK(!A->cplx && A->n == A->m, MK_ERR_ARG, "block PCG: square real operator");
#define MK_BLOCK_PCG(MW)                                                                             \
    return fp64_vcycle ? block_pcg_mw<MW, double>(c, A, mg, B, X, rel_tol, maxiter, check_every) \
                       : block_pcg_mw<MW, float>(c, A, mg, B, X, rel_tol, maxiter, check_every);
    switch (mw) {
        case 8: MK_BLOCK_PCG(8)
        case 16: MK_BLOCK_PCG(16)
        case 32: MK_BLOCK_PCG(32)
        case 64: MK_BLOCK_PCG(64)
        default: throw Error(MK_ERR_ARG, "block PCG: block width must be 8, 16, 32 or 64");
    }
#undef MK_BLOCK_PCG
}

}  // namespace mk
