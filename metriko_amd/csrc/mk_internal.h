// Internal declarations of the metriko HIP library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>
#include <cstdio>
#include <cstdlib>
#include "metriko_hip.h"
#include "graph_key.h"
#include "metriko/core/common/hvec.h"

namespace mk {

// Progress lines on stderr when METRIKO_VERBOSE is set (long runs; no effect otherwise).
inline bool verbose() {
    static const bool v = std::getenv("METRIKO_VERBOSE") != nullptr;
    return v;
}
#define MK_LOG(...)                              \
    do {                                         \
        if (mk::verbose()) {                     \
            std::fprintf(stderr, __VA_ARGS__);   \
            std::fflush(stderr);                 \
        }                                        \
    } while (0)

struct Mg;

struct Error : std::runtime_error {
    mk_status code;
    Error(mk_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MK_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::mk::Error(MK_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define MK_CHECK(cond, code, msg) \
    do { if (!(cond)) throw ::mk::Error((code), (msg)); } while (0)

// Owning device buffer.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; return *this; }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) {
            hipError_t e = hipMalloc(&p, count * sizeof(T));
            if (e != hipSuccess) { p = nullptr; n = 0; throw Error(MK_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e)); }
        }
    }
    void ensure(size_t count) { if (count > n) alloc(count); }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
    void upload(const T* h, size_t count, hipStream_t s) {
        ensure(count);
        if (count) MK_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* h, size_t count, hipStream_t s) const {
        if (count) MK_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// Device-side state of one CG run (scalars kept on the GPU between kernels).
struct CgState {
    double rz;      // r^H z
    double pq;      // p^H A p
    double alpha, beta;
    double rr;      // ||r||^2
    double bb;      // ||b||^2 (reference norm)
    double tol2;    // convergence threshold on rr/bb
    int done;       // 1 = converged or failed
    int status;     // MK_OK / MK_ERR_NOT_SPD / MK_ERR_NAN
    int it;         // iterations taken
    int first;      // 1 until the first beta of the fused FP32-multigrid iteration
};

// Iterations to launch before the next host convergence check.  Iterations launched past
// convergence return at once, but each still costs its few dozen launches (about a third of
// a multigrid PCG iteration), so after the first check interval the count is predicted: the
// residual's geometric decrease per iteration since the previous check, extrapolated to the
// target (rounded up), within [1, interval].  An iteration stops where it would stop with
// any batching, so results do not depend on the prediction.  The solvers use interval 1 in
// kernel-timing mode, so the per-kernel averages hold no launch past convergence.
struct BatchPlanner {
    int interval;
    int prev_it = -1;
    double prev_rr = 0.0;
    explicit BatchPlanner(int iv) : interval(iv > 0 ? iv : 1) {}
    // iterations needed from rr to reach target at the rate seen since (prev_it, prev_rr); 0 =
    // no prediction
    static double need(int it, double rr, int pit, double prr, double target) {
        if (pit < 0 || it <= pit || !(rr > target) || !(prr > 0.0) || !(target > 0.0)) return 0.0;
        const double rate = std::log(rr / prr) / double(it - pit);
        if (!(rate < -1e-3)) return 0.0;
        return std::log(target / rr) / rate;
    }
    static int clamp_need(double nd, int interval) {
        if (!(nd > 0.0) || nd >= double(interval)) return interval;
        return std::max(1, int(std::ceil(nd)));
    }
    int next(int it, double rr, double target) {
        const int b = clamp_need(need(it, rr, prev_it, prev_rr, target), interval);
        prev_it = it;
        prev_rr = rr;
        return b;
    }
};

// per-context scratch vectors
struct Work {
    DevBuf<double> r, z, p, q, t;
    DevBuf<float> r32, z32, p32, q32, x32, d32;
};

// batched (multi-right-hand-side) solves (block.hip): up to 64 columns, row-major blocks
struct BState {
    double rz[64], rr[64], bb[64], alpha[64], beta[64];
    int done[64], status[64], it[64];
    int all_done, pad;
    double tol2;
};
struct BlockResult {
    mk_status status;
    int iters;        // max over the columns
    double max_rel;   // max relative residual (recurrence) over the columns
};
// multigrid (mg.hip)
// order_out: store the finest level in the aggregates' locality order and return pos[old] = new
// there (left empty for a one-level hierarchy); the hierarchy then takes and returns vectors
// in that order
Mg* mg_build_real(mk_ctx* c, int n, const int64_t* ptr, const int* col, const double* val, const mk_solve_opts& o,
                  const std::vector<double>* nullvec = nullptr, int nnull = 0, std::vector<int>* order_out = nullptr);
Mg* mg_build_complex(mk_ctx* c, int n, const int64_t* ptr, const int* col, const double* val, const mk_solve_opts& o);
void mg_apply64(mk_ctx* c, Mg* g, const double* b, double* x, const int* skip);
void mg_apply32(mk_ctx* c, Mg* g, const float* b, float* x, const int* skip);
// V-cycle on an n x mw row-major block (T = float: FP32 values and vectors; double: FP64)
template <typename T>
void mg_apply_block(mk_ctx* c, Mg* g, int mw, const T* b, T* x, const int* skip);
int mg_levels(const Mg* g);
// process-unique id of a hierarchy (never reused after it is freed; keys captured graphs)
uint64_t mg_uid(const Mg* g);
void mg_set_labels(Mg* g, const std::string& base);
// looks up the current device's copy of the allocation-free solve's static scratch (once per
// device; mk_init calls it with the device current); a failure is kept and reported by
// mk_spd_solve_noalloc
mk_status noalloc_setup(int device, std::string* err);
// builds A's multigrid hierarchy when opts ask for MK_PRECOND_AMG (no-op otherwise)
void ensure_mg(mk_ctx* c, const mk_op* A, const mk_solve_opts& o);
mk_op* op_create_owned(mk_ctx* c, int n, metriko::HVec<int64_t>&& ptr, metriko::HVec<int>&& col,
                       metriko::HVec<double>&& val, const mk_solve_opts& o);
int64_t mg_nnz(const Mg* g);
int mg_level_rows(const Mg* g, int l);

}  // namespace mk

// locally_injectivity_check data (mk_igm_injectivity_check): corners of the cut faces and
// fullx2Nfunc (CSR, N functions per cut vertex); nF = 0: no check
struct mk_inj_check_data {
    int N = 0, nF = 0, rows = 0, cols = 0;
    std::vector<int32_t> idx;
    std::vector<int64_t> ptr;
    std::vector<int32_t> col;
    std::vector<double> val;
};

struct mk_ctx {
    // process-unique id (never reused after mk_finalize; keys captured graphs)
    const uint64_t uid = mk::next_uid();
    mk::Work work;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    // reduction scratch: partial sums [kMaxParts * kMaxK], arrival counter, CG state
    mk::DevBuf<double> partials;
    mk::DevBuf<mk::CgState> state;
    mk::DevBuf<double> scal;   // general device scalars
    // SpMV timing (events around each SpMV launch when enabled), attributed per
    // "operator label:kernel:precision" key
    bool time_spmv = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::string> ev_key;   // key of each recorded event pair
    size_t ev_used = 0;
    struct KStat { int64_t calls = 0; double ms = 0.0, bytes = 0.0, bytes_per_call = 0.0; int rows = 0; };
    std::vector<std::pair<std::string, KStat>> kstats;
    double spmv_ms = 0.0;
    int64_t spmv_calls = 0;
    double spmv_bytes = 0.0;
    bool use_graphs = true;  // replay captured PCG iteration batches
    int64_t graph_replays = 0, graph_captures = 0, eager_batches = 0;
    // the next iterative rounding keeps its last LM round's system (mk_keep_igm_state)
    bool keep_igm_state = false;
    // the naive LM pass's iter_cb (mk_set_naive_iter_callback); null = none
    mk_naive_iter_cb naive_cb = nullptr;
    void* naive_user = nullptr;
    struct mk_igm_state* igm_state = nullptr;
    mk_inj_check_data inj_check;
};

namespace mk { struct Mg; void mg_free(Mg*); }

struct mk_op {
    mk_ctx* ctx = nullptr;
    int n = 0;          // rows
    int m = 0;          // columns (== n for square operators)
    mk::Mg* mg = nullptr;  // multigrid hierarchy used as preconditioner (owned)
    // captured PCG iteration batches: key (buffers, mode, batch) -> executable graph, and the
    // SpMV counters one replay adds
    struct Graph {
        std::vector<uintptr_t> key;
        hipGraphExec_t exec = nullptr;
        int64_t spmv_calls = 0;
        double spmv_bytes = 0.0;
    };
    std::vector<Graph> graphs;
    ~mk_op() {
        for (auto& g : graphs) if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (mg) mk::mg_free(mg);
    }
    bool cplx = false;
    int64_t nnz = 0;
    // SELL-C-sigma layout (C = slice height; rows sorted by length inside windows of sigma)
    int C = 64;
    int sigma = 1;
    int nslices = 0;
    int64_t stored = 0;  // stored entries incl. padding
    mk::DevBuf<int64_t> slice_ptr;
    mk::DevBuf<int> slice_w;
    mk::DevBuf<int> col;
    mk::DevBuf<double> val64;   // complex: interleaved
    mk::DevBuf<float> val32;
    mk::DevBuf<int> perm;       // SELL row -> original row (empty if identity)
    int n_short = 0;            // rows in the SELL part
    int nmid = 0;               // rows in the group-of-lanes-per-row part (medium length)
    int mid_lanes = 8;          // lanes per medium row
    mk::DevBuf<int> mid_rows;
    mk::DevBuf<int64_t> mid_ptr;
    mk::DevBuf<int> mid_col;
    mk::DevBuf<double> mid_val64;
    mk::DevBuf<float> mid_val32;
    int nlong = 0;              // rows in the wavefront-per-row part
    mk::DevBuf<int> long_rows;
    mk::DevBuf<int64_t> long_ptr;
    mk::DevBuf<int> long_col;
    mk::DevBuf<double> long_val64;
    mk::DevBuf<float> long_val32;
    // the long rows cut into segments of at most kLongSeg entries for the multi-column kernels
    // (one wavefront per segment, partial sums combined per row in segment order)
    int nseg = 0;
    mk::DevBuf<int64_t> seg_beg, seg_end;  // entry range of each segment in long_col / long_val
    mk::DevBuf<int> lseg_ptr;              // segments of long row w: [lseg_ptr[w], lseg_ptr[w + 1])
    mk::DevBuf<double> seg_part;           // nseg x 64 partial sums (FP64 or FP32 entries)
    mk::DevBuf<double> dinv64;  // inverse diagonal (real part for complex)
    mk::DevBuf<float> dinv32;
    int block = 256;
    double bytes64 = 0, bytes32 = 0;  // algorithmic bytes per SpMV
    std::string label = "op";          // timing attribution
    // near-null vectors (n x nnull, column-major) handed to the multigrid setup; empty = phase rule
    std::vector<double> nullvec;
    int nnull = 0;
    // host CSR copy (sorted columns)
    metriko::HVec<int64_t> h_ptr;
    metriko::HVec<int> h_col;
    metriko::HVec<double> h_val;
    // host CSR used only by the multigrid set-up when the operator was assembled on the
    // device (its values may differ from the device's in the last bit; released after use)
    std::vector<int64_t> mg_ptr;
    metriko::HVec<int> mg_col;     // filled in parallel after a resize that does not zero
    metriko::HVec<double> mg_val;
};

namespace mk {

// ---- kernels / launchers (kernels.hip) -------------------------------------------------
// PLAIN: y = A x; DOT: y = A x and x^H y; RESID: y = b - A x and |y|^2;
// RESID_NORED: y = b - A x; ACCUM: y += A x
enum SpmvMode { SPMV_PLAIN = 0, SPMV_DOT = 1, SPMV_RESID = 2, SPMV_RESID_NORED = 3, SPMV_ACCUM = 4 };
// fused Chebyshev step: r = b - A x_in ; d = a d + s D^-1 r (first: d = s D^-1 r) ; x_out = x_in + d
template <typename T>
void cheb_launch(mk_ctx* c, const mk_op* A, const T* x_in, T* x_out, const T* b, T* d, T a, T s, int first,
                 const int* skip);
// zero-start Chebyshev steps 1+2 fused: x_out = x1 + d, x1 = c0 D^-1 b, d = a x1 + s D^-1 (b - A x1)
template <typename T>
void cheb0_launch(mk_ctx* c, const mk_op* A, const T* b, T* x_out, T* d, T c0, T a, T s, const int* skip);
// y = M [u; v], columns [0, m1) of M act on u, the rest on v
template <typename T>
void spmv2_launch(mk_ctx* c, const mk_op* A, int m1, const T* u, const T* v, T* y, const int* skip);
template <typename T>
void copy_launch(mk_ctx* c, size_t len, const T* x, T* y, const int* skip);
template <typename T>
void spmm_launch(mk_ctx* c, const mk_op* A, int mw, const T* X, T* Y, const T* B, SpmvMode mode, const int* skip);
// one three-term Chebyshev step fused with its residual on a block of mw right-hand sides,
// r = b - A x_in: mode 5 x_out = x_in + cs D^-1 r; 6 x_out = x_in + ca (x_in - x_out) + cs D^-1 r
// (x_out holds the previous iterate on entry); 7 x_out = (1 + ca) x_in + cs D^-1 r; 8 and 9:
// modes 7 and 6 with x_in resp. the previous iterate replaced by x_1 = c0 D^-1 b
template <typename T>
void spmm_cheb_launch(mk_ctx* c, const mk_op* A, int mw, const T* x_in, T* x_out, const T* b, T ca, T cs, int mode,
                      const int* skip, T c0 = T(0));
// A X = B for mw right-hand sides (B, X: n x mw row-major, device); X starts from zero.
// fp64_vcycle: the V-cycle in FP64 (the plain FP64 form of the same algorithm) instead of FP32
BlockResult block_pcg(mk_ctx* c, const mk_op* A, Mg* mg, int mw, const double* B, double* X, double rel_tol,
                      int maxiter, int check_every, bool fp64_vcycle = false);


// values are multiplied by scale on upload.  Renumbered operator (host arrays, either may be
// null): stored row i is CSR row rsrc[i] and CSR column c is stored as cmap[c]; each row keeps
// its entries in CSR order, so products sum in the same order as the unrenumbered operator.
void build_sell(mk_op* op, const int64_t* rowptr, const int32_t* col, const double* val, int C, int sigma,
                double scale = 1.0, const int* rsrc = nullptr, const int* cmap = nullptr);
// y = A x (mode PLAIN); y = A x and partial x^H y (DOT); y = b - A x and partial |y|^2 (RESID)
void spmv64(mk_ctx* c, const mk_op* A, const double* x, double* y, const double* b, SpmvMode mode, double* partials,
            int* nparts);
void spmv32(mk_ctx* c, const mk_op* A, const float* x, float* y, SpmvMode mode, double* partials, int* nparts);

// reductions: host-visible result of sum(partials[k*np + i]) for k < K
void reduce_partials(mk_ctx* c, const double* partials, int np, int K, double* out_dev);

// vector kernels (len = number of scalar entries; complex vectors count 2 per element)
void axpby64(mk_ctx* c, size_t len, double a, const double* x, double b, double* y);  // y = a x + b y
void scale_dev64(mk_ctx* c, size_t len, const double* s_dev, int invert_sqrt, double* y);
void dot64(mk_ctx* c, size_t len, const double* x, const double* y, double* out_dev);  // real dot of interleaved
void mass_mul(mk_ctx* c, int n, const double* mass, const double* u, double* b);      // complex b = M u
void mass_dot(mk_ctx* c, int n, const double* mass, const double* x, double* out_dev); // x^H M x
void cvt64to32_scaled(mk_ctx* c, size_t len, const double* x, const double* s_dev, float* y); // y = x / sqrt(s)
void add32to64_scaled(mk_ctx* c, size_t len, const float* d, const double* s_dev, double* x);  // x += sqrt(s) d

// CG on the device: solves A x = b for FP64 (x,b device) -- returns status, iterations
struct CgResult { int status; int iters; double rr; double bb; };
CgResult cg64(mk_ctx* c, const mk_op* A, const double* b, double* x, double rel_tol, int maxiter, bool jacobi,
              int check_every, bool x_is_zero, Mg* mg = nullptr);
CgResult cg32(mk_ctx* c, const mk_op* A, const float* b, float* x, double rel_tol, int maxiter, bool jacobi,
              int check_every, Mg* mg = nullptr);

// refinement driver: x (device FP64) warm start, b device FP64; pre overrides A's own multigrid.
// target (when below o.tol): the relative residual the iteration aims at; the result only has
// to meet o.tol, so a target under the FP64 floor of the true residual costs no restarts.
struct SolveResult { int status; int iters; int refinements; double rel_residual; };
SolveResult solve_spd(mk_ctx* c, const mk_op* A, const double* b, double* x, const mk_solve_opts& o,
                      Mg* pre = nullptr, double target = 0.0);

inline Work& work(mk_ctx* c) { return c->work; }

void stats_reset(mk_ctx* c);
// y = G^+ g for a Hermitian positive semi-definite m x m G (row-major); eigenvalues below
// 1e-13 of the largest are dropped (Galerkin starts over nearly parallel windows).
std::vector<std::complex<double>> herm_pinv_solve(int m, const std::vector<std::complex<double>>& G,
                                                  const std::vector<std::complex<double>>& g);
void stats_fill(mk_ctx* c, mk_stats* s);
void timing_flush(mk_ctx* c);

}  // namespace mk
