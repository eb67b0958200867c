// Allocation-free variant of mk_spd_solve (mk_spd_solve_noalloc).
//
// Some callers cannot allocate device memory during a solve: code under HIP graph capture,
// tight serving loops, harnesses that forbid it.  This entry point runs mk_spd_solve's FP64
// PCG (Jacobi or no preconditioner) on caller-owned device vectors with every piece of
// scratch -- the Krylov vectors r, z, p, q, the residual check's t, the per-block partial sums
// of the two-launch reductions, the CG state and the device scalars -- taken from one static
// per-device array, `scratch`, instead of the context's buffers (which grow by hipMalloc).
//
// Limits: at most kNoAllocLen scalar unknowns (real n, or 2 n for complex operators); the
// array holds 5 vectors of that length plus the reduction partials, 24 MiB per GPU, kept
// until the process ends.  All calls on one GPU share that GPU's copy of the array and the
// last call's contents stay in it: calls that overlap on one GPU (two host threads, two
// contexts) race.  The CG state and scalars are cleared at the start of every call.
#include <chrono>
#include <cstring>
#include <mutex>
#include "mk_internal.h"
#include "sell.h"

namespace mk {
namespace noalloc {
constexpr size_t kNoAllocLen = size_t(1) << 19;           // scalars per work vector
constexpr size_t kPartials = kMaxParts * size_t(kMaxK);   // two-launch reduction partials
constexpr size_t kStateDoubles = 32;                      // CgState (<= 256 bytes)
constexpr size_t kScalDoubles = 64;                       // device scalars (as mk_ctx::scal)
constexpr size_t kVectors = 5;                            // r, z, p, q, t
constexpr size_t kVecOffset = kPartials + kStateDoubles + kScalDoubles;
constexpr size_t kScratchDoubles = kVecOffset + kVectors * kNoAllocLen;
static_assert(sizeof(CgState) <= kStateDoubles * sizeof(double), "CgState does not fit its scratch slot");
static_assert(kVecOffset % 32 == 0, "work vectors must stay 256-byte aligned");
constexpr int kMaxDevices = 64;
}  // namespace noalloc
}  // namespace mk

// the scratch of mk_spd_solve_noalloc: one copy per GPU, in device memory
static __device__ __align__(256) double scratch[mk::noalloc::kScratchDoubles];

namespace {
// device address of `scratch` per device (host code never uses the array's name directly),
// or why it could not be looked up
double* g_scratch_ptr[mk::noalloc::kMaxDevices] = {};
std::string g_scratch_err[mk::noalloc::kMaxDevices];
std::mutex g_scratch_mu;

__global__ void k_fill_zero(size_t n, double* p) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Points a context buffer at a slice of the scratch for the duration of one call (no
// allocation, no free); the previous pointer and size come back when the call ends.
template <typename T>
struct Rebind {
    mk::DevBuf<T>& buf;
    T* p0;
    size_t n0;
    Rebind(mk::DevBuf<T>& b, T* p, size_t n) : buf(b), p0(b.p), n0(b.n) { b.p = p; b.n = n; }
    ~Rebind() { buf.p = p0; buf.n = n0; }
    bool intact(const T* p) const { return buf.p == p; }
};
}  // namespace

namespace mk {
// Looks up this device's copy of `scratch` (the device must be current); once per device.
mk_status noalloc_setup(int device, std::string* err) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= noalloc::kMaxDevices) {
        if (err) *err = "mk_spd_solve_noalloc: device id beyond the scratch table";
        return MK_ERR_ARG;
    }
    if (g_scratch_ptr[device]) return MK_OK;
    double* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        if (err) *err = g_scratch_err[device];
        return MK_ERR_HIP;
    }
    g_scratch_ptr[device] = p;
    return MK_OK;
}
}  // namespace mk

extern "C" mk_status mk_spd_solve_noalloc(mk_ctx* c, const mk_op* A, const double* b, double* x,
                                          const mk_solve_opts* o_, mk_stats* st) {
    using namespace mk::noalloc;
    if (!c || !A || !b || !x || A->ctx != c) return MK_ERR_ARG;
    mk_status inner = MK_OK;
    try {
        mk_solve_opts o;
        mk_default_opts(&o);
        if (o_) o = *o_;
        MK_CHECK(o.precision == MK_PREC_FP64, MK_ERR_UNSUPPORTED,
                 "mk_spd_solve_noalloc: FP64 PCG only (the FP32 modes need FP32 work vectors)");
        MK_CHECK(o.precond == MK_PRECOND_NONE || o.precond == MK_PRECOND_JACOBI, MK_ERR_UNSUPPORTED,
                 "mk_spd_solve_noalloc: Jacobi or no preconditioner (a multigrid set-up allocates)");
        const size_t len = size_t(A->n) * (A->cplx ? 2 : 1);
        MK_CHECK(len <= kNoAllocLen, MK_ERR_ARG,
                 "mk_spd_solve_noalloc: " + std::to_string(len) + " scalar unknowns exceed the scratch limit of " +
                     std::to_string(kNoAllocLen) + " (real n, or 2 n for complex operators)");
        MK_CHECK(c->device >= 0 && c->device < kMaxDevices, MK_ERR_ARG, "mk_spd_solve_noalloc: device id");
        {
            std::lock_guard<std::mutex> lk(g_scratch_mu);
            MK_CHECK(g_scratch_ptr[c->device], MK_ERR_HIP,
                     "mk_spd_solve_noalloc: no scratch on this device: " +
                         (g_scratch_err[c->device].empty() ? std::string("mk_init did not set it up")
                                                           : g_scratch_err[c->device]));
        }
        MK_HIP(hipSetDevice(c->device));
        const auto t0 = std::chrono::steady_clock::now();
        double* base = g_scratch_ptr[c->device];
        double* partials = base;
        mk::CgState* state = reinterpret_cast<mk::CgState*>(base + kPartials);
        double* scal = base + kPartials + kStateDoubles;
        double* vec = base + kVecOffset;
        mk::Work& w = mk::work(c);
        {
            Rebind<double> rp(c->partials, partials, kPartials);
            Rebind<mk::CgState> rs(c->state, state, 1);
            Rebind<double> rc(c->scal, scal, kScalDoubles);
            Rebind<double> r1(w.r, vec + 0 * kNoAllocLen, kNoAllocLen);
            Rebind<double> r2(w.z, vec + 1 * kNoAllocLen, kNoAllocLen);
            Rebind<double> r3(w.p, vec + 2 * kNoAllocLen, kNoAllocLen);
            Rebind<double> r4(w.q, vec + 3 * kNoAllocLen, kNoAllocLen);
            Rebind<double> r5(w.t, vec + 4 * kNoAllocLen, kNoAllocLen);
            const bool timing = c->time_spmv;
            c->time_spmv = false;  // event timing creates events on first use
            struct RestoreTiming { mk_ctx* c; bool on; ~RestoreTiming() { c->time_spmv = on; } } rt{c, timing};
            // the CG state and the scalars start from zero (a fill kernel, not a memset node)
            const size_t nz = kStateDoubles + kScalDoubles;
            hipLaunchKernelGGL(k_fill_zero, dim3(unsigned((nz + 255) / 256)), dim3(256), 0, c->stream, nz,
                               base + kPartials);
            MK_HIP(hipGetLastError());
            mk::stats_reset(c);
            const mk::SolveResult r = mk::solve_spd(c, A, b, x, o);
            MK_HIP(hipStreamSynchronize(c->stream));
            MK_CHECK(rp.intact(partials) && rs.intact(state) && rc.intact(scal) && r1.intact(vec) &&
                         r5.intact(vec + 4 * kNoAllocLen),
                     MK_ERR_HIP, "mk_spd_solve_noalloc: a work buffer was reallocated");
            if (st) {
                std::memset(st, 0, sizeof(*st));
                st->iterations = r.iters;
                st->refinements = r.refinements;
                st->solves = 1;
                st->rel_residual = r.rel_residual;
                mk::stats_fill(c, st);
                st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            }
            inner = r.status;
        }
        if (inner != MK_OK) c->last_error = "mk_spd_solve_noalloc: " + std::string(mk_status_str(inner));
    } catch (const mk::Error& e) {
        c->last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        c->last_error = e.what();
        return MK_ERR_HIP;
    }
    return inner;
}
