// Batched configuration: many closed meshes of uneven size solved together.  No reference
// counterpart as a whole: per mesh it runs the Smoothest field of face_rosy_field.h:17-44
// (50-step inverse iteration, solveSmallestEig, sparse_linear_solver.h:29-54) and then
// integrates the field's first direction and its quarter turn as two Poisson problems.
//
// MI355X layout: the meshes are laid end to end as one block-diagonal operator (one SELL
// matrix, one multigrid hierarchy whose aggregates never cross meshes), so every kernel
// launch serves all meshes at once whatever their sizes.  What must stay per mesh -- the
// Rayleigh quotient of the start vector, the mass normalisation, the scaling of each
// right-hand side -- runs as segmented reductions over element blocks that never straddle
// two meshes (fixed-order sums, deterministic).  Each mesh's right-hand side is scaled to
// unit norm and the global solve runs to tol / sqrt(#meshes), which bounds every mesh's own
// relative residual by tol.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include "mk_internal.h"
#include "sell.h"
#include "metriko/core/hmesh/hmesh.h"

using mk::DevBuf;
using mk::Error;

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr int kSegBlock = 2048;  // elements per block of the segmented reductions

// Blocks of at most kSegBlock elements, none straddling two segments (meshes).
struct SegTable {
    int nseg = 0, nblk = 0;
    std::vector<int> hseg, hsblk;
    std::vector<int64_t> hbeg, hend;
    DevBuf<int> seg, sblk;
    DevBuf<int64_t> beg, end;
    void build(const std::vector<int64_t>& off, hipStream_t st) {
        nseg = int(off.size()) - 1;
        hsblk.assign(static_cast<size_t>(nseg) + 1, 0);
        for (int m = 0; m < nseg; m++) {
            for (int64_t b = off[static_cast<size_t>(m)]; b < off[static_cast<size_t>(m) + 1]; b += kSegBlock) {
                hseg.push_back(m);
                hbeg.push_back(b);
                hend.push_back(std::min<int64_t>(b + kSegBlock, off[static_cast<size_t>(m) + 1]));
            }
            hsblk[static_cast<size_t>(m) + 1] = int(hseg.size());
        }
        nblk = int(hseg.size());
        MK_CHECK(nblk > 0 && static_cast<size_t>(nblk) * 4 <= mk::kMaxParts * mk::kMaxK, MK_ERR_ARG, "batch too large");
        seg.upload(hseg.data(), hseg.size(), st);
        sblk.upload(hsblk.data(), hsblk.size(), st);
        beg.upload(hbeg.data(), hbeg.size(), st);
        end.upload(hend.data(), hend.size(), st);
    }
};

// q0 = Re(u^H L u), q1 = u^H M u, q2 = |b|^2 over each block's (complex) elements
__global__ void k_seg_rayleigh(const int64_t* __restrict__ beg, const int64_t* __restrict__ end,
                               const double* __restrict__ u, const double* __restrict__ Lu,
                               const double* __restrict__ b, const double* __restrict__ mass,
                               double* __restrict__ part) {
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t i = beg[blockIdx.x] + threadIdx.x; i < end[blockIdx.x]; i += blockDim.x) {
        const double ur = u[2 * i], ui = u[2 * i + 1];
        v[0] += ur * Lu[2 * i] + ui * Lu[2 * i + 1];
        v[1] += mass[i] * (ur * ur + ui * ui);
        v[2] += b[2 * i] * b[2 * i] + b[2 * i + 1] * b[2 * i + 1];
    }
    mk::block_sum<3>(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) part[static_cast<size_t>(blockIdx.x) * 3 + k] = v[k];
}
// Galerkin start of each mesh's inverse iteration on the window of its last kBWin iterates
// (the single-mesh field's start, per segment): per block, w_j^H a and w_j^H b (complex) for
// the nw window columns (column j at W + len j), 4 doubles per column.
constexpr int kBWin = 4;
constexpr int kSegMaxK = 4 * kBWin;
__global__ void k_seg_window_dots(const int64_t* __restrict__ beg, const int64_t* __restrict__ end, int nw,
                                  size_t len, const double* __restrict__ W, const double* __restrict__ a,
                                  const double* __restrict__ b, double* __restrict__ part) {
    double acc[kSegMaxK];
#pragma unroll
    for (int k = 0; k < kSegMaxK; k++) acc[k] = 0.0;
    for (int64_t i = beg[blockIdx.x] + threadIdx.x; i < end[blockIdx.x]; i += blockDim.x) {
        const double ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
#pragma unroll
        for (int j = 0; j < kBWin; j++) {
            if (j < nw) {
                const double wr = W[size_t(j) * len + 2 * i], wi = W[size_t(j) * len + 2 * i + 1];
                acc[4 * j] += wr * ar + wi * ai;
                acc[4 * j + 1] += wr * ai - wi * ar;
                acc[4 * j + 2] += wr * br + wi * bi;
                acc[4 * j + 3] += wr * bi - wi * br;
            }
        }
    }
    mk::block_sum<kSegMaxK>(acc);
    if (threadIdx.x == 0)
        for (int k = 0; k < kSegMaxK; k++) part[static_cast<size_t>(blockIdx.x) * kSegMaxK + k] = acc[k];
}
// x = sum_j y[seg][j] w_j over each block's elements (complex coefficients, 2 kBWin per segment)
__global__ void k_seg_window_combine(const int* __restrict__ seg, const int64_t* __restrict__ beg,
                                     const int64_t* __restrict__ end, int nw, size_t len,
                                     const double* __restrict__ W, const double* __restrict__ Y,
                                     double* __restrict__ x) {
    const double* y = Y + static_cast<size_t>(seg[blockIdx.x]) * 2 * kBWin;
    for (int64_t i = beg[blockIdx.x] + threadIdx.x; i < end[blockIdx.x]; i += blockDim.x) {
        double xr = 0.0, xi = 0.0;
        for (int j = 0; j < nw; j++) {
            const double wr = W[size_t(j) * len + 2 * i], wi = W[size_t(j) * len + 2 * i + 1];
            xr += y[2 * j] * wr - y[2 * j + 1] * wi;
            xi += y[2 * j] * wi + y[2 * j + 1] * wr;
        }
        x[2 * i] = xr;
        x[2 * i + 1] = xi;
    }
}
// q0 = x^H M x (complex elements)
__global__ void k_seg_massnorm(const int64_t* __restrict__ beg, const int64_t* __restrict__ end,
                               const double* __restrict__ x, const double* __restrict__ mass,
                               double* __restrict__ part) {
    double v[1] = {0.0};
    for (int64_t i = beg[blockIdx.x] + threadIdx.x; i < end[blockIdx.x]; i += blockDim.x)
        v[0] += mass[i] * (x[2 * i] * x[2 * i] + x[2 * i + 1] * x[2 * i + 1]);
    mk::block_sum<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}
// q0 = |x|^2 (real elements)
__global__ void k_seg_sumsq(const int64_t* __restrict__ beg, const int64_t* __restrict__ end,
                            const double* __restrict__ x, double* __restrict__ part) {
    double v[1] = {0.0};
    for (int64_t i = beg[blockIdx.x] + threadIdx.x; i < end[blockIdx.x]; i += blockDim.x) v[0] += x[i] * x[i];
    mk::block_sum<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}
// out[seg * K + k] = sum of the K partials of the segment's blocks, in block order
__global__ void k_seg_finish(int nseg, int K, const int* __restrict__ sblk, const double* __restrict__ part,
                             double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nseg * K) return;
    const int m = t / K, k = t - m * K;
    double s = 0.0;
    for (int b = sblk[m]; b < sblk[m + 1]; b++) s += part[static_cast<size_t>(b) * K + k];
    out[t] = s;
}
// y = s[segment] * x over each block's elements of vw doubles (y may alias x)
__global__ void k_seg_scale(const int* __restrict__ seg, const int64_t* __restrict__ beg,
                            const int64_t* __restrict__ end, int vw, const double* __restrict__ s,
                            const double* x, double* y) {
    const double f = s[seg[blockIdx.x]];
    for (int64_t i = beg[blockIdx.x] * vw + threadIdx.x; i < end[blockIdx.x] * vw; i += blockDim.x) y[i] = f * x[i];
}

struct Seg {
    mk_ctx* c;
    SegTable& T;
    DevBuf<double> part, out, s;
    std::vector<double> h;
    Seg(mk_ctx* c_, SegTable& T_) : c(c_), T(T_) {
        part.alloc(static_cast<size_t>(T.nblk) * kSegMaxK);
        out.alloc(static_cast<size_t>(T.nseg) * kSegMaxK);
        s.alloc(static_cast<size_t>(T.nseg));
    }
    // K totals per segment to the host
    const std::vector<double>& finish(int K) {
        const int tot = T.nseg * K;
        hipLaunchKernelGGL(k_seg_finish, dim3((tot + 255) / 256), dim3(256), 0, c->stream, T.nseg, K,
                           (const int*)T.sblk.p, (const double*)part.p, out.p);
        h.resize(static_cast<size_t>(tot));
        out.download(h.data(), h.size(), c->stream);
        MK_HIP(hipStreamSynchronize(c->stream));
        return h;
    }
    void scale(int vw, const std::vector<double>& f, const double* x, double* y) {
        s.upload(f.data(), f.size(), c->stream);
        hipLaunchKernelGGL(k_seg_scale, dim3(T.nblk), dim3(256), 0, c->stream, (const int*)T.seg.p,
                           (const int64_t*)T.beg.p, (const int64_t*)T.end.p, vw, (const double*)s.p, x, y);
    }
};

struct MeshData {
    int nV = 0, nF = 0;
    std::vector<int64_t> lptr;          // field Laplacian rows (complex, columns local)
    std::vector<int> lcol;
    std::vector<double> lval;
    std::vector<double> area;           // face areas (the mass)
    std::vector<double> bx, by, nrm;    // face basis and unit normal, 3 per face
    std::vector<int64_t> sptr;          // stiffness rows (real, columns local), vertex 0 pinned
    std::vector<int> scol;
    std::vector<double> sval;
    std::vector<double> gphi;           // hat gradients x area: 9 per face (3 vertices x xyz)
    std::vector<int> fv;                // face vertices
};

// Per mesh: Hmesh geometry, connection Laplacian rows (face_rosy_field.h:46-76: per face and
// interior neighbour g across h, (f,f) += 1 and (f,g) = -r^N with r = -e^{i farg(h)} /
// e^{i farg(twin h)}; + 1e-9 on the diagonal), face frames, and the piecewise-linear
// stiffness with vertex 0 pinned.
void mesh_setup(int nV, int nF, const double* V, const int32_t* F, int N, MeshData& D) {
    using namespace metriko;
    MatXd Vm(nV, 3);
    std::memcpy(Vm.data(), V, sizeof(double) * static_cast<size_t>(nV) * 3);
    MatXi Fm(nF, 3);
    for (int64_t i = 0; i < int64_t(nF) * 3; i++) {
        if (F[i] < 0 || F[i] >= nV) throw Error(MK_ERR_ARG, "face index out of range");
        Fm.d[static_cast<size_t>(i)] = F[i];
    }
    Hmesh m(Vm, Fm);
    D.nV = nV;
    D.nF = nF;
    const double sgn = (N % 2 == 0) ? -1.0 : 1.0;
    D.lptr.assign(static_cast<size_t>(nF) + 1, 0);
    D.area.resize(static_cast<size_t>(nF));
    D.bx.resize(static_cast<size_t>(nF) * 3);
    D.by.resize(static_cast<size_t>(nF) * 3);
    D.nrm.resize(static_cast<size_t>(nF) * 3);
    for (int f = 0; f < nF; f++) {
        int c[4];
        double vr[4], vi[4];
        int cnt = 0;
        double diag = 0.0;
        int h = m.face2half[f];
        for (int k = 0; k < 3; k++, h = m.next[h]) {
            const int tw = m.twin[h], g = m.face[tw];
            if (g < 0) continue;
            diag += 1.0;
            const double a = double(N) * (m.heArgOnFace[h] - m.heArgOnFace[tw]);
            c[cnt] = g;
            vr[cnt] = sgn * std::cos(a);
            vi[cnt] = sgn * std::sin(a);
            cnt++;
        }
        c[cnt] = f;
        vr[cnt] = diag + 1e-9;
        vi[cnt] = 0.0;
        cnt++;
        for (int i = 1; i < cnt; i++)  // columns ascending
            for (int j = i; j > 0 && c[j - 1] > c[j]; j--) {
                std::swap(c[j - 1], c[j]);
                std::swap(vr[j - 1], vr[j]);
                std::swap(vi[j - 1], vi[j]);
            }
        for (int i = 0; i < cnt; i++) {
            D.lcol.push_back(c[i]);
            D.lval.push_back(vr[i]);
            D.lval.push_back(vi[i]);
        }
        D.lptr[static_cast<size_t>(f) + 1] = int64_t(D.lcol.size());
        D.area[static_cast<size_t>(f)] = m.faceArea[f];
        for (int k = 0; k < 3; k++) {
            D.bx[static_cast<size_t>(3 * f + k)] = m.faceBasisX(f, k);
            D.by[static_cast<size_t>(3 * f + k)] = m.faceBasisY(f, k);
            D.nrm[static_cast<size_t>(3 * f + k)] = m.faceNormal(f, k);
        }
    }
    // hat gradients: grad phi_i = n x (p_{i+2} - p_{i+1}) / (2A) on face (p0, p1, p2)
    D.fv.assign(F, F + static_cast<size_t>(nF) * 3);
    D.gphi.resize(static_cast<size_t>(nF) * 9);
    std::vector<std::vector<std::pair<int, double>>> rows(static_cast<size_t>(nV));
    for (int f = 0; f < nF; f++) {
        const double* p[3] = {V + 3 * static_cast<size_t>(F[3 * f]), V + 3 * static_cast<size_t>(F[3 * f + 1]), V + 3 * static_cast<size_t>(F[3 * f + 2])};
        double e1[3], e2[3], n[3];
        for (int k = 0; k < 3; k++) { e1[k] = p[1][k] - p[0][k]; e2[k] = p[2][k] - p[0][k]; }
        n[0] = e1[1] * e2[2] - e1[2] * e2[1];
        n[1] = e1[2] * e2[0] - e1[0] * e2[2];
        n[2] = e1[0] * e2[1] - e1[1] * e2[0];
        const double dbl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; k++) n[k] /= dbl;
        const double A = 0.5 * dbl;
        double G[3][3];
        for (int i = 0; i < 3; i++) {
            const double* a = p[(i + 1) % 3];
            const double* b = p[(i + 2) % 3];
            const double e[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            G[i][0] = (n[1] * e[2] - n[2] * e[1]) / dbl;
            G[i][1] = (n[2] * e[0] - n[0] * e[2]) / dbl;
            G[i][2] = (n[0] * e[1] - n[1] * e[0]) / dbl;
            for (int k = 0; k < 3; k++) D.gphi[static_cast<size_t>(9 * f + 3 * i + k)] = A * G[i][k];
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                rows[static_cast<size_t>(F[3 * f + i])].push_back({F[3 * f + j], A * (G[i][0] * G[j][0] + G[i][1] * G[j][1] + G[i][2] * G[j][2])});
    }
    D.sptr.assign(static_cast<size_t>(nV) + 1, 0);
    for (int v = 0; v < nV; v++) {
        auto& r = rows[static_cast<size_t>(v)];
        std::stable_sort(r.begin(), r.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        if (v == 0) {  // pinned: identity row
            D.scol.push_back(0);
            D.sval.push_back(1.0);
        } else {
            for (size_t i = 0; i < r.size();) {
                const int cc = r[i].first;
                double s = 0.0;
                while (i < r.size() && r[i].first == cc) s += r[i++].second;
                if (cc == 0) continue;  // column of the pinned vertex
                D.scol.push_back(cc);
                D.sval.push_back(s);
            }
        }
        D.sptr[static_cast<size_t>(v) + 1] = int64_t(D.scol.size());
    }
}

}  // namespace

extern "C" {

mk_status mk_batch_run(mk_ctx* c, int nmesh, const int32_t* nV, const int32_t* nF, const double* const* V,
                       const int32_t* const* F, int N, int n_iter, const double* const* u0,
                       const mk_solve_opts* opts, double* const* compressed, double* const* uv,
                       mk_batch_stats* stats) {
    if (!c || nmesh <= 0 || !nV || !nF || !V || !F || N <= 0 || n_iter < 0) return MK_ERR_ARG;
    for (int m = 0; m < nmesh; m++)
        if (nV[m] <= 0 || nF[m] <= 0 || !V[m] || !F[m]) return MK_ERR_ARG;
    try {
        mk_solve_opts o;
        mk_default_opts(&o);
        if (opts) o = *opts;
        MK_HIP(hipSetDevice(c->device));
        mk_batch_stats S{};
        S.meshes = nmesh;
        const double t_all = now_ms();
        // ---- host: per-mesh geometry and operator rows (OpenMP over meshes) ----
        std::vector<MeshData> D(static_cast<size_t>(nmesh));
        std::vector<std::string> err(static_cast<size_t>(nmesh));
#pragma omp parallel for schedule(dynamic, 1)
        for (int m = 0; m < nmesh; m++) {
            try {
                mesh_setup(nV[m], nF[m], V[m], F[m], N, D[static_cast<size_t>(m)]);
            } catch (const std::exception& e) {
                err[static_cast<size_t>(m)] = e.what();
            }
        }
        for (int m = 0; m < nmesh; m++)
            if (!err[static_cast<size_t>(m)].empty()) throw Error(MK_ERR_ARG, "mesh " + std::to_string(m) + ": " + err[static_cast<size_t>(m)]);
        std::vector<int64_t> foff(static_cast<size_t>(nmesh) + 1, 0), voff(static_cast<size_t>(nmesh) + 1, 0), lnz(static_cast<size_t>(nmesh) + 1, 0),
            snz(static_cast<size_t>(nmesh) + 1, 0);
        for (int m = 0; m < nmesh; m++) {
            foff[static_cast<size_t>(m) + 1] = foff[static_cast<size_t>(m)] + nF[m];
            voff[static_cast<size_t>(m) + 1] = voff[static_cast<size_t>(m)] + nV[m];
            lnz[static_cast<size_t>(m) + 1] = lnz[static_cast<size_t>(m)] + int64_t(D[static_cast<size_t>(m)].lcol.size());
            snz[static_cast<size_t>(m) + 1] = snz[static_cast<size_t>(m)] + int64_t(D[static_cast<size_t>(m)].scol.size());
        }
        const int64_t nFt = foff.back(), nVt = voff.back();
        MK_CHECK(nFt < (int64_t(1) << 31) && nVt < (int64_t(1) << 31), MK_ERR_ARG, "batch too large");
        S.faces = nFt;
        S.verts = nVt;
        // union (block-diagonal) operators
        std::vector<int64_t> Lp(static_cast<size_t>(nFt) + 1, 0);
        metriko::HVec<int64_t> Sp(static_cast<size_t>(nVt) + 1, 0);
        std::vector<int> Lc(static_cast<size_t>(lnz.back()));
        metriko::HVec<int> Sc(static_cast<size_t>(snz.back()));
        std::vector<double> Lv(static_cast<size_t>(lnz.back()) * 2), mass(static_cast<size_t>(nFt));
        metriko::HVec<double> Sv(static_cast<size_t>(snz.back()));
#pragma omp parallel for schedule(dynamic, 1)
        for (int m = 0; m < nmesh; m++) {
            const MeshData& d = D[static_cast<size_t>(m)];
            const int64_t f0 = foff[static_cast<size_t>(m)], v0 = voff[static_cast<size_t>(m)], l0 = lnz[static_cast<size_t>(m)], s0 = snz[static_cast<size_t>(m)];
            for (int f = 0; f < d.nF; f++) Lp[static_cast<size_t>(f0 + f) + 1] = l0 + d.lptr[static_cast<size_t>(f) + 1];
            for (size_t k = 0; k < d.lcol.size(); k++) Lc[static_cast<size_t>(l0) + k] = int(f0) + d.lcol[k];
            std::memcpy(Lv.data() + 2 * l0, d.lval.data(), sizeof(double) * d.lval.size());
            std::memcpy(mass.data() + f0, d.area.data(), sizeof(double) * static_cast<size_t>(d.nF));
            for (int v = 0; v < d.nV; v++) Sp[static_cast<size_t>(v0 + v) + 1] = s0 + d.sptr[static_cast<size_t>(v) + 1];
            for (size_t k = 0; k < d.scol.size(); k++) Sc[static_cast<size_t>(s0) + k] = int(v0) + d.scol[k];
            std::memcpy(Sv.data() + s0, d.sval.data(), sizeof(double) * d.sval.size());
        }
        // ---- autotuned layouts of the field and stiffness operators (samples of their rows) ----
        // the leading rows (up to 262,144), couplings inside the sample kept
        auto tune_layout = [&](int64_t nrows, const auto& P, const auto& Cc, const auto& Vv, int cplx,
                               mk_solve_opts& best, double& us) {
            const int ns = int(std::min<int64_t>(nrows, 262144));
            const int w = cplx ? 2 : 1;
            std::vector<int64_t> sp(static_cast<size_t>(ns) + 1, 0);
            std::vector<int> sc;
            std::vector<double> sv;
            for (int i = 0; i < ns; i++) {
                for (int64_t q = P[static_cast<size_t>(i)]; q < P[static_cast<size_t>(i) + 1]; q++)
                    if (Cc[static_cast<size_t>(q)] < ns) {
                        sc.push_back(Cc[static_cast<size_t>(q)]);
                        for (int k = 0; k < w; k++) sv.push_back(Vv[static_cast<size_t>(w * q + k)]);
                    }
                sp[static_cast<size_t>(i) + 1] = int64_t(sc.size());
            }
            MK_CHECK(mk_autotune_spmv(c, ns, sp.data(), sc.data(), sv.data(), cplx, &o, &best, &us) == MK_OK,
                     MK_ERR_HIP, "autotune: " + c->last_error);
        };
        mk_solve_opts of = o, osl = o;
        {
            double us = 0.0;
            tune_layout(nFt, Lp, Lc, Lv, 1, of, us);
            S.tuned_slice_height = of.slice_height;
            S.tuned_sigma = of.sigma;
            S.tuned_block = of.block_size;
            S.tuned_spmv_us = us;
            tune_layout(nVt, Sp, Sc, Sv, 0, osl, us);
            S.tuned_poisson_slice_height = osl.slice_height;
            S.tuned_poisson_sigma = osl.sigma;
            S.tuned_poisson_block = osl.block_size;
            S.tuned_poisson_spmv_us = us;
        }
        const int64_t replays0 = c->graph_replays, captures0 = c->graph_captures, eager0 = c->eager_batches;
        S.ms_host_setup = now_ms() - t_all;
        // ---- field: batched inverse iteration ----
        double t0 = now_ms();
        mk_op* Lop = nullptr;
        MK_CHECK(mk_op_create_complex(c, int(nFt), Lp.data(), Lc.data(), Lv.data(), &of, &Lop) == MK_OK, MK_ERR_HIP,
                 "field operator: " + c->last_error);
        std::unique_ptr<mk_op, void (*)(mk_op*)> Lhold(Lop, [](mk_op* p) { mk_op_destroy(p); });
        Lop->label = "batch_field_L";
        mk::ensure_mg(c, Lop, of);
        SegTable TF, TV;
        TF.build(foff, c->stream);
        TV.build(voff, c->stream);
        Seg SF(c, TF), SV(c, TV);
        const size_t lenF = static_cast<size_t>(nFt) * 2;
        DevBuf<double> dm, du, dx, db, dLu;
        dm.upload(mass.data(), mass.size(), c->stream);
        {
            std::vector<double> hu(lenF);
            for (int m = 0; m < nmesh; m++) {
                const size_t o2 = static_cast<size_t>(2 * foff[static_cast<size_t>(m)]);
                if (u0 && u0[m]) std::memcpy(hu.data() + o2, u0[m], sizeof(double) * 2 * static_cast<size_t>(nF[m]));
                else
                    for (int f = 0; f < nF[m]; f++) {  // Eigen VecXc::Random: re then im, each -1 + 2 rand()/RAND_MAX
                        hu[o2 + 2 * static_cast<size_t>(f)] = -1.0 + (2.0 * double(std::rand())) / double(RAND_MAX);
                        hu[o2 + 2 * static_cast<size_t>(f) + 1] = -1.0 + (2.0 * double(std::rand())) / double(RAND_MAX);
                    }
            }
            du.upload(hu.data(), hu.size(), c->stream);
        }
        dx.alloc(lenF);
        db.alloc(lenF);
        dLu.alloc(lenF);
        const double gtol = o.tol / std::sqrt(double(nmesh));
        mk_solve_opts og = of;
        og.tol = gtol;
        std::vector<double> f1(static_cast<size_t>(nmesh)), f2(static_cast<size_t>(nmesh));
        // per mesh: Galerkin matrix G = W^H L W of its window (slot order) and W^H b
        DevBuf<double> dW(lenF * kBWin), dY(static_cast<size_t>(nmesh) * 2 * kBWin);
        std::vector<std::complex<double>> G(static_cast<size_t>(nmesh) * kBWin * kBWin),
            gw(static_cast<size_t>(nmesh) * kBWin);
        std::vector<double> hY(static_cast<size_t>(nmesh) * 2 * kBWin, 0.0);
        for (int it = 0; it < n_iter; it++) {
            mk::mass_mul(c, int(nFt), dm.p, du.p, db.p);
            mk::spmv64(c, Lop, du.p, dLu.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
            hipLaunchKernelGGL(k_seg_rayleigh, dim3(TF.nblk), dim3(256), 0, c->stream, (const int64_t*)TF.beg.p,
                               (const int64_t*)TF.end.p, (const double*)du.p, (const double*)dLu.p,
                               (const double*)db.p, (const double*)dm.p, SF.part.p);
            const std::vector<double> h = SF.finish(3);
            // the iterate enters its mesh's window (slot s); the start is the Galerkin solution on
            // the window, x0 = W (W^H L W)^+ W^H b per mesh (with the window {u} alone: the
            // Rayleigh-quotient start u (u^H M u) / (u^H L u)), in the scaled system
            const int slot = it % kBWin, nw = std::min(it + 1, kBWin);
            MK_HIP(hipMemcpyAsync(dW.p + lenF * size_t(slot), du.p, lenF * sizeof(double), hipMemcpyDeviceToDevice,
                                  c->stream));
            hipLaunchKernelGGL(k_seg_window_dots, dim3(TF.nblk), dim3(256), 0, c->stream, (const int64_t*)TF.beg.p,
                               (const int64_t*)TF.end.p, nw, lenF, (const double*)dW.p, (const double*)dLu.p,
                               (const double*)db.p, SF.part.p);
            const std::vector<double>& hd = SF.finish(kSegMaxK);
            for (int m = 0; m < nmesh; m++) {
                const double num = h[static_cast<size_t>(3 * m)], den = h[static_cast<size_t>(3 * m + 1)], bb = h[static_cast<size_t>(3 * m + 2)];
                MK_CHECK(num == num && num > 0.0 && den > 0.0 && bb > 0.0, MK_ERR_NAN,
                         "mesh " + std::to_string(m) + ": Rayleigh quotient not positive");
                const double sb = 1.0 / std::sqrt(bb);
                f2[static_cast<size_t>(m)] = sb;
                const double* q = hd.data() + static_cast<size_t>(m) * kSegMaxK;
                std::complex<double>* Gm = G.data() + static_cast<size_t>(m) * kBWin * kBWin;
                std::complex<double>* gm = gw.data() + static_cast<size_t>(m) * kBWin;
                for (int j = 0; j < nw; j++) {
                    const std::complex<double> gjs(q[4 * j], q[4 * j + 1]);   // w_j^H L w_s
                    Gm[j * kBWin + slot] = gjs;
                    Gm[slot * kBWin + j] = std::conj(gjs);
                    gm[j] = {q[4 * j + 2], q[4 * j + 3]};                      // w_j^H b
                }
                Gm[slot * kBWin + slot] = q[4 * slot];
                std::vector<std::complex<double>> Gn(static_cast<size_t>(nw) * nw), gn(gm, gm + nw);
                for (int i = 0; i < nw; i++)
                    for (int j = 0; j < nw; j++) Gn[static_cast<size_t>(i) * nw + j] = Gm[i * kBWin + j];
                const std::vector<std::complex<double>> y = mk::herm_pinv_solve(nw, Gn, gn);
                for (int j = 0; j < nw; j++) {
                    hY[static_cast<size_t>(m) * 2 * kBWin + 2 * j] = sb * y[static_cast<size_t>(j)].real();
                    hY[static_cast<size_t>(m) * 2 * kBWin + 2 * j + 1] = sb * y[static_cast<size_t>(j)].imag();
                }
            }
            dY.upload(hY.data(), hY.size(), c->stream);
            hipLaunchKernelGGL(k_seg_window_combine, dim3(TF.nblk), dim3(256), 0, c->stream, (const int*)TF.seg.p,
                               (const int64_t*)TF.beg.p, (const int64_t*)TF.end.p, nw, lenF, (const double*)dW.p,
                               (const double*)dY.p, dx.p);
            SF.scale(2, f2, db.p, db.p);
            // each mesh's right-hand side has unit norm: the global tol / sqrt(#meshes) bounds
            // every mesh's own relative residual by tol (aiming at a tenth of it, as the single
            // mesh field does)
            const mk::SolveResult r = mk::solve_spd(c, Lop, db.p, dx.p, og, nullptr, 0.1 * gtol);
            S.field_iterations += r.iters;
            S.field_rel_residual = std::max(S.field_rel_residual, r.rel_residual * std::sqrt(double(nmesh)));
            if (r.status != MK_OK)
                throw Error(r.status, "batched inverse iteration " + std::to_string(it) + ": " + mk_status_str(r.status));
            hipLaunchKernelGGL(k_seg_massnorm, dim3(TF.nblk), dim3(256), 0, c->stream, (const int64_t*)TF.beg.p,
                               (const int64_t*)TF.end.p, (const double*)dx.p, (const double*)dm.p, SF.part.p);
            const std::vector<double>& hs = SF.finish(1);
            for (int m = 0; m < nmesh; m++) {
                MK_CHECK(hs[static_cast<size_t>(m)] > 0.0, MK_ERR_NAN, "mesh " + std::to_string(m) + ": zero iterate");
                f1[static_cast<size_t>(m)] = 1.0 / std::sqrt(hs[static_cast<size_t>(m)]);
            }
            SF.scale(2, f1, dx.p, du.p);
        }
        std::vector<double> hx(lenF);
        du.download(hx.data(), hx.size(), c->stream);
        MK_HIP(hipStreamSynchronize(c->stream));
        S.ms_field = now_ms() - t0;
        // ---- Poisson: u, v from the field's first direction and its quarter turn ----
        t0 = now_ms();
        std::vector<double> du_rhs(static_cast<size_t>(nVt), 0.0), dv_rhs(static_cast<size_t>(nVt), 0.0);
#pragma omp parallel for schedule(dynamic, 1)
        for (int m = 0; m < nmesh; m++) {
            const MeshData& d = D[static_cast<size_t>(m)];
            const double* cm = hx.data() + 2 * foff[static_cast<size_t>(m)];
            if (compressed && compressed[m]) std::memcpy(compressed[m], cm, sizeof(double) * 2 * static_cast<size_t>(d.nF));
            double* bu = du_rhs.data() + voff[static_cast<size_t>(m)];
            double* bv = dv_rhs.data() + voff[static_cast<size_t>(m)];
            for (int f = 0; f < d.nF; f++) {
                const double th = std::atan2(cm[2 * f + 1], cm[2 * f]) / N;
                const double cr = std::cos(th), ci = std::sin(th);
                double gu[3], gv[3];
                for (int k = 0; k < 3; k++) gu[k] = cr * d.bx[static_cast<size_t>(3 * f + k)] + ci * d.by[static_cast<size_t>(3 * f + k)];
                const double* n = &d.nrm[static_cast<size_t>(3 * f)];
                gv[0] = n[1] * gu[2] - n[2] * gu[1];
                gv[1] = n[2] * gu[0] - n[0] * gu[2];
                gv[2] = n[0] * gu[1] - n[1] * gu[0];
                for (int i = 0; i < 3; i++) {
                    const double* g = &d.gphi[static_cast<size_t>(9 * f + 3 * i)];
                    const int v = d.fv[static_cast<size_t>(3 * f + i)];
                    bu[v] += g[0] * gu[0] + g[1] * gu[1] + g[2] * gu[2];
                    bv[v] += g[0] * gv[0] + g[1] * gv[1] + g[2] * gv[2];
                }
            }
            bu[0] = 0.0;  // pinned vertex
            bv[0] = 0.0;
        }
        mk_op* Sop = mk::op_create_owned(c, int(nVt), std::move(Sp), std::move(Sc), std::move(Sv), osl);
        std::unique_ptr<mk_op, void (*)(mk_op*)> Shold(Sop, [](mk_op* p) { mk_op_destroy(p); });
        Sop->label = "batch_stiffness";
        mk::ensure_mg(c, Sop, o);
        mk_solve_opts os = o;
        os.tol = gtol;
        DevBuf<double> drhs, dsol(static_cast<size_t>(nVt));
        std::vector<double> hu_sol(static_cast<size_t>(nVt)), hv_sol(static_cast<size_t>(nVt));
        for (int comp = 0; comp < 2; comp++) {
            drhs.upload(comp == 0 ? du_rhs.data() : dv_rhs.data(), static_cast<size_t>(nVt), c->stream);
            hipLaunchKernelGGL(k_seg_sumsq, dim3(TV.nblk), dim3(256), 0, c->stream, (const int64_t*)TV.beg.p,
                               (const int64_t*)TV.end.p, (const double*)drhs.p, SV.part.p);
            const std::vector<double>& hb = SV.finish(1);
            std::vector<double> sc(static_cast<size_t>(nmesh)), usc(static_cast<size_t>(nmesh));
            for (int m = 0; m < nmesh; m++) {
                const double nb = std::sqrt(hb[static_cast<size_t>(m)]);
                sc[static_cast<size_t>(m)] = nb > 0.0 ? 1.0 / nb : 1.0;
                usc[static_cast<size_t>(m)] = nb > 0.0 ? nb : 1.0;
            }
            SV.scale(1, sc, drhs.p, drhs.p);
            MK_HIP(hipMemsetAsync(dsol.p, 0, static_cast<size_t>(nVt) * sizeof(double), c->stream));
            const mk::SolveResult r = mk::solve_spd(c, Sop, drhs.p, dsol.p, os, nullptr, 0.1 * gtol);
            S.poisson_iterations += r.iters;
            S.poisson_rel_residual = std::max(S.poisson_rel_residual, r.rel_residual * std::sqrt(double(nmesh)));
            if (r.status != MK_OK) throw Error(r.status, std::string("batched Poisson: ") + mk_status_str(r.status));
            SV.scale(1, usc, dsol.p, dsol.p);
            dsol.download(comp == 0 ? hu_sol.data() : hv_sol.data(), static_cast<size_t>(nVt), c->stream);
            MK_HIP(hipStreamSynchronize(c->stream));
        }
        for (int m = 0; m < nmesh; m++) {
            if (!uv || !uv[m]) continue;
            for (int v = 0; v < nV[m]; v++) {
                uv[m][2 * static_cast<size_t>(v)] = hu_sol[static_cast<size_t>(voff[static_cast<size_t>(m)] + v)];
                uv[m][2 * static_cast<size_t>(v) + 1] = hv_sol[static_cast<size_t>(voff[static_cast<size_t>(m)] + v)];
            }
        }
        S.ms_poisson = now_ms() - t0;
        S.ms_total = now_ms() - t_all;
        S.graph_replays = c->graph_replays - replays0;
        S.graph_captures = c->graph_captures - captures0;
        S.eager_batches = c->eager_batches - eager0;
        if (stats) *stats = S;
        return MK_OK;
    } catch (const Error& e) {
        c->last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        c->last_error = e.what();
        return MK_ERR_ARG;
    }
}

}  // extern "C"
