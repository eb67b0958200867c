// metriko (MI355X build) -- seamless Integer-Grid-Map parameterization.
//
// Drop-in for the reference's core/igm/parameterization.h (+ parameterization_setup.ipp,
// parameterization_integ.ipp).  setup() builds the transition/constraint matrices on the
// host with the reference's numbering; integ() assembles the gradient operators and
// hands the whole iterative_rounding to the GPU through the C ABI.
//
// Restated from komietty/metriko include/metriko/core/igm/parameterization.h, parameterization_setup.ipp, parameterization_integ.ipp
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_PARAMETERIZATION_H
#define METRIKO_PARAMETERIZATION_H
#include <memory>
#include <set>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "metriko/core/hmesh/hmesh.h"
#include "metriko/core/solver/matrix_ops.h"
#include "iterative_rounding/iter_rounding.h"

namespace metriko {
    class RosyParameterization {
        const Hmesh& raw;
        const Hmesh& cut;
        const MatXd& ext;
        const VecXi& singlars;
        const VecXi& matching;
        const std::vector<bool>& seam;
        double gridscale;
        int nT = 0, nR = 0, nS = 0;

    public:
        int N, n;
        SprsD vtrans2cut, constraint, uncompress;
        SprsD fullx2Nfunc;  // vtrans2cut * uncompress (a local of integ() in the reference; kept for the uv map)
        VecXi he2matching, he2transidx, integerIdcs, singularIdcs, fixedIdcs;
        VecXd fixedVals;
        MatXi lreductor;
        VecXd fullx;
        MatXd nfn, cfn;
        bool verbose, seamless, roundSeams, localInjectivity;
        // this build: integer decisions of the rounding loop (index, rounded value) in order
        std::vector<int> decisionIndex, decisionValue;
        int transitions() const { return nT; }
        int insideSingulars() const { return nS; }

        RosyParameterization(const Hmesh& raw_, const Hmesh& cut_, const MatXd& ext_, const VecXi& singlars_,
                             const VecXi& matching_, const std::vector<bool>& seam_, const int degree,
                             const double ratio = 0.02)
            : raw(raw_), cut(cut_), ext(ext_), singlars(singlars_), matching(matching_), seam(seam_), gridscale(ratio),
              verbose(false), seamless(false), roundSeams(true), localInjectivity(false) {
            N = degree;
            if (N % 2 == 0) {  // set_sign_symmetry (parameterization.h:84-90)
                n = N / 2;
                lreductor = MatXi::Zero(N, n);
                for (int i = 0; i < n; i++) { lreductor(i, i) = 1; lreductor(n + i, i) = -1; }
            } else {
                n = N;
                lreductor = MatXi::Zero(N, N);
                for (int i = 0; i < N; i++) lreductor(i, i) = 1;
            }
            fixedVals = VecXd::Zero(n);
            compute_he2transidx();
            compute_he2matching();
            nR = raw.nV + nT;
            nS = 0;
            for (int v = 0; v < raw.nV; v++) if (is_inside_singular(v)) nS++;
        }

        void setup();
        bool integ();

    private:
        bool is_inside_singular(int v) const { return !raw.isBV[v] && singlars[v] != 0; }
        int find_first_seam_he(int v) const {
            int s = raw.vert2half[v];
            raw.forVertCw(raw.vert2half[v], [&](int h) { if (seam[size_t(raw.edge[h])]) { s = h; return false; } return true; });
            return s;
        }
        int find_first_bndr_he(int v) const {
            int s = raw.vert2half[v];
            raw.forVertCw(raw.vert2half[v], [&](int h) { if (raw.edgeIsBoundary(raw.edge[h])) { s = h; return false; } return true; });
            return s;
        }
        // setup.ipp:3-9
        void compute_he2matching() {
            he2matching = VecXi(raw.nH);
            for (int h = 0; h < raw.nH; h++) {
                const bool can = raw.edge2half[raw.edge[h]] == h;
                const int m = (can ? -1 : 1) * matching[raw.edge[h]];
                he2matching[h] = m < 0 ? (N + m % N) % N : m % N;
            }
        }
        // setup.ipp:11-50
        void compute_he2transidx() {
            he2transidx = VecXi::Constant(raw.nH, 32767);
            std::vector<int> valence(size_t(raw.nV), 0);
            std::vector<uint8_t> claimed(size_t(raw.nE), 0);
            for (int e = 0; e < raw.nE; e++)
                if (seam[size_t(e)]) { valence[raw.edge2vert(e, 0)]++; valence[raw.edge2vert(e, 1)]++; }
            int tid = 1;
            for (int v = 0; v < raw.nV; v++) {
                if ((valence[v] == 2 && !is_inside_singular(v)) || valence[v] == 0) continue;
                raw.forVertCcw(find_first_bndr_he(v), [&](int cH) {
                    const int cE = raw.edge[cH];
                    if (raw.edgeIsBoundary(cE) || !seam[size_t(cE)] || claimed[cE]) return true;
                    he2transidx[cH] = tid;
                    he2transidx[raw.twin[cH]] = -tid;
                    claimed[cE] = 1;
                    int jv = raw.head[cH];
                    while (valence[jv] == 2 && !is_inside_singular(jv) && !raw.isBV[jv]) {
                        int nh = -1;
                        raw.forVertCcw(raw.vert2half[jv], [&](int h) {
                            if (seam[size_t(raw.edge[h])] && !claimed[raw.edge[h]]) { nh = h; return false; }
                            return true;
                        });
                        if (nh < 0) throw std::runtime_error("compute_he2transidx: broken seam trace");
                        he2transidx[nh] = tid;
                        he2transidx[raw.twin[nh]] = -tid;
                        claimed[raw.edge[nh]] = 1;
                        jv = raw.head[nh];
                    }
                    tid++;
                    return true;
                });
            }
            nT = tid - 1;
        }
    };

    // ---- setup (parameterization_setup.ipp:52-170) -----------------------------------------
    inline void RosyParameterization::setup() {
        using PM = std::vector<int>;  // N x N row-major integer matrix
        auto ident = [&](int s) { PM m(size_t(N * N), 0); for (int i = 0; i < N; i++) m[size_t(i * N + i)] = s; return m; };
        auto mul = [&](const PM& a, const PM& b) {
            PM c(size_t(N * N), 0);
            for (int i = 0; i < N; i++) for (int k = 0; k < N; k++) {
                const int aik = a[size_t(i * N + k)];
                if (!aik) continue;
                for (int j = 0; j < N; j++) c[size_t(i * N + j)] += aik * b[size_t(k * N + j)];
            }
            return c;
        };
        PM unit(size_t(N * N), 0);
        for (int i = 0; i < N; i++) unit[size_t(((i + 1) % N) * N + i)] = 1;
        std::vector<PM> cpm(static_cast<size_t>(N));
        cpm[0] = ident(1);
        for (int i = 1; i < N; i++) cpm[i] = mul(unit, cpm[i - 1]);
        auto assign_block = [&](std::vector<TripD>& T, const PM& m, int r0, int c0) {
            for (int i = 0; i < N; i++) for (int j = 0; j < N; j++)
                if (m[size_t(i * N + j)] != 0) T.emplace_back(r0 + i, c0 + j, double(m[size_t(i * N + j)]));
        };
        // vertices in contiguous ranges, one per thread; each range collects its own triplets
        // and numbers its constraint rows from 0, then the ranges are offset and concatenated
        // in vertex order: the serial loop's triplets, rows and numbering
        int nth = 1;
#ifdef _OPENMP
        if (raw.nV > 100000) nth = std::max(1, omp_get_max_threads());
#endif
        std::vector<std::vector<TripD>> vTs(static_cast<size_t>(nth)), cTs(static_cast<size_t>(nth));
        std::vector<int> curs(static_cast<size_t>(nth), 0);
#pragma omp parallel for schedule(static, 1) num_threads(nth) if (nth > 1)
        for (int t = 0; t < nth; t++) {
            const int v0 = int(int64_t(raw.nV) * t / nth), v1 = int(int64_t(raw.nV) * (t + 1) / nth);
            std::vector<TripD>& vTl = vTs[size_t(t)];
            std::vector<TripD>& cTl = cTs[size_t(t)];
            int cur = 0;
            std::vector<PM> permMats;
            std::vector<int> permIdcs;
            for (int v = v0; v < v1; v++) {
                permMats.assign(1, ident(1));
                permIdcs.assign(1, v);
                int iVcut = -1;
                const int start = raw.isBV[v] ? find_first_bndr_he(v) : find_first_seam_he(v);
                raw.forVertCcw(start, [&](int h) {
                    if (raw.face[h] == -1) return false;
                    const int f = raw.face[h];
                    int jVcut = -1;
                    for (int j = 0; j < 3; j++) if (raw.idx(f, j) == v) jVcut = cut.idx(f, j);
                    if (jVcut != iVcut) {
                        iVcut = jVcut;
                        for (size_t i = 0; i < permIdcs.size(); i++) assign_block(vTl, permMats[i], N * iVcut, N * permIdcs[i]);
                    }
                    const int hn = raw.twin[raw.prev[h]];
                    if (raw.face[hn] != -1 && seam[size_t(raw.edge[hn])]) {
                        const PM& p = cpm[size_t(he2matching[hn])];
                        const int tr = he2transidx[hn];
                        if (tr > 0) {
                            for (auto& m : permMats) m = mul(p, m);
                            permMats.push_back(ident(1));
                            permIdcs.push_back(raw.nV + tr - 1);
                        } else {
                            permMats.push_back(ident(-1));
                            permIdcs.push_back(raw.nV - tr - 1);
                            for (auto& m : permMats) m = mul(p, m);
                        }
                    }
                    return true;
                });
                if (!raw.isBV[v]) {
                    std::vector<int> idcs(permIdcs);
                    std::sort(idcs.begin(), idcs.end());
                    idcs.erase(std::unique(idcs.begin(), idcs.end()), idcs.end());
                    std::vector<PM> mats(idcs.size(), PM(size_t(N * N), 0));
                    bool any = false;
                    for (size_t j = 0; j < idcs.size(); j++) {
                        for (size_t k = 0; k < permIdcs.size(); k++)
                            if (idcs[j] == permIdcs[k]) for (int q = 0; q < N * N; q++) mats[j][size_t(q)] += permMats[k][size_t(q)];
                        if (idcs[j] == v) for (int i = 0; i < N; i++) mats[j][size_t(i * N + i)] -= 1;
                        for (int q = 0; q < N * N; q++) if (mats[j][size_t(q)] != 0) any = true;
                    }
                    if (any) {
                        for (size_t j = 0; j < mats.size(); j++) assign_block(cTl, mats[j], N * cur, N * idcs[j]);
                        cur++;
                    }
                }
            }
            curs[size_t(t)] = cur;
        }
        std::vector<TripD> vT, cT;
        int cur = 0;
        for (int t = 0; t < nth; t++) {
            vT.insert(vT.end(), vTs[size_t(t)].begin(), vTs[size_t(t)].end());
            for (const TripD& q : cTs[size_t(t)]) cT.emplace_back(q.r + N * cur, q.c, q.v);
            cur += curs[size_t(t)];
        }
        vtrans2cut.resize(N * cut.nV, N * nR);
        vtrans2cut.setFromTriplets(vT);
        vtrans2cut.prune(1e-3);
        constraint.resize(N * cur, N * nR);
        constraint.setFromTriplets(cT);
        constraint.prune(1e-3);
        std::vector<TripD> buff;
        for (int i = 0; i < N * nR; i += N)
            for (int a = 0; a < N; a++) for (int b = 0; b < n; b++)
                if (lreductor(a, b) != 0) buff.emplace_back(i + a, i * n / N + b, double(lreductor(a, b)));
        uncompress.resize(N * nR, n * nR);
        uncompress.setFromTriplets(buff);
        fixedIdcs = VecXi(n);
        if (nS == 0) {
            for (int j = 0; j < n; j++) fixedIdcs[j] = j;
        } else {
            int iV = 0;
            for (iV = 0; iV < raw.nV; iV++) if (is_inside_singular(iV)) break;
            for (int j = 0; j < n; j++) fixedIdcs[j] = n * iV + j;
        }
        integerIdcs = VecXi(nT * n);
        for (int i = 0; i < nT; i++) for (int j = 0; j < n; j++) integerIdcs[n * i + j] = n * (raw.nV + i) + j;
        singularIdcs = VecXi(n * nS);
        int c = 0;
        for (int v = 0; v < raw.nV; v++)
            if (is_inside_singular(v)) for (int j = 0; j < n; j++) singularIdcs[c++] = n * v + j;
    }

    // ---- integ (parameterization_integ.ipp:3-120) ------------------------------------------
    inline bool RosyParameterization::integ() {
        fullx2Nfunc = vtrans2cut * uncompress;
        SprsD Cfull = constraint * uncompress;
        reduce_to_linearly_independent(Cfull);
        // G2: per cut face and corner, gradient of the hat function in the face frame.  Rows
        // 2N f + 2k + {0, 1} hold one entry per corner of face f (column N tail + k); written
        // straight into CSR, each row's three entries in ascending column order (what the
        // triplet assembly gives: a face's corners are distinct cut vertices), faces in parallel
        SprsD G2(2 * N * cut.nF, N * cut.nV);
        {
            const int nrows = 2 * N * cut.nF;
#pragma omp parallel for schedule(static, 16384) if (nrows > 100000)
            for (int r = 0; r <= nrows; r++) G2.ptr[size_t(r)] = 3 * int64_t(r);
            G2.col.resize(size_t(3) * size_t(nrows));
            G2.val.resize(size_t(3) * size_t(nrows));
#pragma omp parallel for schedule(static, 4096) if (cut.nF > 100000)
            for (int f = 0; f < cut.nF; f++) {
                const Row3d nrm = row3(cut.faceNormal, f), bx = row3(cut.faceBasisX, f), by = row3(cut.faceBasisY, f);
                int h = cut.face2half[f];
                int tl[3];
                double gx[3], gy[3];
                for (int k3 = 0; k3 < 3; k3++, h = cut.next[h]) {
                    const Row3d g = nrm.cross(cut.hvec(cut.next[h])) / (cut.faceArea[f] * 2);
                    tl[k3] = cut.tail[h];
                    gx[k3] = g.dot(bx);
                    gy[k3] = g.dot(by);
                }
                int o[3] = {0, 1, 2};  // corners by ascending cut vertex
                std::stable_sort(o, o + 3, [&](int a, int b) { return tl[a] < tl[b]; });
                for (int k = 0; k < N; k++) {
                    const int64_t r0 = 2 * int64_t(N) * f + 2 * k;
                    for (int q = 0; q < 3; q++) {
                        const int c = N * tl[o[q]] + k;
                        G2.col[size_t(3 * r0 + q)] = c;
                        G2.val[size_t(3 * r0 + q)] = gx[o[q]];
                        G2.col[size_t(3 * (r0 + 1) + q)] = c;
                        G2.val[size_t(3 * (r0 + 1) + q)] = gy[o[q]];
                    }
                }
            }
        }
        // intrinsic field (integ.ipp:30-43)
        double norm = 0.0;
        for (int i = 0; i < cut.nF; i++)
            for (int j = 0; j < N; j++) {
                double s = 0;
                for (int k = 0; k < 3; k++) s += ext(i, 3 * j + k) * ext(i, 3 * j + k);
                norm += std::sqrt(s);
            }
        norm /= double(N) * cut.nF;
        VecXd F2(2 * N * cut.nF);
#pragma omp parallel for schedule(static, 4096) if (cut.nF > 100000)
        for (int f = 0; f < cut.nF; f++)
            for (int i = 0; i < N; i++) {
                double a = 0, b = 0;
                for (int k = 0; k < 3; k++) { a += ext(f, 3 * i + k) * cut.faceBasisX(f, k); b += ext(f, 3 * i + k) * cut.faceBasisY(f, k); }
                F2[2 * N * f + 2 * i] = a;
                F2[2 * N * f + 2 * i + 1] = b;
            }
        double bb[3];
        for (int k = 0; k < 3; k++) {
            double mn = cut.pos(0, k), mx = cut.pos(0, k);
            for (int v = 1; v < cut.nV; v++) { mn = std::min(mn, cut.pos(v, k)); mx = std::max(mx, cut.pos(v, k)); }
            bb[k] = mx - mn;
        }
        const double length = std::sqrt(bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2]) * gridscale / norm;
        SprsD G2full = G2 * fullx2Nfunc;
        // the naive pass's iter_cb (parameterization_integ.ipp:45-63), registered for this call
        // only (removed again on every exit path)
        struct CheckScope {
            int N = 0;
            ~CheckScope() { if (N) mk_igm_injectivity_check(hip::runtime().get(), N, 0, nullptr, 0, 0, nullptr, nullptr, nullptr); }
        } check_scope;
        if (localInjectivity) {
            std::vector<int32_t> ci(size_t(3) * cut.nF);
            for (int f = 0; f < cut.nF; f++)
                for (int j = 0; j < 3; j++) ci[size_t(3) * f + j] = cut.idx(f, j);
            hip::check(mk_igm_injectivity_check(hip::runtime().get(), N, cut.nF, ci.data(), fullx2Nfunc.rows(),
                                                fullx2Nfunc.cols(), fullx2Nfunc.ptr.data(), fullx2Nfunc.col.data(),
                                                fullx2Nfunc.val.data()),
                       "mk_igm_injectivity_check");
            check_scope.N = N;
        }
        bool success = iterative_rounding(fixedIdcs, fixedVals, singularIdcs, integerIdcs, length, Cfull, G2full, N, n,
                                          cut.nF, seamless, roundSeams, localInjectivity, verbose, F2, fullx,
                                          &decisionIndex, &decisionValue);
        nfn.resize(cut.nV, N);
        cfn.resize(raw.nF, N * 3);
        VecXd NF = fullx2Nfunc * fullx;
#pragma omp parallel for schedule(static, 4096) if (cut.nV > 100000)
        for (int i = 0; i < cut.nV; i++) for (int k = 0; k < N; k++) nfn(i, k) = NF[N * i + k];
#pragma omp parallel for schedule(static, 4096) if (raw.nF > 100000)
        for (int i = 0; i < raw.nF; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < N; k++) cfn(i, N * j + k) = nfn(cut.idx(i, j), k);
        return success;
    }

    // ---- compute_cut_mesh (parameterization.h:126-152) -------------------------------------
    inline std::unique_ptr<Hmesh> compute_cut_mesh(const Hmesh& m, const std::vector<bool>& seam) {
        MatXi cutF(m.nF, 3);
        // vertex v opens a new copy at its fan's start and at every seam crossing; the copies
        // are numbered in vertex order, so each vertex's count is taken first (in parallel),
        // then the fans are walked again with their first number
        const int nV = m.nV;
        auto start = [&](int v) {
            int bgn = m.vert2half[v];
            if (m.isBV[v]) m.forVertCw(m.vert2half[v], [&](int h) { if (m.edgeIsBoundary(m.edge[h])) { bgn = h; return false; } return true; });
            else m.forVertCw(m.vert2half[v], [&](int h) { if (seam[size_t(m.edge[h])]) { bgn = h; return false; } return true; });
            return bgn;
        };
        std::vector<int> first(size_t(nV) + 1, 0);
#pragma omp parallel for schedule(static, 4096) if (nV > 100000)
        for (int v = 0; v < nV; v++) {
            if (m.vert2half[v] < 0) continue;
            const int bgn = start(v);
            int c = 0;
            m.forVertCcw(bgn, [&](int h) {
                if (m.face[h] != -1 && (seam[size_t(m.edge[h])] || h == bgn)) c++;
                return true;
            });
            first[size_t(v) + 1] = c;
        }
        for (int v = 0; v < nV; v++) first[size_t(v) + 1] += first[size_t(v)];
        std::vector<int> cut2v(static_cast<size_t>(first[size_t(nV)]));
#pragma omp parallel for schedule(static, 4096) if (nV > 100000)
        for (int v = 0; v < nV; v++) {
            if (m.vert2half[v] < 0) continue;
            const int bgn = start(v);
            int id = first[size_t(v)] - 1;
            m.forVertCcw(bgn, [&](int h) {
                if (m.face[h] == -1) return true;
                if (seam[size_t(m.edge[h])] || h == bgn) cut2v[size_t(++id)] = v;
                const int f = m.face[h];
                for (int j = 0; j < 3; j++) if (m.idx(f, j) == v) cutF(f, j) = id;
                return true;
            });
        }
        MatXd cutV(int(cut2v.size()), 3);
        for (size_t i = 0; i < cut2v.size(); i++) for (int k = 0; k < 3; k++) cutV(int(i), k) = m.pos(cut2v[i], k);
        return std::make_unique<Hmesh>(cutV, cutF);
    }
}

#endif
