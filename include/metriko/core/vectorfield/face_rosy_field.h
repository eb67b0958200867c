// metriko (MI355X build) -- smoothest N-RoSy face field, matching, seam, combing.
//
// Drop-in for the reference's core/vectorfield/face_rosy_field.h.  The field solve
// (connection Laplacian assembly + 50-step inverse iteration) runs on the GPU
// through the C ABI; matching, singularities, seam and combing are host code with
// the reference's numbering and tie-breaking (they decide the integer outputs).
// The seam restates the libigl v2.5.0 routines the reference calls
// (cut_mesh_from_singularities with a zero matching, dijkstra with unit weights,
// adjacency_list) with linear-time equivalents that produce the same edge set.
//
// Restated from komietty/metriko include/metriko/core/vectorfield/face_rosy_field.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_FACE_ROSY_FIELD_H
#define METRIKO_FACE_ROSY_FIELD_H
#include <algorithm>
#include <deque>
#include <memory>
#include <queue>
#include "base_field.h"
#include "../solver/sparse_linear_solver.h"

namespace metriko {
    class FaceRosyField : public BaseVectorField {
    public:
        VecXd effort;
        FaceRosyField(const Hmesh& m, const int nRosy) : BaseVectorField(m, nRosy) { }
        FaceRosyField(const Hmesh& m, const int nRosy, FieldType type) : BaseVectorField(m, nRosy) {
            if (type == FieldType::Smoothest) solve_smoothest(random_complex_vector(m.nF));
            else if (type == FieldType::CurvatureAligned) solve_curvature_aligned();
        }
        // Smoothest field from an explicit start vector u0 (the reference draws it from std::rand).
        FaceRosyField(const Hmesh& m, const int nRosy, const VecXc& u0, int nIter = 50) : BaseVectorField(m, nRosy) {
            solve_smoothest(u0, nIter);
        }

        // connectionLaplacian (face_rosy_field.h:46-76), assembled on the device.
        DeviceOp connectionLaplacianDevice() const {
            auto& rt = hip::runtime();
            std::vector<int32_t> fh(size_t(mesh.nF) * 3);
            for (int f = 0; f < mesh.nF; f++) {
                int h = mesh.face2half[f];
                for (int k = 0; k < 3; k++, h = mesh.next[h]) fh[size_t(3 * f + k)] = h;
            }
            mk_op* op = nullptr;
            hip::check(mk_field_laplacian(rt.get(), mesh.nF, mesh.nH, fh.data(), mesh.twin.data(), mesh.face.data(),
                                          mesh.heArgOnFace.data(), rosyN, &rt.field_opts, &op),
                       "mk_field_laplacian");
            return DeviceOp(op);
        }
        // Host copy of the device-assembled Laplacian (CSR, sorted columns).
        SprsC connectionLaplacian() const {
            DeviceOp d = connectionLaplacianDevice();
            auto& rt = hip::runtime();
            int64_t nnz = 0;
            hip::check(mk_op_export_csr(rt.get(), d.op, &nnz, nullptr, nullptr, nullptr), "mk_op_export_csr");
            SprsC L(mesh.nF, mesh.nF);
            L.ptr.resize(size_t(mesh.nF) + 1);
            L.col.resize(size_t(nnz));
            L.val.resize(size_t(nnz));
            hip::check(mk_op_export_csr(rt.get(), d.op, &nnz, L.ptr.data(), L.col.data(),
                                        reinterpret_cast<double*>(L.val.data())), "mk_op_export_csr");
            return L;
        }
        // galerkinMassMatrix (face_rosy_field.h:78-84): diagonal of face areas.
        SprsC galerkinMassMatrix() const {
            SprsC S(mesh.nF, mesh.nF);
            std::vector<TripC> T;
            for (int f = 0; f < mesh.nF; f++) T.emplace_back(f, f, complex(mesh.faceArea[f], 0));
            S.setFromTriplets(T);
            return S;
        }

        // principalCurvatureDir (face_rosy_field.h:86-98): per face, a quarter of the sum over
        // its halfedges of -l e^{2i farg} darg (c = l e^{i farg}, -c^2 / l), in the face basis.
        VecXc principalCurvatureDir() const {
            VecXc D(mesh.nF);
            for (int f = 0; f < mesh.nF; f++) {
                complex dir(0, 0);
                int h = mesh.face2half[f];
                for (int k = 0; k < 3; k++, h = mesh.next[h]) {
                    const double l = mesh.hvec(h).norm();
                    const complex c = std::polar(1.0, mesh.heArgOnFace[h]) * l;
                    dir += -c * c / l * mesh.dihedralArg[h];
                }
                D[f] = dir * 0.25;
            }
            return D;
        }

        void computeMatching(const MatchingType type) {
            computeConnection();
            matching = VecXi::Constant(mesh.nE, -1);
            effort = VecXd::Zero(mesh.nE);
            if (type == MatchingType::Curl) cMatching();
            else pMatching();
            computeSingular();
        }

        // computeSingular (face_rosy_field.h:100-107)
        void computeSingular() {
            singular = VecXi::Ones(mesh.nV);
#pragma omp parallel for schedule(static, 4096) if (mesh.nV > 100000)
            for (int v = 0; v < mesh.nV; v++) {
                if (mesh.isBV[v]) continue;
                double sum = mesh.angleDefect[v] * rosyN;
                mesh.forVertCcw(mesh.vert2half[v], [&](int h) {
                    const bool can = mesh.edge2half[mesh.edge[h]] == h;
                    sum += (can ? -1.0 : 1.0) * effort[mesh.edge[h]];
                    return true;
                });
                singular[v] = static_cast<int>(std::round(sum / TwoPI));
            }
        }

        // computeConnection (face_rosy_field.h:109-117)
        void computeConnection() {
            connection = VecXc::Zero(mesh.nE);
#pragma omp parallel for schedule(static, 4096) if (mesh.nE > 100000)
            for (int e = 0; e < mesh.nE; e++) {
                if (mesh.edgeIsBoundary(e)) continue;
                const Row3d v = mesh.hvec(mesh.edge2half[e]).normalized();
                const int f0 = mesh.edge2face(e, 0), f1 = mesh.edge2face(e, 1);
                const complex ef(v.dot(row3(mesh.faceBasisX, f0)), v.dot(row3(mesh.faceBasisY, f0)));
                const complex eg(v.dot(row3(mesh.faceBasisX, f1)), v.dot(row3(mesh.faceBasisY, f1)));
                connection[e] = eg / ef;
            }
        }

    private:
        void solve_smoothest(const VecXc& u0, int nIter = 50) {
            DeviceOp L = connectionLaplacianDevice();
            VecXd mass(mesh.nF);
            for (int f = 0; f < mesh.nF; f++) mass[f] = mesh.faceArea[f];
            compressed = solveSmallestEig(L, mass, u0, nIter);
            roots();
        }
        // FieldType::CurvatureAligned (face_rosy_field.h:25-34) with lambda = 0: D = the principal
        // curvature direction (squared for N = 4), rhs = M D / sqrt(|D^H M D|), one Hermitian
        // positive definite solve L x = rhs on the device operator (multigrid PCG).
        void solve_curvature_aligned() {
            if (rosyN != 2 && rosyN != 4)
                throw std::invalid_argument("FaceRosyField: CurvatureAligned needs N = 2 or 4");
            VecXc D = principalCurvatureDir();
            if (rosyN == 4) for (int f = 0; f < mesh.nF; f++) D[f] = D[f] * D[f];
            double dmd = 0.0;
            for (int f = 0; f < mesh.nF; f++) dmd += mesh.faceArea[f] * std::norm(D[f]);
            const double s = std::sqrt(std::abs(dmd));
            VecXc rhs(mesh.nF);
            for (int f = 0; f < mesh.nF; f++) rhs[f] = mesh.faceArea[f] * D[f] / s;
            DeviceOp L = connectionLaplacianDevice();
            auto& rt = hip::runtime();
            compressed = VecXc::Zero(mesh.nF);
            hip::check(mk_spd_solve(rt.get(), L.op, reinterpret_cast<const double*>(rhs.data()),
                                    reinterpret_cast<double*>(compressed.data()), &rt.field_opts, &rt.last_field),
                       "mk_spd_solve");
            roots();
        }
        // face_rosy_field.h:37-43: principal N-th root times the N rotations
        void roots() {
            field.resize(mesh.nF, rosyN);
#pragma omp parallel for schedule(static, 4096) if (mesh.nF > 100000)
            for (int f = 0; f < mesh.nF; f++) {
                const double th = std::arg(compressed[f]);
                const complex base = std::polar(1.0, th / rosyN);
                for (int i = 0; i < rosyN; i++) field(f, i) = base * std::polar(1.0, TwoPI / rosyN * i);
            }
        }
        Row3d tangent(int f, const complex& c) const {
            return row3(mesh.faceBasisX, f) * c.real() + row3(mesh.faceBasisY, f) * c.imag();
        }
        // cMatching (face_rosy_field.h:169-202)
        void cMatching() {
            const int N = rosyN;
            // every edge from its own two faces: the parallel loop gives the serial result
#pragma omp parallel for schedule(static, 4096) if (mesh.nE > 100000)
            for (int e = 0; e < mesh.nE; e++) {
                if (mesh.edgeIsBoundary(e)) continue;
                const Row3d t = mesh.hvec(mesh.edge2half[e]).normalized();
                const int f0 = mesh.edge2face(e, 0), f1 = mesh.edge2face(e, 1);
                int iM = 0;
                double mn = 32767000.;
                for (int i = 0; i < N; i++) {
                    double curl = 0;
                    for (int j = 0; j < N; j++) {
                        const Row3d v0 = tangent(f0, field(f0, j)), v1 = tangent(f1, field(f1, (i + j) % N));
                        const double d = t.dot(v1 - v0);
                        curl += d * d;
                    }
                    if (curl < mn) { iM = i; mn = curl; }
                }
                matching[e] = iM;
                complex coef(1, 0);
                for (int i = 0; i < N; i++) coef *= field(f1, (i + iM) % N) / (field(f0, i) * connection[e]);
                effort[e] = std::arg(coef);
            }
        }
        // pMatching (face_rosy_field.h:138-167)
        void pMatching() {
            const int N = rosyN;
#pragma omp parallel for schedule(static, 4096) if (mesh.nE > 100000)
            for (int e = 0; e < mesh.nE; e++) {
                if (mesh.edgeIsBoundary(e)) continue;
                const int f0 = mesh.edge2face(e, 0), f1 = mesh.edge2face(e, 1);
                const complex conn = connection[e];
                double minRot = 1000.;
                int offset = 0;
                complex coef(1, 0);
                const complex t0 = conn * field(f0, 0);
                for (int i = 0; i < N; i++) {
                    coef *= field(f1, i) / (conn * field(f0, i));
                    const double r = std::arg(field(f1, i) / t0);
                    if (std::abs(r) < std::abs(minRot)) { offset = i; minRot = r; }
                }
                effort[e] = std::arg(coef);
                double res = 0;
                for (int i = 0; i < N; i++) res += std::arg(field(f1, (i + offset) % N) / (conn * field(f0, i)));
                matching[e] = offset - int(std::round((res - effort[e]) / TwoPI));
            }
        }
    };

    // ---- seam ----------------------------------------------------------------------------------
    // Face-side adjacency (libigl triangle_triangle_adjacency for manifold meshes): side s of
    // face f is the halfedge F(f,s) -> F(f,s+1), i.e. face2half, next, next.next.
    inline void face_side_adjacency(const Hmesh& m, std::vector<int>& TT, std::vector<int>& TTi, std::vector<int>& sideH) {
        TT.assign(size_t(m.nF) * 3, -1);
        TTi.assign(size_t(m.nF) * 3, -1);
        sideH.assign(size_t(m.nF) * 3, -1);
        std::vector<int> h2side(size_t(m.nH), -1);
        for (int f = 0; f < m.nF; f++) {
            int h = m.face2half[f];
            for (int s = 0; s < 3; s++, h = m.next[h]) { sideH[size_t(3 * f + s)] = h; h2side[h] = 3 * f + s; }
        }
        for (int k = 0; k < 3 * m.nF; k++) {
            const int tw = m.twin[sideH[k]];
            if (m.face[tw] < 0) continue;
            const int so = h2side[tw];
            TT[k] = so / 3;
            TTi[k] = so % 3;
        }
    }

    // libigl cut_mesh_from_singularities(V, F, MMatch = 0): BFS spanning tree of the dual
    // graph from face 0 (sides 0,1,2), then Retract.  Retract's passes over (f, s) are
    // replayed as events ordered by (pass, 3f+s): a side is dissolved at the first scan
    // position after its first vertex's cut-edge count has dropped to 1, exactly where
    // the reference's pass loop would dissolve it; passes beyond the reference's guard
    // (pass index > 10001) are not run.
    inline std::vector<uint8_t> cut_mesh_from_singularities_zero(const Hmesh& m, const std::vector<int>& TT,
                                                                 const std::vector<int>& TTi) {
        const int nF = m.nF;
        std::vector<uint8_t> seams(size_t(nF) * 3, 1), visited(size_t(nF), 0);
        std::deque<int> d;
        for (int f0 = 0; f0 < nF; f0++) {
            if (visited[f0]) continue;
            visited[f0] = 1;
            d.push_back(f0);
            while (!d.empty()) {
                const int f = d.front();
                d.pop_front();
                for (int s = 0; s < 3; s++) {
                    int g = TT[size_t(3 * f + s)], j = TTi[size_t(3 * f + s)];
                    if (j == -1) { g = f; j = s; }
                    if (!visited[g]) {
                        seams[size_t(3 * f + s)] = 0;
                        seams[size_t(3 * g + j)] = 0;
                        visited[g] = 1;
                        d.push_back(g);
                    }
                }
            }
        }
        const MatXi& F = m.idx;
        std::vector<int> e(size_t(m.nV), 0);
        for (int f = 0; f < nF; f++)
            for (int s = 0; s < 3; s++)
                if (seams[size_t(3 * f + s)] && TT[size_t(3 * f + s)] <= f) { e[F(f, s)]++; e[F(f, (s + 1) % 3)]++; }
        // sides starting at each vertex
        std::vector<int> vptr(size_t(m.nV) + 1, 0), vside(size_t(nF) * 3);
        for (int k = 0; k < 3 * nF; k++) vptr[size_t(F(k / 3, k % 3)) + 1]++;
        for (int v = 0; v < m.nV; v++) vptr[v + 1] += vptr[v];
        {
            std::vector<int> pos(vptr.begin(), vptr.end() - 1);
            for (int k = 0; k < 3 * nF; k++) vside[size_t(pos[F(k / 3, k % 3)]++)] = k;
        }
        const int64_t S = int64_t(nF) * 3;
        using Ev = int64_t;  // pass * S + position
        std::priority_queue<Ev, std::vector<Ev>, std::greater<Ev>> pq;
        for (int k = 0; k < 3 * nF; k++)
            if (seams[size_t(k)] && e[F(k / 3, k % 3)] == 1) pq.push(Ev(k));
        const int64_t maxPass = 10001;
        while (!pq.empty()) {
            const Ev ev = pq.top();
            pq.pop();
            const int64_t pass = ev / S;
            const int k = int(ev % S);
            if (pass > maxPass) break;
            const int f = k / 3, s = k % 3;
            const int a = F(f, s), b = F(f, (s + 1) % 3);
            if (!seams[size_t(k)] || e[a] != 1) continue;
            seams[size_t(k)] = 0;
            if (TT[size_t(k)] != -1) seams[size_t(3 * TT[size_t(k)] + TTi[size_t(k)])] = 0;
            e[a]--;
            e[b]--;
            for (int v : {a, b}) {
                if (e[v] != 1) continue;
                for (int q = vptr[v]; q < vptr[v + 1]; q++) {
                    const int k2 = vside[size_t(q)];
                    if (!seams[size_t(k2)]) continue;
                    pq.push(k2 > k ? pass * S + k2 : (pass + 1) * S + k2);
                }
            }
        }
        return seams;
    }

    // cut_mesh_with_singularities (face_rosy_field.h:268-343) on the mesh's own topology.
    inline std::vector<uint8_t> cut_mesh_with_singularities(const Hmesh& m, const std::vector<int>& singularities) {
        std::vector<int> TT, TTi, sideH;
        face_side_adjacency(m, TT, TTi, sideH);
        std::vector<uint8_t> cuts = cut_mesh_from_singularities_zero(m, TT, TTi);
        // igl::adjacency_list: sorted unique neighbours
        std::vector<int> aptr(size_t(m.nV) + 1, 0), adj;
        {
            // both ends of every side, bucketed per vertex, then each bucket sorted and
            // deduplicated (in parallel: buckets are independent)
            const int nV = m.nV;
            std::vector<int> cnt(size_t(nV) + 1, 0);
            for (int f = 0; f < m.nF; f++)
                for (int j = 0; j < 3; j++) { cnt[size_t(m.idx(f, j)) + 1]++; cnt[size_t(m.idx(f, (j + 1) % 3)) + 1]++; }
            for (int v = 0; v < nV; v++) cnt[size_t(v) + 1] += cnt[size_t(v)];
            std::vector<int> all(static_cast<size_t>(cnt[size_t(nV)])), pos(cnt.begin(), cnt.end() - 1);
            for (int f = 0; f < m.nF; f++)
                for (int j = 0; j < 3; j++) {
                    const int a = m.idx(f, j), d = m.idx(f, (j + 1) % 3);
                    all[size_t(pos[size_t(a)]++)] = d;
                    all[size_t(pos[size_t(d)]++)] = a;
                }
            std::vector<int> uniq(size_t(nV), 0);
#pragma omp parallel for schedule(static, 4096) if (nV > 100000)
            for (int v = 0; v < nV; v++) {
                auto b = all.begin() + cnt[size_t(v)], e = all.begin() + cnt[size_t(v) + 1];
                std::sort(b, e);
                uniq[size_t(v)] = int(std::unique(b, e) - b);
            }
            for (int v = 0; v < nV; v++) aptr[size_t(v) + 1] = aptr[size_t(v)] + uniq[size_t(v)];
            adj.resize(size_t(aptr[size_t(nV)]));
#pragma omp parallel for schedule(static, 4096) if (nV > 100000)
            for (int v = 0; v < nV; v++)
                std::copy(all.begin() + cnt[size_t(v)], all.begin() + cnt[size_t(v)] + uniq[size_t(v)],
                          adj.begin() + aptr[size_t(v)]);
        }
        std::vector<uint8_t> inCut(size_t(m.nV), 0);
        for (int k = 0; k < 3 * m.nF; k++) if (cuts[size_t(k)]) inCut[m.idx(k / 3, k % 3)] = 1;
        std::vector<int> dist(size_t(m.nV), -1), prev(size_t(m.nV), -1), touched;
        using QE = std::pair<int, int>;
        for (int src : singularities) {
            // igl::dijkstra, unit weights: pop min (distance, id); relax on strict improvement
            for (int v : touched) { dist[v] = -1; prev[v] = -1; }
            touched.clear();
            std::priority_queue<QE, std::vector<QE>, std::greater<QE>> q;
            dist[src] = 0;
            touched.push_back(src);
            q.push({0, src});
            int found = -1;
            while (!q.empty()) {
                const auto [du, u] = q.top();
                q.pop();
                if (du != dist[u]) continue;
                if (inCut[u]) { found = u; break; }
                for (int k = aptr[u]; k < aptr[u + 1]; k++) {
                    const int v = adj[size_t(k)];
                    if (dist[v] == -1 || du + 1 < dist[v]) {
                        if (dist[v] == -1) touched.push_back(v);
                        dist[v] = du + 1;
                        prev[v] = u;
                        q.push({du + 1, v});
                    }
                }
            }
            std::vector<int> path;
            if (found == -1) path.push_back(src);
            else for (int v = found; v != -1; v = prev[v]) path.push_back(v);
            for (int v : path) inCut[v] = 1;
            for (size_t i = 0; i + 1 < path.size(); i++) {
                const int v0 = path[i], v1 = path[i + 1];
                // the two faces of edge (v0, v1); the reference takes the smaller face id
                int hh = -1;
                m.forVertCcw(m.vert2half[v0], [&](int h) { if (m.head[h] == v1) { hh = h; return false; } return true; });
                if (hh < 0) throw std::runtime_error("cut_mesh_with_singularities: path edge not found");
                int fa = m.face[hh], fb = m.face[m.twin[hh]];
                const int fi = (fb >= 0 && (fa < 0 || fb < fa)) ? fb : fa;
                int j = -1;
                for (int z = 0; z < 3; z++) {
                    const int a = m.idx(fi, z), b = m.idx(fi, (z + 1) % 3);
                    if ((a == v0 && b == v1) || (a == v1 && b == v0)) j = z;
                }
                cuts[size_t(3 * fi + j)] = 1;
                const int t = TT[size_t(3 * fi + j)];
                if (t >= 0) cuts[size_t(3 * t + TTi[size_t(3 * fi + j)])] = 1;
            }
        }
        return cuts;
    }

    // compute_seam (face_rosy_field.h:345-360)
    inline std::vector<bool> compute_seam(const FaceRosyField& f) {
        std::vector<int> svids;
        for (int v = 0; v < f.mesh.nV; v++) if (f.singular[v] != 0) svids.push_back(v);
        std::vector<uint8_t> cuts = cut_mesh_with_singularities(f.mesh, svids);
        std::vector<bool> seam(size_t(f.mesh.nE), false);
        for (int k = 0; k < 3 * f.mesh.nF; k++) if (cuts[size_t(k)]) seam[f.mesh.face2edge(k / 3, k % 3)] = true;
        return seam;
    }

    // compute_combbed_field (face_rosy_field.h:207-249)
    inline std::unique_ptr<FaceRosyField> compute_combbed_field(const FaceRosyField& field, const std::vector<bool>& seam) {
        const Hmesh& mesh = field.mesh;
        const int N = field.rosyN;
        auto F = std::make_unique<FaceRosyField>(mesh, N);
        F->field.resize(mesh.nF, N);
        F->connection = field.connection;
        F->compressed = field.compressed;
        F->singular = field.singular;
        std::vector<int> turns(size_t(mesh.nF), 0);
        std::vector<uint8_t> visit(size_t(mesh.nF), 0);
        std::queue<std::pair<int, int>> q;
        q.emplace(0, 0);
        while (!q.empty()) {
            auto [i, mm] = q.front();
            q.pop();
            if (visit[i]) continue;
            visit[i] = 1;
            turns[i] = mm;
            const int r = N - mm;
            for (int k = 0; k < r; k++) F->field(i, k) = field.field(i, mm + k);
            for (int k = 0; k < mm; k++) F->field(i, r + k) = field.field(i, k);
            int h = mesh.face2half[i];
            for (int c = 0; c < 3; c++, h = mesh.next[h]) {
                const int e = mesh.edge[h];
                const bool b = mesh.edge2face(e, 0) == i;
                const int fn = b ? mesh.edge2face(e, 1) : mesh.edge2face(e, 0);
                int n = (b ? 1 : -1) * field.matching[e];
                n = (n + mm + 1000 * N) % N;
                if (!mesh.edgeIsBoundary(e) && !visit[fn] && !seam[size_t(e)]) q.emplace(fn, n);
            }
        }
        F->matching = VecXi::Constant(mesh.nE, -1);
#pragma omp parallel for schedule(static, 4096) if (mesh.nE > 100000)
        for (int e = 0; e < mesh.nE; e++) {
            if (mesh.edgeIsBoundary(e)) continue;
            F->matching[e] = (turns[mesh.edge2face(e, 0)] - turns[mesh.edge2face(e, 1)] + field.matching[e] + 1000 * N) % N;
        }
        return F;
    }

    // compute_extrinsic_field (face_rosy_field.h:251-264)
    inline MatXd compute_extrinsic_field(const Hmesh& mesh, const FaceRosyField& field, int rosyN) {
        (void)rosyN;
        MatXd ext(mesh.nF, 12);
#pragma omp parallel for schedule(static, 4096) if (mesh.nF > 100000)
        for (int f = 0; f < mesh.nF; f++)
            for (int j = 0; j < 4; j++) {
                const complex c = field.field(f, j);
                const Row3d w = (row3(mesh.faceBasisX, f) * c.real() + row3(mesh.faceBasisY, f) * c.imag()).normalized();
                for (int k = 0; k < 3; k++) ext(f, 3 * j + k) = w[k];
            }
        return ext;
    }
}

#endif
