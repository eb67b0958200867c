// metriko (MI355X build) -- halfedge mesh with per-element geometry.
//
// Drop-in for the reference's core/hmesh/hmesh.h (struct Hmesh, hmesh.h:96-147,
// constructor hmesh.h:313-569, dcel hmesh.h:240-311).  Numbering conventions
// are identical because they decide the integer outputs downstream:
//  * edges in the order of libigl's edge_topology (sides keyed (min v, max v,
//    f, i) sorted lexicographically; EF(e,0) is the face where the edge runs
//    EV(e,0) -> EV(e,1));
//  * halfedges of EF(e,0), EF(e,1) numbered in edge order (2e / 2e+1 on
//    closed meshes); boundary halfedges appended loop by loop;
//  * vert2half = the last halfedge assigned to v in edge order;
//  * face2half = side F(f,0) -> F(f,1).
// The element handles (Half, Vert, Edge, Face, ...) keep the reference API;
// the arrays are plain std::vector so the HIP side can upload them directly.
//
// Restated from komietty/metriko include/metriko/core/hmesh/hmesh.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_HMESH_H
#define METRIKO_HMESH_H
#include <algorithm>
#include "metriko/core/common/parallel.h"
#include "metriko/core/common/typedef.h"

namespace metriko {
    struct Hmesh;
    struct Half; struct Vert; struct Edge; struct Face; struct Crnr; struct Loop;

    struct Elem {
        int id;
        const Hmesh* m;
        bool operator==(const Elem& rhs) const { return id == rhs.id; }
    };
    struct Face : Elem {
        Half half() const; Row3d basisX() const; Row3d basisY() const; Row3d normal() const;
        Row3d center() const; double area() const;
    };
    struct Edge : Elem {
        Half half() const; Vert vert0() const; Vert vert1() const; Face face0() const; Face face1() const;
        double len() const; double cot() const; bool isBoundary() const;
    };
    struct Vert : Elem {
        Half half() const; Row3d pos() const; Row3d basisX() const; Row3d basisY() const; Row3d normal() const;
        bool isBoundary() const; bool isInterior() const; double baryArea() const; double circArea() const;
    };
    struct Crnr : Elem { Half half() const; Vert vert() const; Face face() const; };
    struct Loop : Elem { Half half() const; };
    struct Half : Elem {
        Half next() const; Half prev() const; Half twin() const; Vert tail() const; Vert head() const;
        Edge edge() const; Face face() const; Crnr crnr() const; double len() const; double cot() const;
        double varg() const; double farg() const; double darg() const; bool isBoundary() const;
        bool isCanonical() const; Row3d vec() const;
    };

    struct Hmesh {
        int nV = 0, nF = 0, nE = 0, nH = 0, nL = 0, nC = 0, nEularChars = 0;
        std::vector<int> vert2half, edge2half, face2half, loop2half, crnr2half;
        std::vector<int> next, prev, twin, tail, head, edge, face, crnr;
        std::vector<bool> isBV;
        std::vector<Half> halfs; std::vector<Vert> verts; std::vector<Edge> edges;
        std::vector<Face> faces; std::vector<Loop> loops; std::vector<Crnr> crnrs;
        MatXd pos; MatXi idx; MatXi edge2vert, edge2face, face2edge;
        MatXd vertBasisX, vertBasisY, vertNormal, faceBasisX, faceBasisY, faceNormal;
        VecXd faceArea, baryDualArea, circDualArea, halfCotan, edgeCotan;
        MatXd baryCenter;
        VecXd angleDefect, dihedralArg, heArgOnVert, heArgOnFace;

        Hmesh(const MatXd& V, const MatXi& F) { build(V, F); }
        Hmesh(const Hmesh&) = delete;
        Hmesh& operator=(const Hmesh&) = delete;

        bool edgeIsBoundary(int e) const { const int h = edge2half[e]; return face[h] == -1 || face[twin[h]] == -1; }
        Row3d hvec(int h) const { return row3(pos, head[h]) - row3(pos, tail[h]); }

        // clockwise / counter-clockwise outgoing halfedges of v starting at h0
        // (Vert::adjHalfs, hmesh.h:158-161): ccw h -> h.prev().twin(), cw h -> h.twin().next()
        template<typename Fn> void forVertCcw(int h0, Fn&& fn) const {
            int h = h0;
            do { if (!fn(h)) return; h = twin[prev[h]]; } while (h != h0);
        }
        template<typename Fn> void forVertCw(int h0, Fn&& fn) const {
            int h = h0;
            do { if (!fn(h)) return; h = next[twin[h]]; } while (h != h0);
        }

    private:
        void build(const MatXd& V, const MatXi& F);
        void topology();
        void geometry();
    };

    // -- element accessors (hmesh.h:184-238) -------------------------------------------
    inline Half Half::next() const { return {m->next[id], m}; }
    inline Half Half::prev() const { return {m->prev[id], m}; }
    inline Half Half::twin() const { return {m->twin[id], m}; }
    inline Vert Half::tail() const { return {m->tail[id], m}; }
    inline Vert Half::head() const { return {m->head[id], m}; }
    inline Edge Half::edge() const { return {m->edge[id], m}; }
    inline Face Half::face() const { return {m->face[id], m}; }
    inline Crnr Half::crnr() const { return {m->crnr[id], m}; }
    inline Half Vert::half() const { return {m->vert2half[id], m}; }
    inline Half Edge::half() const { return {m->edge2half[id], m}; }
    inline Half Face::half() const { return {m->face2half[id], m}; }
    inline Half Loop::half() const { return {m->loop2half[id], m}; }
    inline Half Crnr::half() const { return {m->crnr2half[id], m}; }
    inline Vert Crnr::vert() const { return half().prev().tail(); }
    inline Face Crnr::face() const { return half().face(); }
    inline Vert Edge::vert0() const { return {m->edge2vert(id, 0), m}; }
    inline Vert Edge::vert1() const { return {m->edge2vert(id, 1), m}; }
    inline Face Edge::face0() const { return {m->edge2face(id, 0), m}; }
    inline Face Edge::face1() const { return {m->edge2face(id, 1), m}; }
    inline bool Half::isCanonical() const { return m->edge2half[m->edge[id]] == id; }
    inline bool Half::isBoundary() const { return m->face[id] == -1; }
    inline bool Vert::isBoundary() const { return m->isBV[id]; }
    inline bool Vert::isInterior() const { return !m->isBV[id]; }
    inline bool Edge::isBoundary() const { return m->edgeIsBoundary(id); }
    inline double Edge::len() const { return half().len(); }
    inline double Half::len() const { return vec().norm(); }
    inline double Half::cot() const { return m->halfCotan[id]; }
    inline double Edge::cot() const { return m->edgeCotan[id]; }
    inline double Half::varg() const { return m->heArgOnVert[id]; }
    inline double Half::farg() const { return m->heArgOnFace[id]; }
    inline double Half::darg() const { return m->dihedralArg[id]; }
    inline double Face::area() const { return m->faceArea[id]; }
    inline double Vert::baryArea() const { return m->baryDualArea[id]; }
    inline double Vert::circArea() const { return m->circDualArea[id]; }
    inline Row3d Half::vec() const { return m->hvec(id); }
    inline Row3d Vert::pos() const { return row3(m->pos, id); }
    inline Row3d Vert::basisX() const { return row3(m->vertBasisX, id); }
    inline Row3d Vert::basisY() const { return row3(m->vertBasisY, id); }
    inline Row3d Vert::normal() const { return row3(m->vertNormal, id); }
    inline Row3d Face::basisX() const { return row3(m->faceBasisX, id); }
    inline Row3d Face::basisY() const { return row3(m->faceBasisY, id); }
    inline Row3d Face::normal() const { return row3(m->faceNormal, id); }
    inline Row3d Face::center() const { return row3(m->baryCenter, id); }

    // -- construction ---------------------------------------------------------------------
    inline void Hmesh::build(const MatXd& V, const MatXi& F) {
        pos = V;
        idx = F;
        nV = V.rows();
        nF = F.rows();
        // edge_topology: sides sorted by (min v, max v, f, i)
        const int64_t nS = int64_t(nF) * 3;
        std::vector<std::pair<uint64_t, int64_t>> keys(static_cast<size_t>(nS));
        for (int f = 0; f < nF; f++)
            for (int i = 0; i < 3; i++) {
                uint64_t a = uint32_t(F(f, i)), b = uint32_t(F(f, (i + 1) % 3));
                if (a > b) std::swap(a, b);
                keys[size_t(3 * int64_t(f) + i)] = {(a << 32) | b, 3 * int64_t(f) + i};
            }
        metriko::parallel_sort(keys);
        std::vector<int> ev0, ev1, ef0, ef1;
        ev0.reserve(size_t(nS / 2 + 16));
        face2edge = MatXi::Constant(nF, 3, -1);
        for (int64_t i = 0; i < nS;) {
            const uint64_t k = keys[i].first;
            const int e = int(ev0.size());
            ev0.push_back(int(k >> 32));
            ev1.push_back(int(k & 0xffffffffu));
            const int64_t s1 = keys[i].second;
            face2edge(int(s1 / 3), int(s1 % 3)) = e;
            if (i + 1 < nS && keys[i + 1].first == k) {
                if (i + 2 < nS && keys[i + 2].first == k) throw std::invalid_argument("Hmesh: mesh is not edge-manifold");
                const int64_t s2 = keys[i + 1].second;
                face2edge(int(s2 / 3), int(s2 % 3)) = e;
                ef0.push_back(int(s1 / 3));
                ef1.push_back(int(s2 / 3));
                i += 2;
            } else {
                ef0.push_back(int(s1 / 3));
                ef1.push_back(-1);
                i += 1;
            }
        }
        nE = int(ev0.size());
        edge2vert.resize(nE, 2);
        edge2face.resize(nE, 2);
        for (int e = 0; e < nE; e++) {
            edge2vert(e, 0) = ev0[e];
            edge2vert(e, 1) = ev1[e];
            int f0 = ef0[e], f1 = ef1[e];
            bool flip = true;
            for (int j = 0; j < 3; j++)
                if (F(f0, j) == ev0[e] && F(f0, (j + 1) % 3) == ev1[e]) flip = false;
            if (flip) std::swap(f0, f1);
            edge2face(e, 0) = f0;
            edge2face(e, 1) = f1;
        }
        nH = 2 * nE;
        nC = 3 * nF;
        nEularChars = nV - nE + nF;
        topology();
        geometry();
    }

    inline void Hmesh::topology() {
        const MatXi& F = idx;
        vert2half.assign(nV, -1); edge2half.assign(nE, -1); face2half.assign(nF, -1); crnr2half.assign(nC, -1);
        next.assign(nH, -1); prev.assign(nH, -1); twin.assign(nH, -1); head.assign(nH, -1);
        tail.assign(nH, -1); edge.assign(nH, -1); face.assign(nH, -1); crnr.assign(nH, -1);
        // dcel (hmesh.h:240-311)
        MatXi EH = MatXi::Constant(nE, 2, -1), FH = MatXi::Constant(nF, 3, -1);
        int numH = 0;
        for (int e = 0; e < nE; e++) {
            if (edge2face(e, 0) != -1) EH(e, 0) = numH++;
            if (edge2face(e, 1) != -1) EH(e, 1) = numH++;
        }
        std::vector<int> HV(numH), HE(numH), HF(numH), twinH(numH, -1);
        std::vector<int> VH(nV, -1);
        // each edge writes only its own halfedges and their (face, side) slots; the loops that
        // follow write disjoint entries too, so the parallel loops give the serial results
#pragma omp parallel for schedule(static, 4096) if (nE > 100000)
        for (int e = 0; e < nE; e++) {
            for (int k = 0; k < 2; k++) {
                const int h = EH(e, k);
                if (h == -1) continue;
                HE[h] = e;
                HV[h] = edge2vert(e, k);
                const int f = edge2face(e, k);
                int j = -1;
                for (int s = 0; s < 3; s++) if (face2edge(f, s) == e) j = s;
                FH(f, j) = h;
                HF[h] = f;
            }
            if (EH(e, 0) != -1 && EH(e, 1) != -1) { twinH[EH(e, 0)] = EH(e, 1); twinH[EH(e, 1)] = EH(e, 0); }
        }
        // a vertex keeps the last halfedge assigned to it in edge order: the largest h
        for (int h = 0; h < numH; h++) VH[HV[h]] = h;
#pragma omp parallel for schedule(static, 4096) if (nF > 100000)
        for (int f = 0; f < nF; f++)
            for (int j = 0; j < 3; j++) { next[FH(f, j)] = FH(f, (j + 1) % 3); prev[FH(f, (j + 1) % 3)] = FH(f, j); }
        for (int v = 0; v < nV; v++) vert2half[v] = VH[v];
        for (int e = 0; e < nE; e++) edge2half[e] = EH(e, 0);
        for (int f = 0; f < nF; f++) face2half[f] = FH(f, 0);
#pragma omp parallel for schedule(static, 4096) if (numH > 100000)
        for (int h = 0; h < numH; h++) {
            head[h] = HV[next[h]]; tail[h] = HV[h]; edge[h] = HE[h]; face[h] = HF[h];
            if (twinH[h] != -1) twin[h] = twinH[h];
        }
        // boundary loops (hmesh.h:391-416)
        for (int iH = 0, iB = nF * 3; iH < nF * 3; ++iH) {
            if (twin[iH] != -1) continue;
            int jH = iH, jB = iB;
            for (;;) {
                tail[iB] = head[jH]; head[iB] = tail[jH]; edge[iB] = edge[jH];
                prev[iB] = iB - 1; next[iB] = iB + 1;
                twin[jH] = iB; twin[iB] = jH;
                jH = prev[jH];
                bool done = false;
                while (twin[jH] != -1) { if (jH == iH) { done = true; break; } jH = prev[twin[jH]]; }
                if (done) break;
                iB++;
            }
            prev[jB] = iB; next[iB] = jB;
            loop2half.push_back(jB);
            iB++;
        }
        nL = int(loop2half.size());
        for (int e = 0; e < nE; e++) if (edge2half[e] == -1) edge2half[e] = twin[EH(e, 1)];
        for (int i = 0; i < nL; i++) loops.push_back(Loop{{i, this}});
        faces.resize(size_t(nF)); verts.resize(size_t(nV)); edges.resize(size_t(nE)); halfs.resize(size_t(nH));
        crnrs.resize(size_t(nC));
#pragma omp parallel if (nH > 100000)
        {
#pragma omp for schedule(static, 16384) nowait
            for (int i = 0; i < nF; i++) faces[size_t(i)] = Face{{i, this}};
#pragma omp for schedule(static, 16384) nowait
            for (int i = 0; i < nV; i++) verts[size_t(i)] = Vert{{i, this}};
#pragma omp for schedule(static, 16384) nowait
            for (int i = 0; i < nE; i++) edges[size_t(i)] = Edge{{i, this}};
#pragma omp for schedule(static, 16384) nowait
            for (int i = 0; i < nH; i++) halfs[size_t(i)] = Half{{i, this}};
#pragma omp for schedule(static, 16384)
            for (int i = 0; i < nC; i++) crnrs[size_t(i)] = Crnr{{i, this}};
        }
        // corners (hmesh.h:430-442)
#pragma omp parallel for schedule(static, 4096) if (nF > 100000)
        for (int f = 0; f < nF; f++)
            for (int it = 0; it < 3; it++) {
                const int vid = F(f, it), cid = f * 3 + it;
                const int hc = face2half[f];
                for (int h : {hc, next[hc], prev[hc]})
                    if (tail[h] != vid && head[h] != vid) { crnr2half[cid] = h; crnr[h] = cid; }
            }
        isBV.assign(nV, false);
        for (int e = 0; e < nE; e++)
            if (edgeIsBoundary(e)) { isBV[tail[edge2half[e]]] = true; isBV[head[edge2half[e]]] = true; }
    }

    inline void Hmesh::geometry() {
        vertNormal = MatXd::Zero(nV, 3); vertBasisX.resize(nV, 3); vertBasisY.resize(nV, 3);
        faceBasisX.resize(nF, 3); faceBasisY.resize(nF, 3); faceNormal.resize(nF, 3); baryCenter.resize(nF, 3);
        faceArea = VecXd(nF); halfCotan = VecXd(nH); edgeCotan = VecXd(nE); angleDefect = VecXd(nV);
        dihedralArg = VecXd(nH); heArgOnVert = VecXd(nH); heArgOnFace = VecXd(nH);
        baryDualArea = VecXd(nV); circDualArea = VecXd(nV);
        // every loop below computes each element from its own inputs, so the OpenMP loops give
        // the serial results bit for bit (the vertex normal sums faces in ascending order)
        // face frames and areas (hmesh.h:461-478)
#pragma omp parallel for schedule(static, 4096) if (nF > 100000)
        for (int f = 0; f < nF; f++) {
            const int h = face2half[f];
            Row3d x = hvec(h), t = hvec(prev[h]) * -1.0, n = x.cross(t);
            set_row3(faceBasisX, f, x.normalized());
            set_row3(faceBasisY, f, x.cross(n).normalized() * -1.0);
            set_row3(faceNormal, f, n.normalized());
            set_row3(baryCenter, f, (row3(pos, tail[h]) + row3(pos, tail[next[h]]) + row3(pos, tail[next[next[h]]])) / 3.0);
            faceArea[f] = n.norm() * 0.5;
        }
        // halfedge argument in the face frame (hmesh.h:481-489)
#pragma omp parallel for schedule(static, 4096) if (nF > 100000)
        for (int f = 0; f < nF; f++) {
            double sum = 0;
            int h = face2half[f];
            const Row3d bx = row3(faceBasisX, f);
            for (int k = 0; k < 3; k++, h = next[h]) {
                Row3d v1 = hvec(h).normalized(), v2 = hvec(prev[h]).normalized();
                if ((v1 - bx).norm() > 1e-10) sum += std::acos(v1.dot(v2));
                heArgOnFace[h] = sum;
            }
        }
        // vertex frames (hmesh.h:492-506): area-weighted face normals summed per vertex over its
        // faces in ascending face order (the order of the reference's loop over faces)
        int64_t fan_corners = 0;
#pragma omp parallel if (nV > 100000) reduction(+ : fan_corners)
        {
            std::vector<int> fs;
#pragma omp for schedule(static, 4096)
            for (int v = 0; v < nV; v++) {
                fs.clear();
                if (vert2half[v] >= 0)
                    forVertCcw(vert2half[v], [&](int h) {
                        if (face[h] != -1) fs.push_back(face[h]);
                        return true;
                    });
                std::sort(fs.begin(), fs.end());
                fan_corners += int64_t(fs.size());
                for (int f : fs)
                    for (int c = 0; c < 3; c++) vertNormal(v, c) += faceNormal(f, c) * faceArea[f];
            }
        }
        if (fan_corners != 3 * int64_t(nF)) {  // a vertex whose faces are not one fan: the face loop
            vertNormal = MatXd::Zero(nV, 3);
            for (int f = 0; f < nF; f++)
                for (int p = 0; p < 3; p++)
                    for (int c = 0; c < 3; c++) vertNormal(idx(f, p), c) += faceNormal(f, c) * faceArea[f];
        }
#pragma omp parallel for schedule(static, 4096) if (nV > 100000)
        for (int v = 0; v < nV; v++) {
            Row3d n = row3(vertNormal, v).normalized();
            set_row3(vertNormal, v, n);
            Row3d d = row3(pos, head[vert2half[v]]) - row3(pos, v);
            Row3d x = (d - n * d.dot(n)).normalized();
            set_row3(vertBasisX, v, x);
            set_row3(vertBasisY, v, n.cross(x));
        }
        // dihedral angles and cotangents (hmesh.h:509-533)
#pragma omp parallel for schedule(static, 4096) if (nH > 100000)
        for (int h = 0; h < nH; h++) {
            if (edgeIsBoundary(edge[h])) { dihedralArg[h] = 0; }
            else {
                Row3d n1 = row3(faceNormal, face[h]), n2 = row3(faceNormal, face[twin[h]]);
                Row3d v = hvec(h) / hvec(h).norm(), c = n1.cross(n2);
                dihedralArg[h] = std::atan2(v.dot(c), n1.dot(n2));
            }
            if (face[h] == -1) { halfCotan[h] = 0.0; continue; }
            Row3d vn = hvec(next[h]), vp = hvec(prev[h]) * -1.0;
            halfCotan[h] = vp.dot(vn) / vp.cross(vn).norm();
        }
#pragma omp parallel for schedule(static, 4096) if (nE > 100000)
        for (int e = 0; e < nE; e++) edgeCotan[e] = (halfCotan[edge2half[e]] + halfCotan[twin[edge2half[e]]]) * 0.5;
        // dual areas, vertex arguments, angle defects (hmesh.h:536-568)
#pragma omp parallel for schedule(static, 4096) if (nV > 100000)
        for (int v = 0; v < nV; v++) {
            double a1 = 0, a2 = 0, sum = 0;
            forVertCcw(vert2half[v], [&](int h) {
                if (face[h] != -1) a1 += faceArea[face[h]];
                a2 += halfCotan[h] * hvec(h).squaredNorm();
                a2 += halfCotan[prev[h]] * hvec(prev[h]).squaredNorm();
                heArgOnVert[h] = sum;
                Row3d v1 = hvec(h).normalized(), v2 = hvec(twin[prev[h]]).normalized();
                sum += std::acos(v1.dot(v2));
                return true;
            });
            baryDualArea[v] = a1 / 3.0;
            circDualArea[v] = a2 * 0.125;
            angleDefect[v] = isBV[v] ? PI - sum : TwoPI - sum;
            forVertCcw(vert2half[v], [&](int h) { heArgOnVert[h] *= TwoPI / sum; return true; });
        }
    }
}

#endif
