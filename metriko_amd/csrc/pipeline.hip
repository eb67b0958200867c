// C ABI drivers built on the drop-in headers: the example_igm pipeline, a host-only
// combinatorics stage (runs without a GPU) and the Hmesh export used by the tests.
#include <chrono>
#include <cstring>
#include "mk_internal.h"
#include "igm_state.h"
#include "metriko/core/vectorfield/face_rosy_field.h"
#include "metriko/core/igm/parameterization.h"
#include "metriko/core/common/io.h"
#include "metriko/core/qex/sanitization.h"
#include "metriko/core/tmesh/motorcycle.h"
#include "metriko/core/tmesh/tmesh.h"
#include "metriko/core/quantization/quantization.h"
#include "metriko/core/qex/gen_q_vert.h"
#include "metriko/core/qex/gen_q_port.h"
#include "metriko/core/qex/gen_q_edge.h"
#include "metriko/core/qex/gen_q_face.h"

namespace {
double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
metriko::MatXd toV(int nV, const double* V) {
    metriko::MatXd M(nV, 3);
    std::memcpy(M.data(), V, sizeof(double) * size_t(nV) * 3);
    return M;
}
metriko::MatXi toF(int nV, int nF, const int32_t* F) {
    metriko::MatXi M(nF, 3);
    for (int64_t i = 0; i < int64_t(nF) * 3; i++) {
        if (F[i] < 0 || F[i] >= nV) throw mk::Error(MK_ERR_ARG, "face index out of range");
        M.d[size_t(i)] = F[i];
    }
    return M;
}
template <typename Fn>
mk_status guarded(mk_ctx* c, Fn&& fn) {
    try {
        fn();
        return MK_OK;
    } catch (const mk::Error& e) {
        if (c) c->last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (c) c->last_error = e.what();
        return MK_ERR_ARG;
    }
}

struct HostStages {
    std::unique_ptr<metriko::FaceRosyField> combed;
    std::vector<bool> seam;
    std::unique_ptr<metriko::Hmesh> cut;
    metriko::MatXd ext;
    double ms_matching = 0, ms_seam = 0, ms_cut = 0;
};

// matching -> singularities -> seam -> cut mesh -> combed field -> extrinsic field
void host_stages(const metriko::Hmesh& mesh, metriko::FaceRosyField& rawf, HostStages& H) {
    using namespace metriko;
    double t = now_ms();
    rawf.computeMatching(MatchingType::Curl);
    H.ms_matching = now_ms() - t;
    t = now_ms();
    H.seam = compute_seam(rawf);
    H.ms_seam = now_ms() - t;
    t = now_ms();
    H.cut = compute_cut_mesh(mesh, H.seam);
    H.combed = compute_combbed_field(rawf, H.seam);
    H.ext = compute_extrinsic_field(mesh, *H.combed, rawf.rosyN);
    H.ms_cut = now_ms() - t;
}
}  // namespace

extern "C" {

mk_status mk_hmesh_export(int nV, int nF, const double* V, const int32_t* F, int* nE, int32_t* edge2vert,
                          int32_t* edge2face, int32_t* face2edge, int32_t* vert2half, double* farg,
                          double* angle_defect) {
    if (nV <= 0 || nF <= 0 || !V || !F || !nE) return MK_ERR_ARG;
    return guarded(nullptr, [&] {
        metriko::Hmesh m(toV(nV, V), toF(nV, nF, F));
        *nE = m.nE;
        if (edge2vert) std::memcpy(edge2vert, m.edge2vert.data(), sizeof(int) * size_t(m.nE) * 2);
        if (edge2face) std::memcpy(edge2face, m.edge2face.data(), sizeof(int) * size_t(m.nE) * 2);
        if (face2edge) std::memcpy(face2edge, m.face2edge.data(), sizeof(int) * size_t(m.nF) * 3);
        if (vert2half) std::memcpy(vert2half, m.vert2half.data(), sizeof(int) * size_t(m.nV));
        if (farg) std::memcpy(farg, m.heArgOnFace.data(), sizeof(double) * size_t(m.nH));
        if (angle_defect) std::memcpy(angle_defect, m.angleDefect.data(), sizeof(double) * size_t(m.nV));
    });
}

mk_status mk_hmesh_halfedges(int nV, int nF, const double* V, const int32_t* F, int32_t* face_halfs,
                             int32_t* twin, int32_t* half_face, double* farg, double* face_area) {
    if (nV <= 0 || nF <= 0 || !V || !F) return MK_ERR_ARG;
    return guarded(nullptr, [&] {
        metriko::Hmesh m(toV(nV, V), toF(nV, nF, F));
        for (int f = 0; f < m.nF; f++) {
            int h = m.face2half[f];
            for (int k = 0; k < 3; k++, h = m.next[h]) if (face_halfs) face_halfs[3 * f + k] = h;
        }
        if (twin) std::memcpy(twin, m.twin.data(), sizeof(int) * size_t(m.nH));
        if (half_face) std::memcpy(half_face, m.face.data(), sizeof(int) * size_t(m.nH));
        if (farg) std::memcpy(farg, m.heArgOnFace.data(), sizeof(double) * size_t(m.nH));
        if (face_area) std::memcpy(face_area, m.faceArea.data(), sizeof(double) * size_t(m.nF));
    });
}

// Host-only stages downstream of the field solve (no GPU): given the compressed field,
// compute roots, curl matching, singularities, seam, cut mesh, combed matching and the
// IGM setup sizes.  Outputs: singular[nV], matching[nE] (combed), seam[nE], cut_idx[3nF],
// sizes[4] = {cutV, nS, nT, constraint rows}, ext[12 nF].
mk_status mk_combinatorics(int nV, int nF, const double* V, const int32_t* F, int N, const double* compressed,
                           int32_t* singular, int32_t* raw_matching, int32_t* matching, uint8_t* seam,
                           int32_t* cut_idx, int32_t* sizes, double* ext) {
    if (nV <= 0 || nF <= 0 || !V || !F || !compressed || N != 4) return MK_ERR_ARG;
    return guarded(nullptr, [&] {
        using namespace metriko;
        Hmesh mesh(toV(nV, V), toF(nV, nF, F));
        FaceRosyField rawf(mesh, N);
        rawf.compressed = VecXc(nF);
        std::memcpy(rawf.compressed.data(), compressed, sizeof(double) * 2 * size_t(nF));
        rawf.field.resize(nF, N);
        for (int f = 0; f < nF; f++) {
            const double th = std::arg(rawf.compressed[f]);
            const complex base = std::polar(1.0, th / N);
            for (int i = 0; i < N; i++) rawf.field(f, i) = base * std::polar(1.0, TwoPI / N * i);
        }
        HostStages H;
        host_stages(mesh, rawf, H);
        RosyParameterization rp(mesh, *H.cut, H.ext, rawf.singular, H.combed->matching, H.seam, N, 0.02);
        rp.seamless = true;
        rp.roundSeams = false;
        rp.setup();
        for (int v = 0; v < nV; v++) if (singular) singular[v] = rawf.singular[v];
        for (int e = 0; e < mesh.nE; e++) {
            if (raw_matching) raw_matching[e] = rawf.matching[e];
            if (matching) matching[e] = H.combed->matching[e];
            if (seam) seam[e] = H.seam[size_t(e)] ? 1 : 0;
        }
        if (cut_idx) std::memcpy(cut_idx, H.cut->idx.data(), sizeof(int) * size_t(nF) * 3);
        if (ext) std::memcpy(ext, H.ext.data(), sizeof(double) * size_t(nF) * 12);
        if (sizes) {
            sizes[0] = H.cut->nV;
            sizes[1] = rp.insideSingulars();
            sizes[2] = rp.transitions();
            sizes[3] = rp.constraint.rows();
        }
    });
}

mk_status mk_curvature_dir(int nV, int nF, const double* V, const int32_t* F, double* D) {
    if (nV <= 0 || nF <= 0 || !V || !F || !D) return MK_ERR_ARG;
    return guarded(nullptr, [&] {
        using namespace metriko;
        Hmesh mesh(toV(nV, V), toF(nV, nF, F));
        FaceRosyField f(mesh, 4);
        VecXc d = f.principalCurvatureDir();
        std::memcpy(D, d.data(), sizeof(double) * 2 * size_t(nF));
    });
}

mk_status mk_field_solve(mk_ctx* c, int nV, int nF, const double* V, const int32_t* F, int N, int field_type,
                         const double* u0, int n_iter, const mk_solve_opts* opts, double* compressed,
                         mk_stats* stats) {
    if (!c || nV <= 0 || nF <= 0 || !V || !F || N <= 0 || !compressed) return MK_ERR_ARG;
    if (field_type != MK_FIELD_SMOOTHEST && field_type != MK_FIELD_CURVATURE_ALIGNED) return MK_ERR_ARG;
    if (field_type == MK_FIELD_CURVATURE_ALIGNED && N != 2 && N != 4) return MK_ERR_ARG;
    return guarded(c, [&] {
        using namespace metriko;
        auto& rt = hip::runtime();
        struct Detach {
            hip::Runtime& r;
            ~Detach() { r.ctx = nullptr; r.external = false; }
        } detach{rt};
        rt.attach(c);
        if (opts) rt.field_opts = *opts; else mk_default_opts(&rt.field_opts);
        Hmesh mesh(toV(nV, V), toF(nV, nF, F));
        std::unique_ptr<FaceRosyField> f;
        if (field_type == MK_FIELD_CURVATURE_ALIGNED) {
            f = std::make_unique<FaceRosyField>(mesh, N, FieldType::CurvatureAligned);
        } else {
            VecXc u(nF);
            if (u0) std::memcpy(u.data(), u0, sizeof(double) * 2 * size_t(nF));
            else u = random_complex_vector(nF);
            f = std::make_unique<FaceRosyField>(mesh, N, u, n_iter);
        }
        std::memcpy(compressed, f->compressed.data(), sizeof(double) * 2 * size_t(nF));
        if (stats) *stats = rt.last_field;
    });
}

mk_status mk_field_matching(int nV, int nF, const double* V, const int32_t* F, int N, const double* compressed,
                            int matching_type, int32_t* singular, int32_t* matching, double* effort) {
    if (nV <= 0 || nF <= 0 || !V || !F || N <= 0 || !compressed) return MK_ERR_ARG;
    if (matching_type != MK_MATCH_PRINCIPAL && matching_type != MK_MATCH_CURL) return MK_ERR_ARG;
    return guarded(nullptr, [&] {
        using namespace metriko;
        Hmesh mesh(toV(nV, V), toF(nV, nF, F));
        FaceRosyField f(mesh, N);
        f.compressed = VecXc(nF);
        std::memcpy(f.compressed.data(), compressed, sizeof(double) * 2 * size_t(nF));
        f.field.resize(nF, N);
        for (int i = 0; i < nF; i++) {
            const complex base = std::polar(1.0, std::arg(f.compressed[i]) / N);
            for (int j = 0; j < N; j++) f.field(i, j) = base * std::polar(1.0, TwoPI / N * j);
        }
        f.computeMatching(matching_type == MK_MATCH_CURL ? MatchingType::Curl : MatchingType::Principal);
        for (int v = 0; v < nV; v++) if (singular) singular[v] = f.singular[v];
        for (int e = 0; e < mesh.nE; e++) {
            if (matching) matching[e] = f.matching[e];
            if (effort) effort[e] = f.effort[e];
        }
    });
}

mk_status mk_pipeline_run(mk_ctx* c, int nV, int nF, const double* V, const int32_t* F, int N, double ratio,
                          const double* u0, int n_iter, const mk_solve_opts* field_opts,
                          const mk_solve_opts* igm_opts, double* compressed, int32_t* singular, int32_t* matching,
                          uint8_t* seam, int32_t* cut_idx, double* uv, int fullx_cap, double* fullx, int dec_cap,
                          int32_t* dec_index, int32_t* dec_value, mk_pipeline_info* info) {
    if (!c || nV <= 0 || nF <= 0 || !V || !F || N != 4 || !info) return MK_ERR_ARG;
    return guarded(c, [&] {
        using namespace metriko;
        auto& rt = hip::runtime();
        struct Detach {
            hip::Runtime& r;
            ~Detach() { r.ctx = nullptr; r.external = false; }
        } detach{rt};
        rt.attach(c);
        if (field_opts) rt.field_opts = *field_opts; else mk_default_opts(&rt.field_opts);
        if (igm_opts) rt.igm_opts = *igm_opts; else mk_default_opts(&rt.igm_opts);
        std::memset(info, 0, sizeof(*info));
        const double t_all = now_ms();
        double t = now_ms();
        Hmesh mesh(toV(nV, V), toF(nV, nF, F));
        info->ms_hmesh = now_ms() - t;
        info->nV = mesh.nV;
        info->nF = mesh.nF;
        info->nE = mesh.nE;
        VecXc u(nF);
        if (u0) std::memcpy(u.data(), u0, sizeof(double) * 2 * size_t(nF));
        else u = random_complex_vector(nF);
        t = now_ms();
        FaceRosyField rawf(mesh, N);
        DeviceOp L = rawf.connectionLaplacianDevice();
        MK_HIP(hipStreamSynchronize(c->stream));
        info->ms_field_asm = now_ms() - t;
        t = now_ms();
        VecXd mass(nF);
        for (int f = 0; f < nF; f++) mass[f] = mesh.faceArea[f];
        rawf.compressed = solveSmallestEig(L, mass, u, n_iter);
        info->field = rt.last_field;
        info->ms_field_solve = now_ms() - t;
        MK_LOG("[pipeline] field solve %.0f ms\n", info->ms_field_solve);
        rawf.field.resize(nF, N);
        for (int f = 0; f < nF; f++) {
            const double th = std::arg(rawf.compressed[f]);
            const complex base = std::polar(1.0, th / N);
            for (int i = 0; i < N; i++) rawf.field(f, i) = base * std::polar(1.0, TwoPI / N * i);
        }
        HostStages H;
        host_stages(mesh, rawf, H);
        info->ms_matching = H.ms_matching;
        info->ms_seam = H.ms_seam;
        info->ms_cut = H.ms_cut;
        MK_LOG("[pipeline] matching/seam/cut %.0f/%.0f/%.0f ms\n", H.ms_matching, H.ms_seam, H.ms_cut);
        t = now_ms();
        RosyParameterization rp(mesh, *H.cut, H.ext, rawf.singular, H.combed->matching, H.seam, N, ratio);
        rp.seamless = true;
        rp.localInjectivity = false;
        rp.roundSeams = false;
        rp.verbose = false;
        rp.setup();
        info->ms_setup = now_ms() - t;
        MK_LOG("[pipeline] parameterization setup %.0f ms\n", info->ms_setup);
        t = now_ms();
        const bool ok = rp.integ();
        info->ms_integ = now_ms() - t;
        if (c->keep_igm_state && c->igm_state && !c->igm_state->has_P) {
            // corner map: uv of face i, corner j = rows N cut.idx(i, j) + {0, 1} of fullx2Nfunc
            mk_igm_state* st = c->igm_state;
            metriko::SprsD& P = st->P;
            P = metriko::SprsD(6 * nF, rp.fullx2Nfunc.cols());
            P.ptr.assign(size_t(6 * nF) + 1, 0);
            for (int i = 0; i < nF; i++)
                for (int j = 0; j < 3; j++)
                    for (int kk = 0; kk < 2; kk++) {
                        const int src = N * H.cut->idx(i, j) + kk, dst = 6 * i + 2 * j + kk;
                        for (int64_t q = rp.fullx2Nfunc.ptr[src]; q < rp.fullx2Nfunc.ptr[src + 1]; q++) {
                            P.col.push_back(rp.fullx2Nfunc.col[q]);
                            P.val.push_back(rp.fullx2Nfunc.val[q]);
                        }
                        P.ptr[size_t(dst) + 1] = int64_t(P.col.size());
                    }
            st->has_P = true;
        }
        info->igm = rt.last_igm;
        info->success = ok ? 1 : 0;
        info->cutV = H.cut->nV;
        info->nS = rp.insideSingulars();
        info->nT = rp.transitions();
        info->k = rt.last_igm.k;
        info->nfullx = rp.fullx.size();
        info->ndec = int(rp.decisionIndex.size());
        if (compressed) std::memcpy(compressed, rawf.compressed.data(), sizeof(double) * 2 * size_t(nF));
        for (int v = 0; v < nV; v++) if (singular) singular[v] = rawf.singular[v];
        for (int e = 0; e < mesh.nE; e++) {
            if (matching) matching[e] = H.combed->matching[e];
            if (seam) seam[e] = H.seam[size_t(e)] ? 1 : 0;
        }
        if (cut_idx) std::memcpy(cut_idx, H.cut->idx.data(), sizeof(int) * size_t(nF) * 3);
        if (uv)
            for (int f = 0; f < nF; f++)
                for (int j = 0; j < 3; j++) {
                    uv[size_t(6 * f + 2 * j)] = rp.cfn(f, N * j);
                    uv[size_t(6 * f + 2 * j + 1)] = rp.cfn(f, N * j + 1);
                }
        if (fullx) for (int i = 0; i < std::min(fullx_cap, rp.fullx.size()); i++) fullx[i] = rp.fullx[i];
        for (int i = 0; i < std::min(dec_cap, info->ndec); i++) {
            if (dec_index) dec_index[i] = rp.decisionIndex[size_t(i)];
            if (dec_value) dec_value[i] = rp.decisionValue[size_t(i)];
        }
        info->ms_total = now_ms() - t_all;
    });
}

}  // extern "C"

extern "C" mk_status mk_qex_extract(mk_ctx* ctx, int nV, int nF, const double* V, const int32_t* F, int N,
                                    const int32_t* matching, const int32_t* singular, const double* uv,
                                    double* uv_out, int cap_q, double* q_uv, double* q_pos, int32_t* q_sid,
                                    int cap_p, int32_t* p_ids, double* p_uvdir, int cap_e, int32_t* e_ports,
                                    int cap_f, int32_t* f_halfs, mk_qex_counts* counts) {
    if (nV <= 0 || nF <= 0 || !V || !F || !matching || !singular || !uv || !counts || N != 4) return MK_ERR_ARG;
    return guarded(ctx, [&] {
        using namespace metriko;
        Hmesh mesh(toV(nV, V), toF(nV, nF, F));
        VecXi mt(mesh.nE), sg(nV);
        for (int e = 0; e < mesh.nE; e++) mt[e] = matching[e];
        for (int v = 0; v < nV; v++) sg[v] = singular[v];
        VecXc cfn(3 * nF);
        for (int c = 0; c < 3 * nF; c++) cfn[c] = complex(uv[2 * c], uv[2 * c + 1]);
        // example_qex.cpp:84-140: sanitize, q-verts, ports (vertex, edge, face), edges, faces
        qex::sanitization(mesh, mt, sg, N, cfn);
        std::vector<qex::Qvert> vq, eq, fq;
        qex::generate_q_vert(mesh, cfn, vq, eq, fq);
        std::vector<qex::Qport> ports;
        qex::generate_vqvert_qport(mesh, cfn, vq, ports);
        qex::generate_eqvert_qport(mesh, cfn, eq, ports);
        qex::generate_fqvert_qport(mesh, fq, ports);
        const std::vector<qex::Qedge> edges = qex::generate_q_edge(mesh, cfn, mt, ports);
        const std::vector<qex::Qface> faces = qex::generate_q_faces(ports, edges);
        *counts = mk_qex_counts{int(vq.size()), int(eq.size()), int(fq.size()), int(ports.size()), int(edges.size()),
                                int(faces.size())};
        if (uv_out)
            for (int c = 0; c < 3 * nF; c++) { uv_out[2 * c] = cfn[c].real(); uv_out[2 * c + 1] = cfn[c].imag(); }
        int k = 0;
        for (const auto* L : {&vq, &eq, &fq})
            for (const qex::Qvert& q : *L) {
                if (k < cap_q) {
                    if (q_uv) { q_uv[2 * k] = q.uv.real(); q_uv[2 * k + 1] = q.uv.imag(); }
                    if (q_pos) for (int j = 0; j < 3; j++) q_pos[3 * k + j] = q.pos[j];
                    if (q_sid) q_sid[k] = q.sid;
                }
                k++;
            }
        for (int i = 0; i < int(ports.size()) && i < cap_p; i++) {
            const qex::Qport& p = ports[size_t(i)];
            if (p_ids) {
                int32_t* o = p_ids + 6 * i;
                o[0] = p.vid; o[1] = p.eid; o[2] = p.fid; o[3] = p.prev_id; o[4] = p.next_id; o[5] = p.isConnected;
            }
            if (p_uvdir) {
                double* o = p_uvdir + 4 * i;
                o[0] = p.uv.real(); o[1] = p.uv.imag(); o[2] = p.dir.real(); o[3] = p.dir.imag();
            }
        }
        for (int i = 0; i < int(edges.size()) && i < cap_e; i++)
            if (e_ports) { e_ports[2 * i] = edges[size_t(i)].port1.idx; e_ports[2 * i + 1] = edges[size_t(i)].port2.idx; }
        for (int i = 0; i < int(faces.size()) && i < cap_f; i++)
            if (f_halfs) for (int j = 0; j < 4; j++) f_halfs[4 * i + j] = faces[size_t(i)].qhalfs[size_t(j)].idx;
    });
}

extern "C" mk_status mk_read_obj(const char* path, int cap_v, int cap_f, double* V, int32_t* F, int* nV, int* nF) {
    if (!path || !nV || !nF) return MK_ERR_ARG;
    return guarded(nullptr, [&] {
        metriko::MatXd Vm;
        metriko::MatXi Fm;
        if (!metriko::read_obj(path, Vm, Fm)) throw mk::Error(MK_ERR_ARG, "cannot open OBJ file");
        *nV = Vm.rows();
        *nF = Fm.rows();
        if (V && Vm.rows() <= cap_v) for (size_t i = 0; i < Vm.d.size(); i++) V[i] = Vm.d[i];
        if (F && Fm.rows() <= cap_f) for (size_t i = 0; i < Fm.d.size(); i++) F[i] = Fm.d[i];
    });
}

extern "C" mk_status mk_cfn_write_json(const char* path, int n, const double* cfn) {
    if (!path || n < 0 || (n && !cfn)) return MK_ERR_ARG;
    return guarded(nullptr, [&] {
        metriko::VecXc c(n);
        for (int i = 0; i < n; i++) c[i] = metriko::complex(cfn[2 * i], cfn[2 * i + 1]);
        metriko::write_cfn_json(path, c);
    });
}

extern "C" mk_status mk_cfn_read_json(const char* path, int cap, double* cfn, int* n) {
    if (!path || !n) return MK_ERR_ARG;
    return guarded(nullptr, [&] {
        metriko::MatXd uv1;
        metriko::VecXc uv2;
        if (!metriko::read_cfn_json(path, uv1, uv2)) throw mk::Error(MK_ERR_ARG, "cannot open cfn file");
        *n = uv2.size();
        if (cfn && uv2.size() <= cap)
            for (int i = 0; i < uv2.size(); i++) { cfn[2 * i] = uv2[i].real(); cfn[2 * i + 1] = uv2[i].imag(); }
    });
}

extern "C" mk_status mk_motorcycle_graph(mk_ctx* ctx, int nV, int nF, const double* V, const int32_t* F,
                                         const int32_t* matching, const int32_t* singular, const double* uv,
                                         int cap_p, int32_t* port_ids, double* port_uvdir, int cap_s,
                                         int32_t* seg_ids, double* seg_uv, int* n_ports, int* n_segs) {
    if (nV <= 0 || nF <= 0 || !V || !F || !matching || !singular || !uv || !n_ports || !n_segs) return MK_ERR_ARG;
    return guarded(ctx, [&] {
        using namespace metriko;
        Hmesh mesh(toV(nV, V), toF(nV, nF, F));
        VecXi mt(mesh.nE), sg(nV);
        for (int e = 0; e < mesh.nE; e++) mt[e] = matching[e];
        for (int v = 0; v < nV; v++) sg[v] = singular[v];
        VecXc cfn(3 * nF);
        for (int c = 0; c < 3 * nF; c++) cfn[c] = complex(uv[2 * c], uv[2 * c + 1]);
        const MotorcycleGraph g(mesh, cfn, mt, sg);
        *n_ports = int(g.mports.size());
        int ns = 0;
        for (const Mcurv& c : g.mcurvs) ns += int(c.sgmts.size());
        *n_segs = ns;
        for (int i = 0; i < int(g.mports.size()) && i < cap_p; i++) {
            const Mport& p = g.mports[size_t(i)];
            if (port_ids) {
                int32_t* o = port_ids + 5 * i;
                o[0] = p.id; o[1] = p.vert.id; o[2] = p.face.id; o[3] = p.prev; o[4] = p.next;
            }
            if (port_uvdir) {
                double* o = port_uvdir + 4 * i;
                o[0] = p.uv.real(); o[1] = p.uv.imag(); o[2] = p.dir.real(); o[3] = p.dir.imag();
            }
        }
        auto cid = [&](const Mcurv* c) { return c ? int32_t(c - g.mcurvs.data()) : int32_t(-1); };
        int k = 0;
        for (size_t ci = 0; ci < g.mcurvs.size(); ci++)
            for (const Msgmt& s : g.mcurvs[ci].sgmts) {
                if (k < cap_s) {
                    if (seg_ids) {
                        int32_t* o = seg_ids + 6 * k;
                        o[0] = int32_t(ci); o[1] = s.face.id; o[2] = s.fr.type; o[3] = s.to.type;
                        o[4] = cid(s.fr.crash); o[5] = cid(s.to.crash);
                    }
                    if (seg_uv) {
                        double* o = seg_uv + 4 * k;
                        o[0] = s.fr.uv.real(); o[1] = s.fr.uv.imag(); o[2] = s.to.uv.real(); o[3] = s.to.uv.imag();
                    }
                }
                k++;
            }
    });
}

extern "C" mk_status mk_tmesh_build(mk_ctx* ctx, int nV, int nF, const double* V, const int32_t* F,
                                    const int32_t* matching, const int32_t* singular, const double* uv, int cap_e,
                                    int32_t* tedges, int32_t* th2sing, int cap_q, int32_t* quad_ptr, int cap_t,
                                    int32_t* quad_th, int32_t* quad_side, int* n_tedges, int* n_tquads,
                                    int* n_entries) {
    if (nV <= 0 || nF <= 0 || !V || !F || !matching || !singular || !uv || !n_tedges || !n_tquads || !n_entries)
        return MK_ERR_ARG;
    return guarded(ctx, [&] {
        using namespace metriko;
        Hmesh mesh(toV(nV, V), toF(nV, nF, F));
        VecXi mt(mesh.nE), sg(nV);
        for (int e = 0; e < mesh.nE; e++) mt[e] = matching[e];
        for (int v = 0; v < nV; v++) sg[v] = singular[v];
        VecXc cfn(3 * nF);
        for (int c = 0; c < 3 * nF; c++) cfn[c] = complex(uv[2 * c], uv[2 * c + 1]);
        const MotorcycleGraph g(mesh, cfn, mt, sg);
        const Tmesh t(g.mcurvs);
        *n_tedges = int(t.nTE);
        *n_tquads = int(t.nTQ);
        int tot = 0;
        for (const Tquad& q : t.tquads) tot += int(q.thids.size());
        *n_entries = tot;
        for (int i = 0; i < int(t.nTE) && i < cap_e; i++) {
            const Tedge& e = t.tedges[size_t(i)];
            if (tedges) {
                tedges[3 * i] = int32_t(e.seg_fr.curv - g.mcurvs.data());
                tedges[3 * i + 1] = e.seg_fr.id;
                tedges[3 * i + 2] = e.seg_to.id;
            }
            if (th2sing) { th2sing[2 * i] = t.th2sing[2 * i]; th2sing[2 * i + 1] = t.th2sing[2 * i + 1]; }
        }
        if (quad_ptr && int(t.nTQ) <= cap_q) {
            int k = 0;
            quad_ptr[0] = 0;
            for (size_t q = 0; q < t.nTQ; q++) {
                for (size_t j = 0; j < t.tquads[q].thids.size(); j++, k++)
                    if (k < cap_t) {
                        if (quad_th) quad_th[k] = t.tquads[q].thids[j];
                        if (quad_side) quad_side[k] = t.tquads[q].sides[j];
                    }
                quad_ptr[q + 1] = k;
            }
        }
    });
}

extern "C" mk_status mk_qgp_quantize(mk_ctx* ctx, int nV, int nF, const double* V, const int32_t* F,
                                     const int32_t* matching, const int32_t* singular, const double* uv, int cap,
                                     double* ideal, double* lengths, int* n_tedges, int* valid) {
    if (nV <= 0 || nF <= 0 || !V || !F || !matching || !singular || !uv || !n_tedges) return MK_ERR_ARG;
    return guarded(ctx, [&] {
        using namespace metriko;
        Hmesh mesh(toV(nV, V), toF(nV, nF, F));
        VecXi mt(mesh.nE), sg(nV);
        for (int e = 0; e < mesh.nE; e++) mt[e] = matching[e];
        for (int v = 0; v < nV; v++) sg[v] = singular[v];
        VecXc cfn(3 * nF);
        for (int c = 0; c < 3 * nF; c++) cfn[c] = complex(uv[2 * c], uv[2 * c + 1]);
        const MotorcycleGraph g(mesh, cfn, mt, sg);
        const Tmesh t(g.mcurvs);
        // ideal lengths: the uv length of each T-edge's segments (example_qgp.cpp:112-123)
        VecXd R = VecXd::Zero(int(t.nTE));
        for (int i = 0; i < int(t.nTE); i++)
            for (const Msgmt& s : t.tedges[size_t(i)].segments()) R[i] += std::abs(s.diff());
        const VecXd X = compute_quantization(t, R);
        *n_tedges = int(t.nTE);
        if (valid) *valid = (validate_quantization(t, X) && compute_validation(t, X)) ? 1 : 0;
        for (int i = 0; i < int(t.nTE) && i < cap; i++) {
            if (ideal) ideal[i] = R[i];
            if (lengths) lengths[i] = X[i];
        }
    });
}
