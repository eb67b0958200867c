/* metriko_hip.h -- C ABI of the MI355X (gfx950) implementation of metriko's
 * field + IGM hot path.
 *
 * Plain C: opaque handles, int status codes, caller-owned host buffers,
 * library-owned device buffers, no exceptions across the boundary.  Every
 * call is stream-ordered on the context's stream and synchronous at return.
 * One context per host thread.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference's include/metriko/core).
 */
#ifndef METRIKO_HIP_H
#define METRIKO_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int mk_status;
enum {
    MK_OK = 0,
    MK_ERR_ARG = 1,          /* invalid argument / shape */
    MK_ERR_HIP = 2,          /* HIP runtime error */
    MK_ERR_NOT_SPD = 3,      /* CG breakdown: p^H A p <= 0 (matrix not SPD/HPD) */
    MK_ERR_NO_CONVERGENCE = 4, /* maxiter reached above tolerance */
    MK_ERR_NAN = 5,          /* NaN/Inf in inputs or iterates */
    MK_ERR_STAGNATION = 6,   /* iterative refinement stopped improving */
    MK_ERR_ALLOC = 7,        /* device allocation failed */
    MK_ERR_UNSUPPORTED = 8,  /* option not supported */
    MK_ERR_SINGULAR = 9      /* singular system (KKT / kernel basis) */
};

typedef struct mk_ctx mk_ctx;
typedef struct mk_op mk_op;

/* MK_PREC_FP64: FP64 PCG.  MK_PREC_MIXED: FP32 PCG corrections inside FP64 iterative
 * refinement.  MK_PREC_MIXED_MG: FP64 (flexible) PCG whose multigrid V-cycle runs in FP32;
 * the FP64 Krylov loop refines the FP32 preconditioner's corrections every iteration (with a
 * non-multigrid preconditioner it is plain FP64 PCG). */
enum { MK_PREC_FP64 = 0, MK_PREC_MIXED = 1, MK_PREC_MIXED_MG = 2 };
enum { MK_FMT_CSR = 0, MK_FMT_SELL = 1 };
enum { MK_PRECOND_NONE = 0, MK_PRECOND_JACOBI = 1, MK_PRECOND_AMG = 2 };  /* AMG: smoothed aggregation V-cycle */

typedef struct {
    int precision;      /* MK_PREC_*, default MK_PREC_MIXED_MG */
    double tol;         /* FP64 relative residual ||b - A x|| / ||b||, default 1e-10 */
    int maxiter;        /* max inner iterations per solve (all refinement steps), default 20000 */
    int format;         /* MK_FMT_* (device operator layout), default SELL */
    int slice_height;   /* SELL slice height C (multiple of 64), default 64 */
    int sigma;          /* SELL sorting window (rows), 1 = unsorted, default 1 */
    int block_size;     /* threads per block for SpMV: 64, 128 or 256, default 256 */
    int precond;        /* MK_PRECOND_*, default AMG */
    double inner_tol;   /* FP32 inner relative tolerance per refinement step, default 1e-5 */
    int max_refine;     /* max refinement steps, default 12 */
    int check_every;    /* iterations between host convergence checks, default 16 */
} mk_solve_opts;

typedef struct {
    int iterations;       /* inner CG iterations (all solves / refinement steps) */
    int refinements;      /* FP64 refinement steps (0 for plain FP64) */
    int solves;           /* number of linear solves */
    double rel_residual;  /* final FP64 relative residual of the last solve */
    double ms;            /* wall time of the call (ms) */
    double spmv_ms;       /* GPU time inside SpMV kernels (hipEvent) */
    int64_t spmv_calls;   /* SpMV launches */
    double spmv_bytes;    /* algorithmic bytes moved by those SpMVs */
    double setup_ms;      /* preconditioner (multigrid) setup time (ms), not in ms */
    int mg_levels;        /* multigrid levels (0 = no multigrid) */
} mk_stats;

const char* mk_status_str(mk_status s);
void mk_default_opts(mk_solve_opts* o);
mk_status mk_init(mk_ctx** ctx, int device);
mk_status mk_finalize(mk_ctx* ctx);
mk_status mk_last_error(mk_ctx* ctx, char* buf, int len);
/* Enable per-launch hipEvent timing of SpMV kernels (reported in mk_stats.spmv_ms).  While
 * timing is on, the PCG iterations run as individual launches (no graph replay). */
mk_status mk_set_timing(mk_ctx* ctx, int on);
/* Multigrid PCG iterations are captured once per operator and buffer set as a hipGraph of
 * check_every iterations and replayed (default on; off = individual launches, same results). */
mk_status mk_set_graphs(mk_ctx* ctx, int on);
/* JSON array of per-kernel SpMV timings recorded since the last reset: one object per
 * "<operator>:<kernel>:<fp32|fp64>" key with calls, ms (hipEvent), algorithmic bytes,
 * bytes per call and rows.  reset != 0 clears the table afterwards. */
mk_status mk_timing_report(mk_ctx* ctx, int reset, char* buf, int len);

/* --- generic sparse operators ---------------------------------------------------------
 * Square sparse matrix in CSR (int32 column ids, int64 row pointers), stored on the
 * device in the layout of opts->format with FP64 and FP32 value copies.
 * Replaces the factorize() step of CholeskyWrapper (igm/iterative_rounding/iter_rounding.h:14-31)
 * and suitesparse::PositiveDefiniteSolver's constructor (solver/suitesparse/suitesparse_pdefinite.h:26-41). */
mk_status mk_op_create_real(mk_ctx* ctx, int n, const int64_t* rowptr, const int32_t* col, const double* val,
                            const mk_solve_opts* opts, mk_op** op);
mk_status mk_op_create_complex(mk_ctx* ctx, int n, const int64_t* rowptr, const int32_t* col,
                               const double* val_interleaved, const mk_solve_opts* opts, mk_op** op);
mk_status mk_op_destroy(mk_op* op);
/* y = A x (FP64), host vectors (complex: interleaved re,im). */
mk_status mk_op_apply(mk_ctx* ctx, const mk_op* op, const double* x, double* y);

/* SPD / HPD solve A x = b by PCG (FP64 or FP32+FP64 refinement); x is the warm start on
 * input.  Replaces CholeskyWrapper::solve (iter_rounding.h:19) and LinearSolver::solve
 * (solver/suitesparse/suitesparse_utilities.h:16-24). */
mk_status mk_spd_solve(mk_ctx* ctx, const mk_op* op, const double* b, double* x, const mk_solve_opts* opts,
                       mk_stats* stats);

/* Host callback of the naive LM pass: iterative_rounding's iter_cb (igm/iterative_rounding/
 * iter_rounding.h:49, called through NaiveIntegration::post_iteration, iter_rounding_init.h:28,
 * after every iteration of LMSolver::solve, levenberg_marquardt.h:221-224).  It receives
 * fullx = UFull x of the current iterate (nfull values, host memory valid during the call);
 * a nonzero return ends the naive pass there, as the reference's iter_cb returning true does.
 * Set (cb != NULL) or cleared (cb == NULL) per context; while set it takes the place of the
 * device check of mk_igm_injectivity_check. */
typedef int (*mk_naive_iter_cb)(void* user, const double* fullx, int nfull);
mk_status mk_set_naive_iter_callback(mk_ctx* ctx, mk_naive_iter_cb cb, void* user);

/* Allocation-free variant of mk_spd_solve.  No counterpart in the reference: for callers that
 * cannot allocate device memory during a solve (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  b and x are DEVICE arrays (complex interleaved;
 * x is the warm start on input); FP64 PCG with opts->precond MK_PRECOND_JACOBI or _NONE
 * (other modes: MK_ERR_UNSUPPORTED).  Every work vector, reduction partial and device scalar
 * comes from the library's static per-device scratch (24 MiB, held until the process ends):
 * at most 524,288 scalar unknowns (n real, 2 n complex), larger operators get MK_ERR_ARG.
 * The call makes no device allocation and frees nothing; the operator itself is created
 * beforehand (mk_op_create_*).  All calls on one GPU share its scratch: calls that overlap
 * on one GPU (two host threads, two contexts) race. */
mk_status mk_spd_solve_noalloc(mk_ctx* ctx, const mk_op* op, const double* b_dev, double* x_dev,
                               const mk_solve_opts* opts, mk_stats* stats);

/* A X = B for m right-hand sides (B, X: n x m column-major host arrays; X starts from zero).
 * Real operators with the multigrid preconditioner: batched PCG, up to 64 columns per block
 * (one matrix read serves every column), FP64 Krylov vectors and an FP32 V-cycle (FP64 with
 * MK_PREC_FP64).  Other
 * configurations solve the columns one by one with mk_spd_solve's algorithm.  Replaces
 * repeated CholeskyWrapper::solve calls (iter_rounding.h:19). */
mk_status mk_spd_solve_many(mk_ctx* ctx, const mk_op* op, int m, const double* B, double* X,
                            const mk_solve_opts* opts, mk_stats* stats);

/* --- field (a2-a4) --------------------------------------------------------------------
 * Connection Laplacian of FaceRosyField::connectionLaplacian (vectorfield/face_rosy_field.h:46-76)
 * assembled on the device from halfedge topology: face_halfs[3f+k] = k-th halfedge of face f
 * (face2half, next, next.next), twin[h], half_face[h] (-1 on boundary), farg[h] = heArgOnFace.
 * Returns the HPD operator L = S + 1e-9 I (complex). */
mk_status mk_field_laplacian(mk_ctx* ctx, int nF, int nH, const int32_t* face_halfs, const int32_t* twin,
                             const int32_t* half_face, const double* farg, int N, const mk_solve_opts* opts,
                             mk_op** L);
/* Copies the assembled L back as CSR (sorted columns): rowptr[nF+1], col/val[nnz] (val interleaved).
 * Pass NULL arrays to query nnz. */
mk_status mk_op_export_csr(mk_ctx* ctx, const mk_op* op, int64_t* nnz, int64_t* rowptr, int32_t* col, double* val);

/* Inverse power iteration of solveSmallestEig (solver/sparse_linear_solver.h:29-54):
 * n_iter times x = L^{-1}(M u); s = sqrt|x^H M x|; x /= s; u = x.  mass[nF] = diagonal of M,
 * u0/x_out interleaved complex [2 nF].  Each solve starts from the Galerkin solution on the
 * last four iterates and aims at opts->tol / 10 (the parameterization amplifies field error
 * about a thousandfold) as far as the FP64 floor of the true residual allows; every solve
 * meets opts->tol. */
mk_status mk_field_smallest_eig(mk_ctx* ctx, const mk_op* L, const double* mass, const double* u0, int n_iter,
                                const mk_solve_opts* opts, double* x_out, mk_stats* stats);

/* --- IGM (a12-a16) ---------------------------------------------------------------------
 * Whole iterative_rounding (igm/iterative_rounding/iter_rounding.h:33-113), every flag
 * combination: NaiveIntegration KKT (iter_rounding_init.h:74-232), naive LM pass,
 * SeamlessIntegration rounding loop (iter_rounding_loop.h).  flags->local_injectivity runs
 * both LM passes with DiagonalDamping (lambda from 0.01) and the InjectiveBarrier rows
 * (injective_barrier.h:10-110), also when !seamless (then the naive pass alone).  Inputs mirror the
 * reference's 17 parameters: G2 (= G2_cut * fullx2Nfunc, CSR rows 2N*nF x cols n*nR),
 * Cfull (CSR, already reduced to independent rows), intrinsic field F2[2N*nF].
 * Outputs: fullx[n*nR]; decisions (index, value) in rounding order (ndec on return).
 * Returns MK_OK with *success = 0 when a post-check fails (reference: success=false). */
typedef struct {
    int seamless, round_seams, local_injectivity, verbose;
} mk_igm_flags;
typedef struct {
    int rounds, lm_rounds, k, kernel_rank, n_fixed;
    int null_vectors;  /* near-null vectors given to the A0 multigrid (0 = one-vector rule) */
    int batched_solves, batch_iterations;  /* Woodbury columns solved in blocks; their PCG iterations */
    int device_order;  /* 1: the A0 system ran in its multigrid aggregates' locality order on the device */
    double kkt_rel_residual, setup_ms, kkt_ms, naive_ms, rounding_ms;
    mk_stats solve;
} mk_igm_stats;
mk_status mk_iterative_rounding(mk_ctx* ctx, int nfixed, const int32_t* fixedIdcs, const double* fixedVals,
                                int nsing, const int32_t* singularIdcs, int nint, const int32_t* integerIdcs,
                                double length,
                                int c_rows, int c_cols, const int64_t* c_ptr, const int32_t* c_col, const double* c_val,
                                int g_rows, int g_cols, const int64_t* g_ptr, const int32_t* g_col, const double* g_val,
                                int N, int n, int nF, const mk_igm_flags* flags, const double* F2,
                                const mk_solve_opts* opts, double* fullx, int* success,
                                int max_dec, int32_t* dec_index, int32_t* dec_value, double* dec_margin, int* ndec,
                                mk_igm_stats* stats);

/* locally_injectivity_check (parameterization_integ.ipp:45-63), the iter_cb RosyParameterization
 * hands iterative_rounding (iter_rounding.h:49, called per naive LM iteration by
 * NaiveIntegration::post_iteration, iter_rounding_init.h:28): after this call, every
 * mk_iterative_rounding with local_injectivity whose G2 has f_cols columns ends its naive LM pass
 * once every cut face f keeps a non-negative orientation in (u, v) = functions 0 and 1 of
 * fullx2Nfunc * fullx at its corners cut_idx[3f .. 3f+2].  fullx2Nfunc: CSR, f_rows = N * cut
 * vertices, f_cols = fullx entries.  nF = 0 removes the check. */
mk_status mk_igm_injectivity_check(mk_ctx* ctx, int N, int nF, const int32_t* cut_idx, int f_rows, int f_cols,
                                   const int64_t* f_ptr, const int32_t* f_col, const double* f_val);

/* --- QEx quad extraction (host) -----------------------------------------------------------
 * example_qex.cpp:84-140 on the corner uv of an injective seamless parameterization:
 * qex::sanitization (qex/sanitization.h), generate_q_vert, generate_vqvert_qport,
 * generate_eqvert_qport, generate_fqvert_qport (qex/gen_q_port.h), generate_q_edge
 * (qex/gen_q_edge.h), generate_q_faces (qex/gen_q_face.h).  Closed meshes, N = 4.
 * matching[nE] (combed, edges in Hmesh order), singular[nV], uv[6 nF] (corner uv, corner
 * 3 f + j at F[f, j]).  Outputs fill up to their capacity; counts are always set:
 * uv_out[6 nF] the sanitized uv; q-vertices (vertex ones, then edge, then face) q_uv[2],
 * q_pos[3], q_sid[1] each; ports p_ids[6] = (vid, eid, fid, prev, next, connected),
 * p_uvdir[4] = (u, v, dir u, dir v); edges e_ports[2] (port indices); faces f_halfs[4]
 * (half 2 i = edge i from its first port, 2 i + 1 reversed).  ctx may be NULL (then no
 * error text). */
typedef struct { int n_vq, n_eq, n_fq, n_ports, n_edges, n_faces; } mk_qex_counts;
mk_status mk_qex_extract(mk_ctx* ctx, int nV, int nF, const double* V, const int32_t* F, int N,
                         const int32_t* matching, const int32_t* singular, const double* uv, double* uv_out,
                         int cap_q, double* q_uv, double* q_pos, int32_t* q_sid, int cap_p, int32_t* p_ids,
                         double* p_uvdir, int cap_e, int32_t* e_ports, int cap_f, int32_t* f_halfs,
                         mk_qex_counts* counts);

/* QGP motorcycle graph (host): MotorcycleGraph(mesh, uv, matching, singular)
 * (tmesh/motorcycle.h, example_qgp.cpp:110) on corner uv[6 nF] (as mk_qex_extract); fills up
 * to capacity, n_ports / n_segs always set: ports port_ids[5] = (id, vertex, face, prev,
 * next), port_uvdir[4] = (u, v, dir u, dir v); segments in curve order then segment order,
 * seg_ids[6] = (curve, face, from type, to type, from crash curve, to crash curve; types
 * 0 None, 1 HitL, 2 HitR, 3 HitB; -1 no curve), seg_uv[4] = (from u, v, to u, v). */
mk_status mk_motorcycle_graph(mk_ctx* ctx, int nV, int nF, const double* V, const int32_t* F,
                              const int32_t* matching, const int32_t* singular, const double* uv, int cap_p,
                              int32_t* port_ids, double* port_uvdir, int cap_s, int32_t* seg_ids, double* seg_uv,
                              int* n_ports, int* n_segs);

/* QGP T-mesh (host): Tmesh(MotorcycleGraph(...).mcurvs) (tmesh/tmesh.h, example_qgp.cpp:111).
 * T-edge i: tedges[3 i] = (curve, first segment, last segment) (segment indices within the
 * curve, as mk_motorcycle_graph lists them), half-edges 2 i (along the curve) and 2 i + 1;
 * th2sing[2 nTE] (singular vertex a half-edge leaves, -1 otherwise); T-quad q: half-edges
 * quad_th[quad_ptr[q] .. quad_ptr[q + 1]) with their sides quad_side (0..3).  Filled up to
 * capacity (cap_q = T-quads, cap_t = entries); the three counts are always set. */
mk_status mk_tmesh_build(mk_ctx* ctx, int nV, int nF, const double* V, const int32_t* F, const int32_t* matching,
                         const int32_t* singular, const double* uv, int cap_e, int32_t* tedges, int32_t* th2sing,
                         int cap_q, int32_t* quad_ptr, int cap_t, int32_t* quad_th, int32_t* quad_side,
                         int* n_tedges, int* n_tquads, int* n_entries);

/* QGP quantization (host): compute_quantization(Tmesh, R) (quantization/quantization.h) on
 * the T-mesh of mk_tmesh_build, R = each T-edge's uv length (example_qgp.cpp:112-123):
 * ideal[nTE] = R, lengths[nTE] = the integer T-edge lengths X, valid = opposite T-quad
 * sides equal and compute_validation passed.  n_tedges always set. */
mk_status mk_qgp_quantize(mk_ctx* ctx, int nV, int nF, const double* V, const int32_t* F, const int32_t* matching,
                          const int32_t* singular, const double* uv, int cap, double* ideal, double* lengths,
                          int* n_tedges, int* valid);

/* Files (host): mk_read_obj = the triangle OBJ reader example_igm / example_qex use
 * (igl::readOBJ, `v` and `f a b c` lines; call with capacities 0 for nV / nF first);
 * mk_cfn_write_json / mk_cfn_read_json = the corner-uv dump {"cfn": [re, im, ...]} of
 * example_qex.cpp:14-43 (n complex values, interleaved). */
mk_status mk_read_obj(const char* path, int cap_v, int cap_f, double* V, int32_t* F, int* nV, int* nF);
mk_status mk_cfn_write_json(const char* path, int n, const double* cfn);
mk_status mk_cfn_read_json(const char* path, int cap, double* cfn, int* n);

/* --- differentiation: the sizing hook (no reference counterpart) ---------------------------
 * mk_keep_igm_state(ctx, 1) makes the next iterative rounding (mk_iterative_rounding or
 * mk_pipeline_run) keep its last LM round's system
 *   (A0 + wK^2 U^T U) x = length wI^2 X2F^T (sigma .* F2) + wC^2 xPrev + wK^2 U^T r
 * (integer decisions, proximal point xPrev and Woodbury columns of that round held fixed);
 * mk_igm_state_take hands it to the caller.  sigma is a per-face scale of the target field
 * (sigma = exp(-mean log-size of the face's vertices); sigma = 1 reproduces the rounding's
 * x).  mk_igm_sized_uv: fullx[nfull] and corner uv[6 nF] (as mk_pipeline_run's uv; needs a
 * state from mk_pipeline_run) for a given sigma[nF].  mk_igm_sized_uv_grad: the adjoint
 * solve H lambda = UFull^T P^T g_uv (Woodbury against A0, multigrid PCG) and
 * g_scale[f] = d(g_uv . uv)/d sigma_f. */
typedef struct mk_igm_state mk_igm_state;
mk_status mk_keep_igm_state(mk_ctx* ctx, int on);
mk_status mk_igm_state_take(mk_ctx* ctx, mk_igm_state** state);
mk_status mk_igm_state_destroy(mk_igm_state* state);
mk_status mk_igm_state_sizes(const mk_igm_state* state, int* nF, int* nfull, int* k, int* nfixed,
                             int* has_corner_map);
mk_status mk_igm_sized_uv(mk_ctx* ctx, mk_igm_state* state, const double* face_scale, double* fullx, double* uv);
mk_status mk_igm_sized_uv_grad(mk_ctx* ctx, mk_igm_state* state, const double* face_scale, const double* g_uv,
                               double* g_scale);

/* --- driver ---------------------------------------------------------------------------
 * example/example_igm.cpp:14-36 in one call: Hmesh -> Smoothest field -> Curl matching ->
 * seam -> cut mesh -> combed field -> extrinsic field -> RosyParameterization(seamless,
 * !localInjectivity, !roundSeams) setup + integ.  All buffers caller-owned; sizes:
 * compressed[2nF], singular[nV], matching[nE] (combed), seam[nE], cut_idx[3nF], uv[6nF]
 * (corner uv, 3 per face), fullx[fullx_cap].  u0 may be NULL (glibc rand() stream as the
 * reference's VecXc::Random).  Query sizes with mk_pipeline_sizes first. */
typedef struct {
    int nV, nF, nE, cutV, nS, nT, k, nfullx, ndec, success;
    double ms_hmesh, ms_field_asm, ms_field_solve, ms_matching, ms_seam, ms_cut, ms_setup, ms_integ, ms_total;
    mk_stats field;
    mk_igm_stats igm;
} mk_pipeline_info;
mk_status mk_pipeline_run(mk_ctx* ctx, int nV, int nF, const double* V, const int32_t* F, int N, double ratio,
                          const double* u0, int n_iter, const mk_solve_opts* field_opts,
                          const mk_solve_opts* igm_opts,
                          double* compressed, int32_t* singular, int32_t* matching, uint8_t* seam,
                          int32_t* cut_idx, double* uv, int fullx_cap, double* fullx,
                          int dec_cap, int32_t* dec_index, int32_t* dec_value, mk_pipeline_info* info);

/* Host-only stages after the field solve (no GPU needed): roots, Curl matching,
 * singularities, seam, cut mesh, combed matching, extrinsic field and the IGM setup sizes.
 * compressed[2nF] in; singular[nV], raw_matching[nE], matching[nE] (combed), seam[nE],
 * cut_idx[3nF], sizes[4] = {cutV, nS, nT, constraint rows}, ext[12nF] out (any may be NULL). */
mk_status mk_combinatorics(int nV, int nF, const double* V, const int32_t* F, int N, const double* compressed,
                           int32_t* singular, int32_t* raw_matching, int32_t* matching, uint8_t* seam,
                           int32_t* cut_idx, int32_t* sizes, double* ext);

/* --- field types and matchings beyond the example_igm flow -------------------------------
 * mk_field_solve: the FaceRosyField(mesh, N, type) constructor (face_rosy_field.h:17-44) in
 * one call.  MK_FIELD_SMOOTHEST: 50-step inverse iteration as mk_field_smallest_eig (u0 NULL:
 * glibc rand() stream).  MK_FIELD_CURVATURE_ALIGNED (N = 2 or 4): D = principal curvature
 * direction (face_rosy_field.h:86-98), squared for N = 4, and one Hermitian positive definite
 * solve L x = M D / sqrt|D^H M D| (face_rosy_field.h:25-34; solveSquare,
 * sparse_linear_solver.h:16-27).  compressed[2nF] out (interleaved). */
enum { MK_FIELD_SMOOTHEST = 0, MK_FIELD_CURVATURE_ALIGNED = 1 };
mk_status mk_field_solve(mk_ctx* ctx, int nV, int nF, const double* V, const int32_t* F, int N, int field_type,
                         const double* u0, int n_iter, const mk_solve_opts* opts, double* compressed,
                         mk_stats* stats);
/* principalCurvatureDir (face_rosy_field.h:86-98), host only: D[2nF] interleaved. */
mk_status mk_curvature_dir(int nV, int nF, const double* V, const int32_t* F, double* D);
/* computeMatching(MatchingType) (face_rosy_field.h:119-135) on a given compressed field, host
 * only: Principal (pMatching, :138-167) or Curl (cMatching, :169-202), then computeSingular
 * (:100-107).  singular[nV], matching[nE], effort[nE] out (any may be NULL). */
enum { MK_MATCH_PRINCIPAL = 0, MK_MATCH_CURL = 1 };
mk_status mk_field_matching(int nV, int nF, const double* V, const int32_t* F, int N, const double* compressed,
                            int matching_type, int32_t* singular, int32_t* matching, double* effort);

/* --- batched configuration (no reference counterpart as a whole) ---------------------------
 * mk_autotune_spmv: times the FP64 SpMV of an n x n CSR matrix (complex when cplx, values
 * interleaved) in each candidate device layout -- the CSR-like layout (slice height 64,
 * unsorted) and SELL slice heights 64/128/256 x sorting windows 1/256 x block sizes 128/256 --
 * and writes the fastest into *best (a copy of *base with format, slice_height, sigma and
 * block_size set); *us receives its average launch time in microseconds. */
mk_status mk_autotune_spmv(mk_ctx* ctx, int n, const int64_t* rowptr, const int32_t* col, const double* val,
                           int cplx, const mk_solve_opts* base, mk_solve_opts* best, double* us);
/* mk_batch_run: for each of nmesh closed meshes (V[m]: nV[m] x 3, F[m]: nF[m] x 3), the
 * Smoothest N-RoSy field (n_iter inverse-iteration steps, face_rosy_field.h:17-44 with
 * solveSmallestEig, sparse_linear_solver.h:29-54; u0[m] interleaved [2 nF[m]], or NULL for the
 * std::rand stream) and the u/v Poisson integration of its first direction g_u (root 0 in the
 * face basis) and g_v = n x g_u: min sum_f A_f |grad u - g_f|^2 with vertex 0 pinned to 0.
 * All meshes share one block-diagonal operator per stage (one launch per kernel for the whole
 * batch, the field and stiffness layouts each picked by mk_autotune_spmv on a row sample; every
 * batch of PCG iterations replays one captured hipGraph over all meshes); every mesh's solves
 * meet opts->tol relative to its own right-hand side.  Outputs (either may be NULL per mesh):
 * compressed[m] [2 nF[m]] interleaved, uv[m] [2 nV[m]] = (u, v) per vertex. */
typedef struct {
    int meshes;
    int64_t faces, verts;
    int field_iterations, poisson_iterations;          /* PCG iterations over all solves */
    double field_rel_residual, poisson_rel_residual;   /* bound on every mesh's own relative residual */
    double ms_host_setup, ms_field, ms_poisson, ms_total;
    int tuned_slice_height, tuned_sigma, tuned_block;  /* layout picked for the field operator */
    double tuned_spmv_us;                              /* its SpMV time on the row sample */
    int tuned_poisson_slice_height, tuned_poisson_sigma, tuned_poisson_block;  /* the stiffness operator's */
    double tuned_poisson_spmv_us;
    int64_t graph_replays;   /* PCG iteration batches replayed as one captured hipGraph each */
    int64_t graph_captures;  /* graphs captured (one per operator and buffer set) */
    int64_t eager_batches;   /* iteration batches launched kernel by kernel (graphs off or timing on) */
} mk_batch_stats;
mk_status mk_batch_run(mk_ctx* ctx, int nmesh, const int32_t* nV, const int32_t* nF, const double* const* V,
                       const int32_t* const* F, int N, int n_iter, const double* const* u0,
                       const mk_solve_opts* opts, double* const* compressed, double* const* uv,
                       mk_batch_stats* stats);

/* Host-only multigrid PCG (no GPU; test and tuning aid): builds the same smoothed-aggregation
 * hierarchy the device path uses for the CSR matrix (complex when cplx != 0, values
 * interleaved) and solves A x = b with a host V-cycle.  Returns iterations, levels and the
 * final relative residual; maxit = 0 applies one V-cycle (x = M^-1 b).  nullvec (n x nnull,
 * column-major, real operators) are near-null vectors for the aggregation. */
mk_status mk_mg_host_pcg(int n, const int64_t* rowptr, const int32_t* col, const double* val, int cplx,
                         const double* nullvec, int nnull, const double* b, double* x, double tol, int maxit,
                         int* iters, int* levels, double* relres);

/* Halfedge arrays of the library's Hmesh, as mk_field_laplacian takes them: face_halfs[3nF],
 * twin[nH], half_face[nH], farg[nH] (heArgOnFace), face_area[nF]; nH = 2 nE (query with
 * mk_hmesh_export).  Host code, no GPU needed. */
mk_status mk_hmesh_halfedges(int nV, int nF, const double* V, const int32_t* F, int32_t* face_halfs,
                             int32_t* twin, int32_t* half_face, double* farg, double* face_area);

/* Hmesh topology/geometry export for tests: edge2vert[2nE], edge2face[2nE], face2edge[3nF],
 * vert2half[nV], farg[2nE], angle_defect[nV]; nE returned in *nE (pass NULL arrays to query). */
mk_status mk_hmesh_export(int nV, int nF, const double* V, const int32_t* F, int* nE, int32_t* edge2vert,
                          int32_t* edge2face, int32_t* face2edge, int32_t* vert2half, double* farg,
                          double* angle_defect);

#ifdef __cplusplus
}
#endif
#endif
