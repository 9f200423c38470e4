/*
 * emfa_hip.h -- C ABI of the MI355X (gfx950) method-of-moments library.
 *
 * Plain C: opaque handles, pointers and sizes; no exceptions cross this
 * boundary.  Every call returns an emfa_status; emfa_last_error() gives
 * the message of the last failure on the calling thread.
 *
 * Each entry point replaces a piece of the reference's in-process C++ path
 * (cited as reference file:line).  The reference has no device boundary;
 * these calls are what its solver classes (EMSolver/EFIE.cpp, CFIE.cpp,
 * MFIE.cpp) bind instead of their CPU loops.
 *
 * Layouts (device side, owned by handles):
 *   complex matrices: row-major (row = test function), interleaved re/im,
 *                     complex<float> (EMFA_F32) or complex<double> (EMFA_F64)
 *   mesh:             vertices double[nv][3], triangles int[nt][3] (0-based)
 *   RWG list:         int[n][6] = {v1, v2, vxp, vxm, tp, tm}
 *                     (EMComponent/CommonEdge.h:14-43, SCE fields)
 */
#ifndef EMFA_HIP_H
#define EMFA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    EMFA_OK = 0,
    EMFA_E_INVALID = 1,  /* bad argument / shape                     */
    EMFA_E_OOM = 2,      /* device allocation failed / over budget     */
    EMFA_E_SINGULAR = 3, /* exactly zero pivot during factorisation    */
    EMFA_E_NOCONV = 4,   /* iterative method did not converge          */
    EMFA_E_HIP = 5,      /* HIP runtime error                          */
    EMFA_E_NODEVICE = 6  /* no usable gfx950 device                    */
} emfa_status;

enum { EMFA_F32 = 0, EMFA_F64 = 1 };
enum { EMFA_EFIE = 0, EMFA_CFIE = 1, EMFA_MFIE = 2 /* excitation only: MFIE.cpp:289-322 */ };

typedef struct emfa_ctx emfa_ctx;
typedef struct emfa_mesh emfa_mesh;
typedef struct emfa_mat emfa_mat;
typedef struct emfa_lu emfa_lu;

/* ---- library / context --------------------------------------------- */
const char *emfa_last_error(void);
int emfa_version(void);
/* device: HIP ordinal.  Creates one HIP stream owned by the context.
 * Replaces nothing in the reference (it has no device). */
emfa_status emfa_ctx_create(int device, emfa_ctx **out);
emfa_status emfa_ctx_destroy(emfa_ctx *ctx);
emfa_status emfa_ctx_sync(emfa_ctx *ctx);
/* the context's hipStream_t, as an opaque pointer */
void *emfa_ctx_stream(emfa_ctx *ctx);
/* bytes currently allocated by this context's handles */
size_t emfa_ctx_bytes_in_use(const emfa_ctx *ctx);
/* number of compute units of the context's device */
int emfa_ctx_cu_count(const emfa_ctx *ctx);
/* The compute units this context may count on while other contexts share
 * the device (0: all): its trailing updates run at most that many
 * workgroups, and (unless emfa_ctx_set_panel_cap set a bound) so does its
 * LU panel, whose workgroups wait on each other.  May be changed from
 * another thread while the context works; each launch reads the current
 * value.  No counterpart in the reference (it has no device). */
emfa_status emfa_ctx_set_cu_share(emfa_ctx *ctx, int cus);
/* Caps the workgroups of this context's LU panel (whose workgroups wait on
 * each other) at `wgs` (0: the CU share).  Contexts that factor concurrently
 * keep the sum of their caps within emfa_ctx_cu_count so every panel's
 * workgroups can be resident together. */
emfa_status emfa_ctx_set_panel_cap(emfa_ctx *ctx, int wgs);
/* on != 0: FP32 factorisations without interchanges and with the bf16-split
 * update (panel variant 2, update mode 1) run as HIP graphs.  The launch
 * sequence of a factorisation depends only on its size, the batch count and
 * the context's LU configuration (block, panel caps, CU share), so it is
 * captured once per such configuration and replayed for every later matrix
 * or batch of matrices of that configuration; the matrices are reached
 * through a device array of their bases that is rewritten before each
 * replay.  Single factorisations take this path as batches of one.  No
 * counterpart in the reference (it has no device). */
emfa_status emfa_ctx_set_graphs(emfa_ctx *ctx, int on);
/* graphs captured and graph launches made by this context so far */
emfa_status emfa_ctx_graph_stats(const emfa_ctx *ctx, long long *captures, long long *launches);

/* ---- host-side geometry (no device work) ---------------------------- */
/* Synthetic meshes of the workloads: icosahedron subdivided `subdiv` times
 * and projected on the sphere (T=20*4^s), scaled per axis for ellipsoids.
 * Replaces the reference's file loaders EMComponent/Mesh.cpp:22-49 for the
 * synthetic configs.  Sizes: nv = 10*4^s+2, nt = 20*4^s. */
emfa_status emfa_geom_icosphere(int subdiv, double ax, double ay, double az, double *xyz, int nv_cap, int *tri,
                                int nt_cap, int *nv, int *nt);
/* Square plate in z=0, side `size`, n x n cells, 2 triangles per cell. */
emfa_status emfa_geom_plate(int cells, double size, double *xyz, int nv_cap, int *tri, int nt_cap, int *nv, int *nt);
/* RWG list with the reference's ordering (EMComponent/CommonEdge.cpp:74-106,
 * 704-743: first-appearance order of shared edges, v1->v2 as in tp),
 * built in O(E) with a hash map.  Returns the RWG count in *n. */
emfa_status emfa_rwg_build(int nt, const int *tri, int *rwg, int cap, int *n, int *n_boundary);
/* Half-RWG list of the IP-SIE solver (FastWT): replaces
 * component::ContourEdge::buildEdge(mesh, policy::HALF_RWG)
 * (EMComponent/ContourEdge.cpp:15-95; the edge list of :216-252).  Conformal
 * pairs from shared edges and from boundary edges whose end points coincide
 * with a later boundary edge's (reversed), then the nonconformal half-RWGs:
 * boundary edges overlapped by anti-parallel, nearly collinear (cosines >
 * 0.95) boundary edges, each with its contour lines, all in FP32 vertex
 * arithmetic as the reference's.  Outputs (pass NULL to get the counts only):
 *   hrwg     cap x 6 ints per half-RWG: v1, v2, vx (free vertex), triangle,
 *            edge number (conformal pair / nonconformal index), type (0
 *            conformal, 1 nonconformal); length[cap]: FP32 edge length;
 *   line_ptr (n_nc + 1) over the nonconformal half-RWGs, line_tri[line_cap]
 *            (the triangle of the overlapping edge), line_seg[line_cap x 6]
 *            (start xyz, end xyz), as ContourEdge::line_list_. */
emfa_status emfa_contour_build(int nv, const double *xyz, int nt, const int *tri, int *hrwg, double *length,
                               int cap, int *n, int *line_ptr, int *line_tri, double *line_seg, int line_cap,
                               int *n_nc, int *n_lines);

/* ---- mesh upload ---------------------------------------------------- */
emfa_status emfa_mesh_upload(emfa_ctx *ctx, const double *xyz, int nv, const int *tri, int nt, const int *rwg, int n,
                             emfa_mesh **out);
emfa_status emfa_mesh_destroy(emfa_mesh *mesh);
int emfa_mesh_unknowns(const emfa_mesh *mesh);

/* ---- dense complex matrices ----------------------------------------- */
emfa_status emfa_mat_create(emfa_ctx *ctx, int rows, int cols, int prec, emfa_mat **out);
emfa_status emfa_mat_destroy(emfa_mat *m);
/* host buffers are interleaved complex, row-major, double precision
 * (converted to the matrix precision on upload) */
emfa_status emfa_mat_upload(emfa_mat *m, const double *host);
emfa_status emfa_mat_download(const emfa_mat *m, double *host);
emfa_status emfa_mat_copy(emfa_mat *dst, const emfa_mat *src); /* converts precision */
int emfa_mat_rows(const emfa_mat *m);
int emfa_mat_cols(const emfa_mat *m);
int emfa_mat_prec(const emfa_mat *m);
/* raw device pointer of the matrix storage */
void *emfa_mat_device_ptr(emfa_mat *m);

/* ---- impedance fill (K1 regular, K2 singular, gather) --------------- */
typedef struct {
    int kind;        /* EMFA_EFIE or EMFA_CFIE                                  */
    int quad;        /* 3 (reference rule) or 7 points per triangle             */
    double k;        /* wave number 2 pi f / c                                  */
    double alpha;    /* CFIE combination (EMSolver/Miscellaneous.cpp:29)        */
    double sigma_re; /* EFIE surface impedance (EMSolver/EFIE.cpp:297-306)      */
    double sigma_im;
    int cfie_mode;   /* 0: reference self term (Miscellaneous.cpp:32-77), 1: EFIE-style extraction */
} emfa_fill_params;

/* Fills Z (n x n, precision of Z).  Replaces EFIE::fillZ
 * (EMSolver/EFIE.cpp:603-696) and CFIE::fillZ (EMSolver/CFIE.cpp:485-568),
 * EDM far-field shortcut (EFIE.cpp:647-651) not applied. */
emfa_status emfa_fill(emfa_ctx *ctx, const emfa_mesh *mesh, const emfa_fill_params *p, emfa_mat *Z);

/* ---- excitation (K3) ------------------------------------------------- */
/* inc: nrhs x 9 doubles {k_hat(3), e_hat(3), h_hat(3)} (EMSolver/EM.cpp:33-41).
 * B: n x nrhs.  Replaces EFIE::fillV (EFIE.cpp:568-601,698-722),
 * CFIE::fillV (CFIE.cpp:570-597, Miscellaneous.cpp:152-189) and, for kind
 * EMFA_MFIE, MFIE::fillV (MFIE.cpp:289-322,425-437) scaled by Z0 to match a
 * CFIE fill with alpha = 0. */
emfa_status emfa_rhs_fill(emfa_ctx *ctx, const emfa_mesh *mesh, int kind, double k, double alpha, int nrhs,
                          const double *inc, emfa_mat *B);

/* ---- dense LU (K4 panel, K5 MFMA trailing update, K6 triangular solve) */
/* Factor a copy of Z in precision `prec` with partial pivoting (LAPACK
 * getrf semantics).  Replaces arma::solve(...,fast) -> cgesv
 * (EMSolver/MFIE.cpp:64).  Returns EMFA_E_SINGULAR for an exactly zero
 * pivot (the factor is still produced, like LAPACK's info > 0). */
emfa_status emfa_lu_factor(emfa_ctx *ctx, const emfa_mat *Z, int prec, emfa_lu **out);
/* Factor `count` matrices of one size: out[z] as emfa_lu_factor(Z[z]) would
 * give it.  In FP32 with panel variant 2 and update mode 1 the batch shares
 * one launch sequence (every kernel takes the batch on blockIdx.z; a factor
 * that grows is redone alone with interchanges); otherwise the matrices are
 * factored one after another.  EMFA_E_SINGULAR if any factor has an exactly
 * zero pivot (all factors are still produced).  No reference counterpart:
 * the reference factors one matrix per solve (EMSolver/MFIE.cpp:64). */
emfa_status emfa_lu_factor_batch(emfa_ctx *ctx, const emfa_mat *const *Z, int count, int prec, emfa_lu **out);
emfa_status emfa_lu_destroy(emfa_lu *lu);
/* In place: B (n x nrhs, any precision) <- A^{-1} B using the factor. */
emfa_status emfa_lu_solve(emfa_ctx *ctx, const emfa_lu *lu, emfa_mat *B);
/* pivots (0-based, like the oracle) and the factor itself */
emfa_status emfa_lu_pivots(const emfa_lu *lu, int *ipiv_host);
emfa_mat *emfa_lu_matrix(emfa_lu *lu);
int emfa_lu_info(const emfa_lu *lu);

/* LU parameters of a context: outer block nb (0: 256), panel workgroup cap
 * (0: all CUs or the context's CU share), panel variant (1: packet exchange,
 * 0: arrival counter, -1: built-in choice, 2: FP32 factorisations without
 * row interchanges -- meant for FP32 + FP64-refinement solves; a zero pivot
 * or a multiplier above 100 in modulus makes emfa_lu_factor redo the
 * factorisation with partial pivoting; FP64 factorisations always pivot).
 * No counterpart in the reference (LAPACK chooses its own blocking). */
emfa_status emfa_lu_set_params(emfa_ctx *ctx, int nb, int panel_workgroups, int panel_variant);
/* Trailing-update arithmetic of FP32 factorisations: 0 = FP32 MFMA
 * (V_MFMA_F32_32X32X2_F32, default), 1 = both operands split once per outer
 * step into bf16 planes hi = bf16(x), lo = bf16(x - hi) and multiplied on
 * bf16 MFMA with FP32 accumulation (hi*hi + hi*lo + lo*hi, about 16
 * significant bits per product; meant for FP32 + FP64-refinement solves).
 * No reference counterpart. */
emfa_status emfa_lu_set_update_mode(emfa_ctx *ctx, int mode);
/* How a factor was computed: interchanges 1 (partial pivoting) or 0 (none);
 * fallback 1 when a factorisation without interchanges grew and was redone
 * with them. */
emfa_status emfa_lu_get_pivoting(const emfa_lu *lu, int *interchanges, int *fallback);
typedef struct {
    int nb, panel_variant, panel_workgroups, update_mode; /* chosen */
    float best_ms;                                        /* its factorisation time */
    int candidates;                                       /* candidates timed */
    int cand_nb[16], cand_variant[16], cand_workgroups[16], cand_update_mode[16];
    float cand_ms[16];
} emfa_lu_tuning;
/* Autotuner: times the LU of a generated n x n matrix in precision prec for
 * each candidate (nb 256/512 x panel variant x panel workgroup cap all/128;
 * with for_refinement and FP32 also the bf16-split update and the panel
 * without interchanges), keeps the fastest in the context and reports all
 * timings. */
emfa_status emfa_lu_autotune(emfa_ctx *ctx, int n, int prec, int for_refinement, emfa_lu_tuning *out);

/* ---- mixed-precision refinement -------------------------------------- */
typedef struct {
    int iterations;      /* refinement steps taken                       */
    double relres_first; /* max over columns of ||b - A x|| / ||b||, before */
    double relres_final; /* ... after the last step                       */
} emfa_refine_report;
/* X (F64, n x nrhs) <- solution of Z X = B, starting from the factor's
 * solution and refining with FP64 residuals R = B - Z X (Z read in its own
 * precision, products accumulated in FP64) until every column's relative
 * residual <= tol or maxit steps.  No reference counterpart: it is the
 * precision form built beside the reference's dense solve. */
emfa_status emfa_refine(emfa_ctx *ctx, const emfa_mat *Z, const emfa_lu *lu, const emfa_mat *B, emfa_mat *X,
                        double tol, int maxit, emfa_refine_report *rep);

/* ---- products -------------------------------------------------------- */
/* C (F64) = alpha_sign * (A * Xin) + beta * C, A any precision, Xin F64;
 * accumulated in FP64.  alpha_sign is +1 or -1, beta 0 or 1. */
emfa_status emfa_gemm_residual(emfa_ctx *ctx, const emfa_mat *A, const emfa_mat *Xin, emfa_mat *C, int negate,
                               int accumulate);

/* ---- GMRES(m) (K7 matvec, K8 FP64 orthogonalisation) ----------------- */
typedef struct {
    int iterations;
    int restarts;
    double relres;
} emfa_gmres_report;
/* Solves Z x = b column by column (b, x: n x nrhs F64 matrices) with
 * restarted GMRES(m), FP64 Gram-Schmidt, relative tolerance.  Replaces
 * iml::GMRES (EMIterative/iml.cpp:156-207, called at EMSolver/EFIE.cpp:79)
 * with its conjugated Arnoldi coefficient and unbounded restarts fixed. */
emfa_status emfa_gmres(emfa_ctx *ctx, const emfa_mat *Z, const emfa_mat *B, emfa_mat *X, int m, double tol,
                       int max_restart, emfa_gmres_report *rep);

/* ---- BiCGSTAB --------------------------------------------------------- */
typedef struct {
    int iterations;  /* steps, summed over the right-hand sides */
    int max_steps;   /* most steps any right-hand side took */
    double relres;   /* largest ||b - Z x|| / ||b|| of the returned x */
} emfa_bicgstab_report;
/* Solves Z x = b column by column from x0 = 0 (X is overwritten), FP64
 * vectors and inner products, the matrix in its own precision.
 * variant 0 replaces iml::BICGSTAB(const Qcx_mat&, ...) (EMIterative/
 * iml.cpp:3-39; EMSolver/CFIE.cpp:76 calls it with tol 0.01, 500 steps):
 * stops when ||s|| / ||b|| < tol, EMFA_E_NOCONV (X still written) when a
 * column ran all max_iter steps, as the reference's `false`.
 * variant 1 replaces the matrix-free overload (iml.cpp:88-154): stops when
 * ||r|| / ||b|| <= tol or after max_iter steps, restarts the shadow
 * residual when rho vanishes, and never reports failure (as the reference). */
emfa_status emfa_bicgstab(emfa_ctx *ctx, const emfa_mat *Z, const emfa_mat *B, emfa_mat *X, double tol, int max_iter,
                          int variant, emfa_bicgstab_report *rep);

/* emfa_bicgstab without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The solver's vectors and partial sums live in a fixed 8 MiB
 * device array inside the library (one per device, held until the process
 * ends), so Z may have at most EMFA_BICGSTAB_NOALLOC_MAX_N rows; larger
 * systems get EMFA_E_INVALID.  Calls on one device share that array: calls
 * that overlap in time on one device race and must be serialised by the
 * caller.  Same results as emfa_bicgstab. */
#define EMFA_BICGSTAB_NOALLOC_MAX_N 58000
emfa_status emfa_bicgstab_noalloc(emfa_ctx *ctx, const emfa_mat *Z, const emfa_mat *B, emfa_mat *X, double tol,
                                  int max_iter, int variant, emfa_bicgstab_report *rep);

/* ---- far field (K9) --------------------------------------------------- */
/* RCS in dB for ndir directions dirs[ndir][3]; direction d uses column
 * cols[d] of X (cols may be NULL: column 0).  Replaces getBiRCS
 * (EMSolver/EFIE.cpp:788-832, CFIE.cpp:643-685). */
emfa_status emfa_rcs(emfa_ctx *ctx, const emfa_mesh *mesh, double k, const emfa_mat *X, int ndir, const double *dirs,
                     const int *cols, double *out_db);

/* ---- instrumentation -------------------------------------------------- */
/* ---------------------------------------------------------------------
 * MLFMA operator (EMSolver/MLFMA.cpp).  emfa_mlfma_create replaces
 * MLFMA::buildBox, levelRootsAndWeights, calculateInterpolationMtr,
 * calculateLevelK, calculateLevelTrans, calculateRadAndRecv and
 * calculateNearZ (MLFMA.cpp:688-1422); emfa_mlfma_matvec replaces
 * MLFMA::matrixVectorMultiply (MLFMA.cpp:454-509) and emfa_mlfma_gmres the
 * iterative solve (MLFMA.cpp:402-452, GMRES(m) here).  Vectors are in RWG
 * order (the reference keeps them in box order internally, as does this
 * operator).  Near and far parts are complex FP32; self terms FP64. */
typedef struct emfa_mlfma emfa_mlfma;
typedef struct {
    int kind;             /* EMFA_CFIE (MLFMA.cpp's formulation) or EMFA_EFIE */
    double k;             /* wavenumber 2 pi f / c0 */
    double alpha;         /* CFIE combination factor (alpha:) */
    double box_threshold; /* bound on the finest box edge, in wavelengths (mlfma_box_threshold:) */
    int cfie_mode;        /* CFIE self terms: 0 as Miscellaneous.cpp:32-77, 1 accurate */
} emfa_mlfma_params;
typedef struct {
    int levels;           /* finest level index; boxes are at levels 1..levels */
    long long near_nnz;   /* entries of the near matrix */
    size_t bytes;         /* device bytes held by the operator */
    int boxes[16];        /* non-empty boxes per level */
    int multipoles[16];   /* L per level (levels >= 2) */
    double box_edge[16];  /* box edge per level (m) */
    long long far_pairs[16]; /* (box, far box) translations per level */
    int grid[3];          /* single-level operator: boxes per axis of its grid (0 for the MLFMA) */
} emfa_mlfma_info;
emfa_status emfa_mlfma_create(emfa_ctx *ctx, const emfa_mesh *mesh, const emfa_mlfma_params *p, emfa_mlfma **out);
emfa_status emfa_mlfma_destroy(emfa_mlfma *m);
emfa_status emfa_mlfma_get_info(const emfa_mlfma *m, emfa_mlfma_info *info);
/* Per box of a level (1..info.levels, box order): near boxes, far boxes and
 * unknowns; each output may be null.  FMM::reportBoxInfo (FMM.cpp:759-772)
 * writes these counts to ffm_box_log.txt. */
emfa_status emfa_mlfma_boxes(const emfa_mlfma *m, int level, int *nnear, int *nfar, int *nunk);

/* Single-level fast multipole operator: replaces FMM::init's box set-up,
 * preCalculateKArray / preCalculateTrans / preCalculateRadAndRecv and
 * calculateNearZ (EMSolver/FMM.cpp:23-58,157-206,322-573,
 * EMComponent/FMM_Boundary.cpp:15-53).  p->box_threshold is the box length
 * in wavelengths (`fmm_box_length:`); the grid, the box of each RWG and L
 * are computed in FP32 as the reference computes them.  Every box that is
 * not one of the 27 grid neighbours of a box is far from it (one level, no
 * interpolation).  The handle is an emfa_mlfma: emfa_mlfma_matvec
 * (FMM::matrixVectorMultiply, FMM.cpp:270-320), emfa_mlfma_bicgstab (variant
 * 0 is FMM::iterativeSolve, FMM.cpp:222-268), emfa_mlfma_gmres,
 * emfa_mlfma_get_info (levels = 1, the grid at index 1) and
 * emfa_mlfma_destroy apply to it. */
emfa_status emfa_fmm_create(emfa_ctx *ctx, const emfa_mesh *mesh, const emfa_mlfma_params *p, emfa_mlfma **out);

/* IP-SIE fast operator on half-RWGs: replaces FastWT::init's box set-up and
 * FastWT::solve's set-up stages (EMSolver/FastWT.cpp:26-70,167-400 tree and
 * samples as the MLFMA's, calculateRadAndRecv :583-653, calculateNearZ
 * :689-784).  The half-RWG list is emfa_contour_build's (n entries of hrwg /
 * length, the contour lines of its n_nc nonconformal edges); the mesh is the
 * one it was built from (its RWG list is not used).  Near entry of field f,
 * source s: l_f l_s K(f, s) + alpha (line + ip), K the CFIE triangle-pair
 * kernel, line the consistency terms of both edges against the other's
 * triangle, ip the interior penalty 2 beta Z0 / k times a length (the edge's
 * own, or that of f's contour line on s's triangle); beta =
 * stabilization_factor |log10(average_size)| (FastWT.cpp:48).  Patterns carry
 * the scalar line channel besides the vector ones (far part of
 * FastWT::matrixVectorMultiply, :1006-1267).  p->kind must be EMFA_CFIE.  The
 * handle is an emfa_mlfma: emfa_mlfma_matvec, emfa_mlfma_bicgstab (variant 0
 * = FastWT::iterativeSolve, :916-962), emfa_mlfma_gmres and
 * emfa_mlfma_precondition (FastWT::calculatePreconditioner, :808-914) apply;
 * vectors are in half-RWG order. */
emfa_status emfa_fastwt_create(emfa_ctx *ctx, const emfa_mesh *mesh, int n, const int *hrwg, const double *length,
                               int n_nc, const int *line_ptr, const int *line_tri, const double *line_seg,
                               const emfa_mlfma_params *p, double beta, emfa_mlfma **out);
/* Excitation of every half-RWG (FastWT::fillVbyBox / cVKernel,
 * FastWT.cpp:786-806,1475-1492; 3-point rule): B is F64 n x nrhs,
 * inc nrhs x 9 {k, e, h} as for emfa_rhs_fill. */
emfa_status emfa_fastwt_rhs(emfa_ctx *ctx, const emfa_mlfma *m, const emfa_mesh *mesh, emfa_mat *B, int nrhs,
                            const double *inc);
/* Bistatic RCS in dB of column col of X (half-RWG order) at ndir unit
 * directions (FastWT::getBiRCS, FastWT.cpp:1393-1420; 3-point rule). */
emfa_status emfa_fastwt_rcs(emfa_ctx *ctx, const emfa_mlfma *m, const emfa_mesh *mesh, const emfa_mat *X, int col,
                            int ndir, const double *dirs, double *rcs_db);
/* Y = Z_mlfma X for F64 n x ncols matrices (column by column). */
emfa_status emfa_mlfma_matvec(emfa_ctx *ctx, emfa_mlfma *m, const emfa_mat *X, emfa_mat *Y);
/* GMRES(restart) with FP64 orthogonalisation on the MLFMA operator; X holds
 * the initial guess.  EMFA_E_NOCONV when a column misses tol. */
emfa_status emfa_mlfma_gmres(emfa_ctx *ctx, emfa_mlfma *m, const emfa_mat *B, emfa_mat *X, int restart, double tol,
                             int max_restart, emfa_gmres_report *rep);

/* BiCGSTAB on the MLFMA operator (MLFMA::iterativeSolve, MLFMA.cpp:402-452,
 * runs variant 0 of emfa_bicgstab); X is overwritten.  With a
 * preconditioner (emfa_mlfma_precondition) this and emfa_mlfma_gmres solve
 * P Z X = P B and report residuals relative to ||P B||, as the reference
 * does. */
emfa_status emfa_mlfma_bicgstab(emfa_ctx *ctx, emfa_mlfma *m, const emfa_mat *B, emfa_mat *X, double tol,
                                int max_iter, int variant, emfa_bicgstab_report *rep);
/* Sparse approximate inverse of MLFMA::calculatePreconditioner
 * (MLFMA.cpp:714-741,1294-1391; `preconditioning: true`): for each leaf box,
 * the near entries between the unknowns of its near boxes whose RWG edge
 * midpoints lie within row_threshold / col_threshold leaf edges of its
 * centre (`row_threshold:`, `col_threshold:`), solved in the least-squares
 * sense by Householder QR (FP64, on the host's threads) for each of the
 * box's own unknowns; P is kept on the device (FP32, CSR).  Afterwards
 * emfa_mlfma_matvec returns P Z X (MLFMA.cpp:506-508).  `entries` (may be
 * null) receives P's nonzero count.  The MLFMA operator only. */
emfa_status emfa_mlfma_precondition(emfa_ctx *ctx, emfa_mlfma *m, const emfa_mesh *mesh, double row_threshold,
                                    double col_threshold, long long *entries);

/* When enabled, each launch of a named kernel family is bracketed by HIP
 * events on the context stream; totals are read with emfa_prof_get.
 * Families: "fill_pairs", "fill_gather", "lu_gemm", "lu_panel", "lu_trsm",
 * "lu_laswp", "solve_trsm", "refine_gemm", "rhs", "rcs". */
emfa_status emfa_prof_enable(emfa_ctx *ctx, int on);
emfa_status emfa_prof_get(emfa_ctx *ctx, const char *family, double *total_ms, long long *launches,
                          double *flops_or_bytes);
emfa_status emfa_prof_reset(emfa_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* EMFA_HIP_H */
