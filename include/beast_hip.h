/*
 * beast_hip.h -- C ABI of the MI355X-native Galerkin BEM library.
 *
 * Drop-in boundary for the reference's
 *     assemble(op, X, Y; quadstrat)           src/operator.jl:100-116
 * on op in {Maxwell3D.singlelayer, Maxwell3D.doublelayer}
 * (MWSingleLayer3D / MWDoubleLayer3D, src/maxwell/mwops.jl:11-64) and
 * X, Y raviartthomas spaces (src/bases/rtspace.jl:3-105), plus the dense solve
 * behind it (LAPACK getrf/getrs via src/solvers/lusolver.jl:26-34).
 *
 * Conventions: plain pointers and sizes; 0-based indices; complex numbers as
 * interleaved (re, im) doubles; matrices in column-major order (the layout of
 * the reference's Matrix{ComplexF64}) unless a function says otherwise.
 * Every function returns BH_OK (0) or a negative error code; the message of
 * the last failure on the calling thread is available from bh_last_error().
 * Handles are not thread safe; distinct contexts may be used concurrently.
 */
#ifndef BEAST_HIP_H
#define BEAST_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BH_VERSION 100

enum bh_status {
    BH_OK = 0,
    BH_ERR_INVALID = -1,      /* bad argument (reference: ArgumentError / error()) */
    BH_ERR_OOM = -2,          /* device allocation failed */
    BH_ERR_HIP = -3,          /* HIP runtime error */
    BH_ERR_ASSERT = -4,       /* a condition the reference @asserts failed (e.g. hits > 3) */
    BH_ERR_UNSUPPORTED = -5,  /* valid input outside the GPU path (caller falls back) */
    BH_ERR_SINGULAR = -6,     /* exactly singular pivot in the LU (LAPACK info > 0) */
    BH_ERR_NOCONV = -7        /* iterative refinement did not reach the tolerance */
};

/* Operator kinds:
 *   BH_OP_MW_SINGLELAYER          MWSingleLayer3D (src/maxwell/mwops.jl:11-15,101-117)
 *   BH_OP_MW_DOUBLELAYER          MWDoubleLayer3D (mwops.jl:61-64,141-155)
 *   BH_OP_NCROSS                  NCross, <f, n x g> on the common cells of one mesh
 *                                 (src/identityop.jl:14-22); quadstrat BH_QS_SINGLENUM
 *   BH_OP_MW_DOUBLELAYER_ROTATED  DoubleLayerRotatedMW3D, f . (n x (grad G x g)) with n the
 *                                 test cell's normal (src/maxwell/nxdbllayer.jl:17-44); its
 *                                 default quadstrat is DoubleNumQStrat(6,7) (:23); it has no
 *                                 singularity extraction, so BH_QS_WILTON_SAUTER is
 *                                 BH_ERR_UNSUPPORTED (the reference has no regular part for it)
 *   BH_OP_IDENTITY                Identity, <f, g> on the common cells of one mesh
 *                                 (src/identityop.jl:6-11); quadstrat BH_QS_SINGLENUM
 * gamma/alpha/beta are unused by the local operators; like the reference's integrands,
 * the two double layers use gamma only. */
enum bh_op_kind {
    BH_OP_MW_SINGLELAYER = 0,
    BH_OP_MW_DOUBLELAYER = 1,
    BH_OP_NCROSS = 2,
    BH_OP_MW_DOUBLELAYER_ROTATED = 3,
    BH_OP_IDENTITY = 4
};

/* Quadrature strategy kinds; the integer parameters follow the field order of
 * the reference structs:
 *   BH_QS_WILTON_SAUTER  DoubleNumWiltonSauterQStrat(outer_far, inner_far,
 *                        outer_near, inner_near, ss_tetr, ss_face, ss_edge, ss_vert)
 *                        src/quadrature/doublenumwiltonsauterqstrat.jl:1-10
 *   BH_QS_SAUTER         DoubleNumSauterQstrat(outer, inner, ss_tetr, ss_face,
 *                        ss_edge, ss_vert)   src/quadrature/doublenumsauterqstrat.jl:1-8
 *   BH_QS_DOUBLENUM      DoubleNumQStrat(outer, inner)   src/quadrature/doublenumqstrat.jl:1-4
 *   BH_QS_SINGLENUM      SingleNumQStrat(quad_rule)   src/quadrature/quadstrats.jl:48-50
 *                        (local operators; RT x RT default 6, src/identityop.jl:66-67)
 * Triangle rule parameters are Dunavant polynomial degrees 1..10; Sauter-Schwab
 * parameters are Gauss-Legendre orders 1..16.  A degree or order beyond those (valid for
 * the reference, e.g. the rules 12 and 13 of test/test_sauterschwabints.jl:29-30) is
 * BH_ERR_UNSUPPORTED, so the caller falls back to its CPU path; one below 1 is
 * BH_ERR_INVALID. */
enum bh_qs_kind { BH_QS_WILTON_SAUTER = 0, BH_QS_SAUTER = 1, BH_QS_DOUBLENUM = 2, BH_QS_SINGLENUM = 3 };

/* Arithmetic.  Assembly: BH_FP64 everywhere, or BH_FP32 for the far-field
 * double-Gauss only (classification, Sauter-Schwab and Wilton blocks stay FP64
 * and are rounded when stored).  Factorisation: BH_FP64 (getrf), BH_FP32 (FP32
 * storage, FP32 MFMA products), or BH_FP32_SPLIT (FP32 storage and accumulation;
 * the trailing-update products use split bf16 pairs on the bf16 MFMA, about 16
 * significant bits per product); the FP32 modes are meant for GMRES-IR. */
enum bh_precision { BH_FP64 = 0, BH_FP32 = 1, BH_FP32_SPLIT = 2 };

typedef struct bh_ctx bh_ctx;
typedef struct bh_mesh bh_mesh;
typedef struct bh_space bh_space;
typedef struct bh_matrix bh_matrix;

/* Operator description: gamma, alpha, beta as in MWSingleLayer3D(gamma, alpha,
 * beta) (mwops.jl:11-15); MWDoubleLayer3D uses gamma only (its alpha is not
 * used by the kernel, mwops.jl:141-155). */
typedef struct {
    int32_t kind;          /* bh_op_kind */
    double gamma[2];
    double alpha[2];
    double beta[2];
} bh_operator;

typedef struct {
    int32_t kind;          /* bh_qs_kind */
    int32_t params[8];
} bh_quadstrat;

int bh_version(void);
const char *bh_last_error(void);

/* Context: one GPU and one stream.  Replaces the Threading{:multi} driver of
 * src/operator.jl:215-231 (the device is the parallel resource). */
int bh_ctx_create(int device, bh_ctx **out);
int bh_ctx_destroy(bh_ctx *ctx);
int bh_ctx_synchronize(bh_ctx *ctx);
/* Buffer pool (no counterpart in the reference): the device buffers of destroyed
 * resident matrices stay with the context for the next matrix of a similar size
 * (at least the size asked for, at most 5/4 of it), so that a repeated
 * assemble-factor-solve allocates nothing.  limit_bytes >= 0 sets the most the pool
 * keeps (default 64 GiB; 0 disables it and frees what it holds), a negative value
 * leaves it; *pooled_bytes (may be NULL) receives what the pool holds now.  A
 * device allocation that fails frees the pool and tries once more. */
int bh_ctx_pool(bh_ctx *ctx, int64_t limit_bytes, int64_t *pooled_bytes);
/* Measurement (no counterpart in the reference): while profiling is on, every
 * launch of the library's kernels is bracketed by HIP events on the context
 * stream; bh_ctx_kernel_stats returns, per kernel name (e.g. "k_cgemm"), the
 * launch count, summed device seconds and the algorithmic work (flops or
 * bytes, see DESIGN) the launches were given.  Enabling clears the counters. */
int bh_ctx_profile(bh_ctx *ctx, int enable);
int bh_ctx_kernel_stats(bh_ctx *ctx, const char *kernel, int64_t *launches, double *seconds, double *work);
/* Measurement (no counterpart in the reference): restrict profiling to the
 * launches of one kernel name (NULL or "" restores all), so that timing a
 * dominant kernel adds no events around the others. */
int bh_ctx_profile_only(bh_ctx *ctx, const char *kernel);

/* Triangle mesh: vertices[3*nv] (x,y,z per vertex), cells[3*nf] 0-based vertex
 * indices in chart order (CompScienceMeshes chart: r(u,v) = u v1 + v v2 + (1-u-v) v3). */
int bh_mesh_create(bh_ctx *ctx, const double *vertices, int64_t nv, const int32_t *cells, int64_t nf,
                   bh_mesh **out);
int bh_mesh_destroy(bh_mesh *mesh);

/* Raviart-Thomas space in the form of RTBasis.fns (rtspace.jl:3-7): the shapes
 * of dof b are entries shape_ptr[b] .. shape_ptr[b+1]-1 of (cell, refid, coeff),
 * refid 0-based (the local vertex opposite the edge).  From this the library
 * builds the element-to-dof map of assemblydata (src/bases/basis.jl:235-294). */
int bh_space_create(bh_mesh *mesh, int64_t ndofs, const int32_t *shape_ptr, const int32_t *shape_cell,
                    const int32_t *shape_ref, const double *shape_coeff, bh_space **out);
int bh_space_destroy(bh_space *space);
/* Element-to-dof map as built (dof[(cell*3+ref)*kmax + k], -1 if unused). */
int bh_space_e2d(const bh_space *space, int32_t *kmax, int32_t *dof, double *coeff, int64_t cap);

/* Dense assembly into a host matrix (drop-in for assemble/assemble!,
 * operator.jl:100-116,168-184): Z[row_off + m + ld*(col_off + n)] (+)= scale * A[m, n]
 * with m a test dof of X and n a trial dof of Y.  With accumulate = 0 the
 * target block is overwritten, with 1 it is added to (the store(v,m,n)
 * semantics of assemble!).  Z is complex double (precision = BH_FP64) or complex
 * float (precision = BH_FP32), column-major with leading dimension ld. */
int bh_assemble(bh_ctx *ctx, const bh_operator *op, const bh_quadstrat *qs, int precision,
                const bh_space *X, const bh_space *Y, const double *scale, int accumulate, void *Z,
                int64_t ld, int64_t row_off, int64_t col_off);

/* Whether (op, qs) has a GPU path, without a context or a device: BH_OK, BH_ERR_UNSUPPORTED
 * (a valid combination the caller's CPU path must take: a rule degree beyond the tables, a
 * strategy the operator's reference kernel does not define on the GPU) or BH_ERR_INVALID. */
int bh_check_strategy(const bh_operator *op, const bh_quadstrat *qs);

/* Pair classes (0 far, 1 Wilton near, 2 common vertex, 3 common edge, 4 common
 * face) for all test x trial cell pairs, row-major [nf_test][nf_trial];
 * the classification of doublenumwiltonsauterqstrat.jl:32-69. */
int bh_classify(bh_ctx *ctx, const bh_operator *op, const bh_quadstrat *qs, const bh_mesh *test,
                const bh_mesh *trial, int8_t *classes);

/* Local 3x3 blocks of a list of cell pairs (column-major 3x3 complex each, i.e.
 * 18 doubles per pair), computed by the device kernels of the given class path;
 * debug/inspection entry for per-regime parity tests. */
int bh_local_blocks(bh_ctx *ctx, const bh_operator *op, const bh_quadstrat *qs, const bh_mesh *test,
                    const bh_mesh *trial, int64_t npairs, const int32_t *test_cells,
                    const int32_t *trial_cells, double *blocks);

/* Plane-wave excitation e = (n x E) x n tested with X (excitation.jl:17-115,
 * mwexc.jl:36-42): E(x) = amplitude * exp(-gamma d.x) * pol, rule = triangle
 * degree (SingleNumQStrat(8) by default).  b[m] (+)= sum. */
int bh_assemble_planewave(bh_ctx *ctx, const bh_space *X, const double *direction, const double *pol_re,
                          const double *pol_im, const double *gamma, const double *amplitude, int rule,
                          int accumulate, double *b);
/* The plane wave through either trace of the reference (src/excitation.jl:145-180):
 * trace BH_TRACE_TANGENTIAL is (n x E) x n (TangTraceMW, as bh_assemble_planewave),
 * BH_TRACE_CROSS is n x E (CrossTraceMW), n the cell's normal; e.g. the right-hand side
 * n x H of the MFIE (examples/nxmfie.jl:14-22) with H = -1/(i mu omega) curl E, a plane
 * wave itself (curl, src/maxwell/mwexc.jl:44-52). */
#define BH_TRACE_TANGENTIAL 0
#define BH_TRACE_CROSS 1
int bh_assemble_excitation(bh_ctx *ctx, const bh_space *X, int trace, const double *direction, const double *pol_re,
                           const double *pol_im, const double *gamma, const double *amplitude, int rule,
                           int accumulate, double *b);

/* Far field of a surface current (potential(op, points, coeffs, X), src/postproc.jl:138-150,174-211,
 * with the far-field operators of src/maxwell/farfield.jl:1-120): for each point y (a direction,
 * points[3*p..3*p+2], used as given),
 *     ff[p] = amplitude * sum over cells and rule points x of w I(y, K(y, x) J(x)),
 * J(x) = sum_n coeffs[n] f_n(x) (coeffs: ndofs complex), K = exp(gamma y.x), and
 *     BH_FF_SINGLELAYER          I = (y x KJ) x y           MWFarField3D (:61-63)
 *     BH_FF_DOUBLELAYER          I =  y x KJ                MWDoubleLayerFarField3D (:65-67)
 *     BH_FF_DOUBLELAYER_ROTATED  I =  y x (KJ x n)          MWDoubleLayerRotatedFarField3D (:98-100)
 *     BH_FF_DROPCONSTANT         I = (y x KJ) x y, K = expm1(gamma y.x)   MWFarField3DDropConstant (:69-75)
 * rule = triangle rule degree 1..10 (the default SingleNumQStrat(10), src/postproc/farfield.jl:3).
 * ff: 3 complex per point (6 doubles), overwritten.  FP64 throughout. */
#define BH_FF_SINGLELAYER 0
#define BH_FF_DOUBLELAYER 1
#define BH_FF_DOUBLELAYER_ROTATED 2
#define BH_FF_DROPCONSTANT 3
int bh_farfield(bh_ctx *ctx, int kind, const double *gamma, const double *amplitude, const bh_space *X,
                const double *coeffs, int64_t npoints, const double *points, int rule, double *ff);

/* ---------------------------------------------------------------- resident matrices
 * A square system matrix kept in HBM with its LU factors (no counterpart in the
 * reference: its Matrix{ComplexF64} lives on the host and is factored by
 * LAPACK, lusolver.jl:26-34 / BEAST.lu, lu.jl:29-47). */
typedef struct {
    int32_t iterations;     /* refinement iterations (outer) */
    int32_t inner;          /* total inner (GMRES) iterations */
    double residual;        /* final ||b - A x|| / ||b|| in FP64, max over rhs */
    double seconds_factor;  /* device time of the last factorisation */
    double seconds_solve;   /* device time of the solve */
} bh_solve_stats;

/* Assemble op on (X, X) into a resident FP64 matrix (row-major in HBM).  The same
 * launch also writes the matrix rounded to FP32 (8 N^2 further bytes, held until the
 * matrix is factored: an FP32 factorisation starts from it, an FP64 one frees it). */
int bh_matrix_assemble(bh_ctx *ctx, const bh_operator *op, const bh_quadstrat *qs, const bh_space *X,
                       bh_matrix **out);
/* Resident matrices of several requests assembled in shared launches (no counterpart in
 * the reference, which assembles one operator per call, src/operator.jl:100-116): out[i]
 * = ops[i] on (spaces[i], spaces[i]) under qs[i], for integral operators (local operators
 * are added with bh_matrix_accumulate).  The classification, the Sauter-Schwab and Wilton
 * blocks and the far-field tiles of all requests run in one launch per kernel
 * (configuration), so that requests too small to fill the GPU share it; every matrix (and
 * its FP32 copy) is bit-identical to the one bh_matrix_assemble gives the request alone.
 * On failure no matrix is returned. */
int bh_matrix_assemble_batch(bh_ctx *ctx, int64_t count, const bh_operator *ops, const bh_quadstrat *qs,
                             const bh_space *const *spaces, bh_matrix **out);
/* Z += scale * op(X, X) on a resident FP64 matrix (before it is factored); the
 * resident form of a LinearCombinationOfOperators term (operator.jl:294-306). */
int bh_matrix_accumulate(bh_matrix *M, const bh_operator *op, const bh_quadstrat *qs, const bh_space *X,
                         const double *scale);
/* Resident matrix from host data (column-major complex double, n x n). */
int bh_matrix_from_host(bh_ctx *ctx, int64_t n, const double *A, int64_t ld, bh_matrix **out);
int bh_matrix_destroy(bh_matrix *M);
int bh_matrix_size(const bh_matrix *M, int64_t *n);
int bh_matrix_download(bh_matrix *M, double *A, int64_t ld);
/* LU of the resident matrix (getrf semantics: sequential row interchanges, unit
 * lower L).  BH_FP64 factors the FP64 matrix with LAPACK's partial pivoting (the
 * same ipiv as zgetrf).  BH_FP32 / BH_FP32_SPLIT factor a rounded FP32 copy; their
 * panels choose pivot rows by tournament pivoting, so ipiv may differ from
 * zgetrf's.  Calls on one matrix are serialised with its solves.  The FP32
 * factorisation's workspace (about 1 GB at N = 40k) stays allocated in the context
 * for the next factorisation until the context is destroyed. */
int bh_matrix_factor(bh_matrix *M, int precision);
/* Several resident matrices (all of ctx) factored in shared launches (no counterpart in the
 * reference, which factors one matrix per `\` or LU call, src/solvers/lusolver.jl:26-34):
 * with BH_FP32_SPLIT, the matrices of up to 24,576 rows follow the schedule of
 * bh_matrix_factor with every launch issued once for all matrices that have it at that
 * step, so a batch of small matrices fills the GPU; each gets the factors and pivots
 * bh_matrix_factor gives it alone.  Larger matrices and the other precisions are factored
 * one by one.  Each matrix records the batch's factorisation time.  A singular matrix is
 * left unfactored and reported (BH_ERR_SINGULAR) after the others are done. */
int bh_matrix_factor_batch(bh_ctx *ctx, int64_t count, bh_matrix *const *mats, int precision);
/* Solve A X = B for nrhs column-major right-hand sides (ld = n).  With FP64
 * factors: getrs.  With FP32 factors: FP64-residual iterative refinement with
 * GMRES (GMRES-IR), stopped at ||b - A x|| / ||b|| <= tol or maxit outer steps. */
int bh_matrix_solve(bh_matrix *M, int64_t nrhs, const double *B, double *X, double tol, int maxit,
                    bh_solve_stats *stats);
/* bh_matrix_solve without device allocations (no counterpart in the reference: for callers
 * that cannot allocate device memory, e.g. code under HIP graph capture or a tight serving
 * loop of repeated right-hand sides on resident factors).  Same arguments, results and
 * errors as bh_matrix_solve; its workspace is a fixed 64 MiB scratch array per device instead
 * of a fresh allocation, which limits 67 * n * nrhs * 16 bytes (plus alignment and 0.13 MB
 * per right-hand side) to 64 MiB, about n * nrhs <= 62,000; larger solves are refused
 * (BH_ERR_INVALID).
 * The factors' own solve buffers are allocated by bh_matrix_factor.  All calls on one device
 * share the array: calls that overlap in time on one GPU race (serialise them). */
int bh_matrix_solve_noalloc(bh_matrix *M, int64_t nrhs, const double *B, double *X, double tol, int maxit,
                            bh_solve_stats *stats);
/* Pivot vector of the last factorisation (1-based like LAPACK ipiv). */
int bh_matrix_pivots(bh_matrix *M, int64_t *ipiv);

#ifdef __cplusplus
}
#endif

#endif /* BEAST_HIP_H */
