/*
 * hipfem.h -- C ABI of the MI355X (gfx950) assembly-and-solve path.
 *
 * Three drop-in points of the reference are served:
 *   hipcsr_*  a SparseMtrx (src/core/sparsemtrx.h:60-281), registered in the
 *             reference as SMT_HipCSR through REGISTER_SparseMtrx
 *             (src/core/classfactory.h:149);
 *   hippcg_*  a SparseLinearSystemNM (src/core/sparselinsystemnm.h:56-97),
 *             registered as ST_HipPCG through REGISTER_SparseLinSolver
 *             (src/core/classfactory.h:150);
 *   hfem_model_*  the batched element/material evaluator that replaces the
 *             element loops of EngngModel::assemble (src/core/engngm.C:889-1037)
 *             and EngngModel::assembleVectorFromElements (engngm.C:1351-1535)
 *             for LSpace / LTRSpace / QSpace with IsoLE / MisesMat.
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative HIPFEM_E* code on
 *    failure; hipfem_last_error() gives the message (thread-local).  The
 *    reference raises OOFEM_ERROR (RuntimeException, src/core/error.h:79) on
 *    the same conditions; its C++ wrapper maps a non-zero status to that.
 *  - Equation numbers at the boundary are 1-based with 0 = "no equation",
 *    exactly like the reference's location arrays (element.C:429-455).
 *  - Dense matrices are row-major.  Host vectors are plain double arrays;
 *    the *_dev variants take device pointers that live on the context's GPU.
 *  - One context = one GPU + one HIP stream.  Handles are not re-entrant: a
 *    handle is used from one host thread at a time, as the reference does.
 */
#ifndef HIPFEM_H
#define HIPFEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HIPFEM_OK            0
#define HIPFEM_EINVAL      (-1)   /* bad argument / state */
#define HIPFEM_EHIP        (-2)   /* HIP runtime error */
#define HIPFEM_ENOMEM      (-3)   /* allocation failed */
#define HIPFEM_ERANGE      (-4)   /* index outside the matrix / pattern */
#define HIPFEM_ENUMERIC    (-5)   /* zero pivot / non-positive volume / NaN */
#define HIPFEM_ETOOBIG     (-6)   /* size beyond the int32 index range */

/* ConvergedReason values of the reference (src/core/convergedreason.h:41-50) */
#define HIPFEM_CR_CONVERGED      1
#define HIPFEM_CR_DIVERGED_ITS (-1)

/* element types (LSpace, LTRSpace, QSpace) and material models */
#define HIPFEM_LSPACE   1
#define HIPFEM_LTRSPACE 2
#define HIPFEM_QSPACE   3
#define HIPFEM_MAT_ISOLE 0
#define HIPFEM_MAT_MISES 1
#define HIPFEM_MATP      8  /* material row: [model, E, nu, sig0, H, 0, 0, 0] */

typedef struct hipfem_ctx    hipfem_ctx;
typedef struct hipcsr_mtx    hipcsr_mtx;
typedef struct hippcg_solver hippcg_solver;
typedef struct hfem_model    hfem_model;

/* ------------------------------------------------------------------ context */
int         hipfem_ctx_create(int device, hipfem_ctx **out);
void        hipfem_ctx_destroy(hipfem_ctx *ctx);
int         hipfem_ctx_synchronize(hipfem_ctx *ctx);
/* the HIP stream (hipStream_t) every launch of this context goes to */
void       *hipfem_ctx_stream(hipfem_ctx *ctx);
const char *hipfem_last_error(void);
const char *hipfem_version(void);

/* Device vectors of FP64 on the context's GPU (plumbing for callers that keep
 * their vectors resident; no counterpart in the reference). */
int hipfem_dev_alloc(hipfem_ctx *ctx, int64_t n, double **out);
int hipfem_dev_free(hipfem_ctx *ctx, double *p);
int hipfem_dev_upload(hipfem_ctx *ctx, double *dst, const double *src, int64_t n);
int hipfem_dev_download(hipfem_ctx *ctx, double *dst, const double *src, int64_t n);
/* y = a*x + b*y on the context stream (x may be NULL: y = b*y) */
int hipfem_dev_axpby(hipfem_ctx *ctx, int64_t n, double a, const double *x, double b, double *y);
/* y = x on the context stream: method 0 hipMemcpyAsync, 1 a 16-byte-wide
 * copy kernel (bandwidth reference for the counters; even n, aligned) */
int hipfem_dev_copy(hipfem_ctx *ctx, int64_t n, const double *x, double *y, int method);
/* dot product x.y and 2-norm, FP64, fixed summation order */
int hipfem_dev_dot(hipfem_ctx *ctx, int64_t n, const double *x, const double *y, double *out);
/* y[i incy] = a x[i incx] + b y[i incy], i < n (x may be NULL: y = b y);
 * gathers and scatters one column of interleaved multi-column vectors */
int hipfem_dev_axpby_strided(hipfem_ctx *ctx, int64_t n, double a, const double *x, int64_t incx, double b,
                             double *y, int64_t incy);
/* max |x_i| (NRSolver's increment cap, nrsolver.C:340-353) */
int hipfem_dev_amax(hipfem_ctx *ctx, int64_t n, const double *x, double *out);

/* ------------------------------------------------------------------- hipcsr */
/* SparseMtrx constructor / clone()  (sparsemtrx.h:95-99) */
int  hipcsr_create(hipfem_ctx *ctx, hipcsr_mtx **out);
void hipcsr_destroy(hipcsr_mtx *m);
int  hipcsr_clone(const hipcsr_mtx *m, hipcsr_mtx **out);

/* buildInternalStructure(EngngModel*, di, s) (sparsemtrx.h:198; compcol.C:174-272):
 * the pattern is the union of loc x loc over the element location arrays the
 * caller passes (nelem arrays of ndofe 1-based equations, 0 = prescribed). */
int hipcsr_build_structure(hipcsr_mtx *m, int neq, int nelem, int ndofe, const int *loc);
/* buildInternalStructure(EngngModel*, n, m, I, J) (sparsemtrx.h:192): pattern from
 * a full CSR description (0-based sorted columns). */
int hipcsr_build_structure_csr(hipcsr_mtx *m, int n, const int64_t *rowptr, const int32_t *colind);

/* assemble(loc, mat) / assemble(rloc, cloc, mat) (sparsemtrx.h:220,231):
 * add mat(i,j) at (loc[i], loc[j]) for loc > 0.  Contributions are staged and
 * reach the device at assembleEnd(). */
int hipcsr_assemble(hipcsr_mtx *m, int n, const int *loc, const double *mat);
int hipcsr_assemble_rc(hipcsr_mtx *m, int nr, const int *rloc, int nc, const int *cloc, const double *mat);
int hipcsr_assemble_begin(hipcsr_mtx *m);   /* sparsemtrx.h:234 */
int hipcsr_assemble_end(hipcsr_mtx *m);     /* sparsemtrx.h:236 */

int hipcsr_zero(hipcsr_mtx *m);                          /* zero()        :254 */
int hipcsr_scale(hipcsr_mtx *m, double a);               /* times(double)  :155 */
int hipcsr_add(hipcsr_mtx *m, double a, const hipcsr_mtx *other); /* add(x, m) :157 (same pattern) */
int hipcsr_add_diagonal(hipcsr_mtx *m, double a, const double *d);  /* addDiagonal :160 */

/* times(x, answer) / timesT (sparsemtrx.h:137,143): y = A x, y = A^T x */
int hipcsr_times(hipcsr_mtx *m, const double *x, double *y);
int hipcsr_times_t(hipcsr_mtx *m, const double *x, double *y);
int hipcsr_times_dev(hipcsr_mtx *m, const double *dx, double *dy);
/* times(FloatMatrix B, answer) (sparsemtrx.h:149): B is n x ncol row-major */
int hipcsr_times_mat(hipcsr_mtx *m, int ncol, const double *B, double *Y);

/* at(i,j) (sparsemtrx.h:265-267), 1-based; a missing entry reads as 0 and
 * cannot be written (HIPFEM_ERANGE), like CompCol::at. */
int hipcsr_get(hipcsr_mtx *m, int i, int j, double *v);
int hipcsr_set(hipcsr_mtx *m, int i, int j, double v);
int hipcsr_is_allocated(hipcsr_mtx *m, int i, int j);    /* isAllocatedAt :270 */
/* position of entry (i, j) (1-based) in the value array of hipcsr_download,
 * -1 when it is not in the pattern (the monotone search of CompCol::at,
 * compcol.C:339) */
int hipcsr_find(const hipcsr_mtx *m, int i, int j, int64_t *pos);

int     hipcsr_nrows(const hipcsr_mtx *m);
int64_t hipcsr_nnz(const hipcsr_mtx *m);
long    hipcsr_version(const hipcsr_mtx *m);             /* giveVersion()  :82 */
/* copies of the pattern / values for inspection (host arrays sized n+1, nnz) */
int hipcsr_download(hipcsr_mtx *m, int64_t *rowptr, int32_t *colind, double *val);
int hipcsr_upload_values(hipcsr_mtx *m, const double *val);
/* SpMV kernel shape (no counterpart in the reference): lanes per row (8, 16,
 * 32), loads in flight per lane and pass (1, 2, 4, 6), non-temporal value and
 * column streams (0/1).  A default is chosen from the mean row length when
 * the pattern is built. */
int hipcsr_set_spmv_shape(hipcsr_mtx *m, int lanes, int unroll, int nontemporal);
int hipcsr_get_spmv_shape(const hipcsr_mtx *m, int *lanes, int *unroll, int *nontemporal);
/* streaming SpMV form: rows_per_block (8, 16, 32, 64; 0 = back to the lane
 * form) consecutive rows per 256-thread block, their contiguous value and
 * column range read with unit stride; unroll 2, 4 or 6. */
int hipcsr_set_spmv_stream(hipcsr_mtx *m, int rows_per_block, int unroll);
/* node-block SpMV forms over the same CSR arrays: the 3x3 block form when
 * every free node is fully free and numbered consecutively (then CSR order
 * is a 3x3 block layout), else the row-group form when rows of a node share
 * one column list (partially constrained nodes); on by default where one
 * applies.  hipcsr_is_nodeblocked: 1 block form, 2 row-group form, 0 none. */
int hipcsr_set_spmv_nodeblock(hipcsr_mtx *m, int on);
int hipcsr_is_nodeblocked(const hipcsr_mtx *m);
/* node rows (block form) or row groups, and 3x3 blocks or column chunks */
int hipcsr_nodeblock_sizes(const hipcsr_mtx *m, int64_t *groups, int64_t *blocks);
/* element-by-element operator (no counterpart in the reference): after
 * hfem_model_assemble_stiffness on an LSpace mesh the matrix remembers the
 * evaluator, and while its values stay unchanged the FP32 PCG of precision
 * 2 applies A as sum_e K_e x_e recomputed in FP32 instead of reading the
 * stored values; on by default, off falls back to the FP32 copy. */
int hipcsr_set_matrix_free(hipcsr_mtx *m, int on);
int hipcsr_is_matrix_free(const hipcsr_mtx *m);
/* y = A x through the element-by-element FP32 operator (host arrays) */
int hipcsr_times_matrix_free(hipcsr_mtx *m, const double *x, double *y);
/* device pointers of the CSR arrays (row pointers and columns are int32) */
int hipcsr_device_arrays(hipcsr_mtx *m, const int32_t **rowptr, const int32_t **colind, double **val);

/* ------------------------------------------------------------------- hippcg */
/* IMLSolver-compatible options (imlsolver.C:67-101): lstol (default 1e-5),
 * lsiter (default 200), lsprecond (0 void, 1 diagonal). */
int  hippcg_create(hipfem_ctx *ctx, hippcg_solver **out);
void hippcg_destroy(hippcg_solver *s);
int  hippcg_set_options(hippcg_solver *s, double lstol, int lsiter, int lsprecond);
/* precision: 0 = FP64 PCG; 1 = FP32 inner PCG inside FP64 iterative
 * refinement (true residual, update and dot products in FP64), inner_tol:
 * relative reduction asked of each inner solve; 2 = one PCG on FP32
 * vectors and matrix with FP64 solution and dot accumulation, the
 * recursive residual replaced by the FP64 true residual b - A x whenever it
 * has dropped by the reliable-update factor (hippcg_set_reliable, default
 * 0.1) and convergence decided on the true residual. */
int  hippcg_set_precision(hippcg_solver *s, int precision, double inner_tol, int max_outer);
int  hippcg_set_reliable(hippcg_solver *s, double delta);
/* 1 when the last precision-2 solve applied A element by element (A was the
 * unchanged result of hfem_model_assemble_stiffness on LSpace), 0 when it
 * used the FP32 copy of the stored values */
int  hippcg_last_matrix_free(hippcg_solver *s);
/* solve(A, b, x) (sparselinsystemnm.h:81; imlsolver.C:102-146): x holds the
 * initial guess on entry.  *reason receives HIPFEM_CR_CONVERGED or
 * HIPFEM_CR_DIVERGED_ITS; *iters the (inner) iterations; *achieved the final
 * relative residual. */
int hippcg_solve(hippcg_solver *s, hipcsr_mtx *A, const double *b, double *x,
                 int *reason, int *iters, double *achieved);
int hippcg_solve_dev(hippcg_solver *s, hipcsr_mtx *A, const double *db, double *dx,
                     int *reason, int *iters, double *achieved);
/* solve(A, B, X) for nrhs right-hand sides (sparselinsystemnm.C:47-70; B
 * and X n x nrhs, column-major, host; X holds the initial guesses): column
 * by column, stopping at the first that does not converge.  With
 * precision 2, the diagonal preconditioner and zero initial guesses the
 * columns go six at a time through the batched block solver (hfem_batch,
 * one block), which reads each matrix entry once per iteration for six
 * columns.  *iters and *achieved are the largest over the columns. */
int hippcg_solve_multi(hippcg_solver *s, hipcsr_mtx *A, int nrhs, const double *B, double *X,
                       int *reason, int *iters, double *achieved);
/* outer iterations (precision 1) or reliable updates (precision 2) of the
 * last mixed-precision solve */
int hippcg_last_outer(hippcg_solver *s);
/* Instrumentation (no counterpart in the reference): when on, every PCG
 * SpMV launch of a solve is bracketed by HIP events on the context stream;
 * hippcg_profile returns the number of timed launches of the last solve
 * (iterations up to convergence) and their summed duration in ms, and
 * the FP32 inner launches separately. */
int hippcg_set_profiling(hippcg_solver *s, int on);
/* Iteration layout (no counterpart in the reference; same arithmetic):
 * fused 1 forms p = z + beta p inside the SpMV gather (lane, node-block and
 * row-group forms of the plain PCG), fused 0 (default) updates p in place in
 * its own pass and gathers p once per nonzero. */
int hippcg_set_fused(hippcg_solver *s, int fused);
int hippcg_profile(hippcg_solver *s, int *n_spmv64, double *ms_spmv64, int *n_spmv32, double *ms_spmv32);
/* Allocation-free plain PCG.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same contract and
 * result as hippcg_solve_dev with precision 0 (the same kernels in the same
 * order), but the diagonal inverse, the work vectors, the partial sums and
 * the solver state live in one fixed device array per GPU (about 48 MiB,
 * kept until the process ends) instead of the solver's own buffers, so the
 * call allocates nothing.  Limits: precision 0 only, at most
 * HIPPCG_NOALLOC_MAX_N equations (HIPFEM_ETOOBIG beyond), no residual
 * history and no profiling.  Every call on one GPU shares that array:
 * calls that overlap on the same GPU race. */
#define HIPPCG_NOALLOC_MAX_N 1048576
int hippcg_solve_dev_noalloc(hippcg_solver *s, hipcsr_mtx *A, const double *db, double *dx, int *reason, int *iters,
                             double *achieved);
/* Residual history of the plain PCG (precision 0, and the inner solves of
 * precision 1): with a capacity cap > 0 every solve records ||r_k|| / ||b||
 * for k = 0, 1, ... (the quantity cg.h compares with lstol, iml/cg.h:34-60,
 * as IMLSolver::solve runs it, imlsolver.C:126-129); hippcg_get_history
 * copies the last solve's min(cap, iterations + 1) values and gives their
 * count in *len.  The reference logs only the final value; the history is
 * the observable its CG template is compared on. */
int hippcg_set_history(hippcg_solver *s, int cap);
int hippcg_get_history(hippcg_solver *s, double *out, int cap, int *len);

/* --------------------------------------------------- batched evaluator */
int  hfem_model_create(hipfem_ctx *ctx, hfem_model **out);
void hfem_model_destroy(hfem_model *md);
/* mesh in the reference's local node order; conn is nelem x nen (0-based);
 * emat indexes the material table mat (nmat x HIPFEM_MATP); eqn is
 * nnode x 3 with >0 free equation, <0 minus the prescribed equation, 0 none
 * (forceEquationNumbering engngm.C:469-531). */
int hfem_model_set_mesh(hfem_model *md, int nnode, const double *coords, int etype, int nelem,
                        const int *conn, const int *emat, int nmat, const double *mat,
                        const int *eqn, int neq, int npeq);
/* pattern of the free-free stiffness from the model's location arrays
 * (same result as hipcsr_build_structure over every element), plus the
 * element->pattern offset map and the element coloring used by assembly. */
int hfem_model_build_structure(hfem_model *md, hipcsr_mtx *K);
/* K = sum_e K_e (TangentAssembler assemblercallback.C:246 -> NLStructuralElement::
 * computeStiffnessMatrix nlstructuralelement.C:332-448).  mode 0: elastic
 * stiffness; mode 1: consistent tangent from the temp GP state (MisesMat
 * give3dMaterialStiffnessMatrix misesmat.C:496-544). */
int hfem_model_assemble_stiffness(hfem_model *md, hipcsr_mtx *K, int mode);
/* f = sum_e B^T sigma dV (NLStructuralElement::giveInternalForcesVector
 * nlstructuralelement.C:151-241), MisesMat return mapping per GP
 * (misesmat.C:181-263) writing the temp state when update_state != 0.
 * f_free (neq), f_presc (npeq, may be NULL), ebe3 (3 doubles, may be NULL):
 * sum of squared element force components per dof id (u, v, w). */
int hfem_model_internal_forces(hfem_model *md, const double *u_free, const double *u_presc,
                               double *f_free, double *f_presc, double *ebe3, int update_state);
int hfem_model_internal_forces_dev(hfem_model *md, const double *du_free, const double *du_presc,
                                   double *df_free, double *df_presc, double *ebe3, int update_state);
/* y = sum_e K_e x_e on device vectors (EngngModel::assembleExtrapolatedForces
 * engngm.C:1536-1610, element matrix times element vector, assembled by
 * equation number), evaluated as sum_g B^T D B x_e dV without forming K_e.
 * mode 0: elastic D; 1: TangentStiffness D of the temp Gauss-point state
 * (as hfem_model_assemble_stiffness).  dx_free / dx_presc may be NULL
 * (zero); dy_free (neq) and dy_presc (npeq) are overwritten, either may be
 * NULL.  The state is not modified. */
int hfem_model_tangent_product_dev(hfem_model *md, int mode, const double *dx_free, const double *dx_presc,
                                   double *dy_free, double *dy_presc);
/* Newton residual on device vectors: rhs = RT - F (RT may be NULL = 0) and
 * sums6 = {sum rhs^2 over u, v, w equations, sum RT^2 over u, v, w}, the
 * per-dof-id quantities of NRSolver::checkConvergence (nrsolver.C:648-660). */
int hfem_model_residual(hfem_model *md, const double *dRT, const double *dF, double *drhs, double *sums6);
/* per dof id squared sums of two device vectors over the free equations:
 * sums6 = {sum x^2 (u, v, w), sum y^2 (u, v, w)} -- the dg_dispErr /
 * dg_totalDisp sums of NRSolver::checkConvergence (nrsolver.C:648-660) */
int hfem_model_group_sums(hfem_model *md, const double *dx, const double *dy, double *sums6);
/* MisesMatStatus::updateYourself (misesmat.C:720-728) for every Gauss point */
int hfem_model_commit(hfem_model *md);
/* form of the FP32 element-by-element operator used by hippcg precision 2
 * (no counterpart in the reference): 0 = one launch per element color
 * adding into the FP64 result; 1 = every element in one launch, its forces
 * stored and summed per node in FP64; 2 (default) = every element in one
 * launch, each force stored in its color's FP32 slot vector and summed per
 * equation in color order (the result of form 0 bit for bit; measured on
 * MI355X for 1M hexahedra: 0.248 ms, against 0.255 for form 0 and 0.34 for
 * form 1) */
int hfem_model_set_element_operator(hfem_model *md, int form);
/* Element::initForNewStep for every element (element.C:878-887, called by
 * EngngModel::initStepIncrements before a restarted step): each Gauss
 * point's temp values back to its committed ones */
int hfem_model_restart(hfem_model *md);
/* reset every Gauss point to the virgin state */
int hfem_model_reset_state(hfem_model *md);
/* Gauss point state, ngp_total x 38 doubles in the oracle's layout */
int hfem_model_get_state(hfem_model *md, double *state);
/* the inverse of hfem_model_get_state: every Gauss point's 38 fields
 * (a restart from a saved state, as the reference's status objects are
 * restored from a context file) */
int hfem_model_set_state(hfem_model *md, const double *state);
int hfem_model_ngp_total(const hfem_model *md);
int hfem_model_ncolors(const hfem_model *md);

/* ---------------------------------------------------- batched block solves */
/* Many independent SPD systems in one persistent launch (the RVE batch; no
 * single counterpart in the reference, whose RVE solves go one by one
 * through PrescribedGradient::computeField, prescribedgradient.C:161-181).
 * A is block diagonal: block b is rows [rows[b], rows[b+1]) and no entry
 * leaves its block (checked).  Six right-hand sides per block ([row][6]
 * interleaved, FP64 device arrays).  Jacobi PCG on FP32 values/vectors with
 * FP64 solution and dots and FP64 reliable updates (hippcg precision 2).
 * Blocks with more stored entries than the split are solved first, all
 * together, by a bulk-synchronous phase that spreads each block's rows
 * over many workgroups; the rest by one persistent launch, one workgroup
 * per block, blocks taken largest first. */
typedef struct hfem_batch hfem_batch;
int  hfem_batch_create(hipcsr_mtx *A, int nblocks, const int32_t *rows, hfem_batch **out);
void hfem_batch_destroy(hfem_batch *bt);
int  hfem_batch_set_options(hfem_batch *bt, double tol, double delta, int maxit);
/* kernel shapes (workgroup size, lanes per row) and the one to use */
int  hfem_batch_nshapes(void);
int  hfem_batch_shape_info(int shape, int *block_size, int *lanes_per_row);
int  hfem_batch_set_shape(hfem_batch *bt, int shape, int grid);
/* per block: iterations, reliable updates, final max relative true
 * residual of its six columns (negative: not converged within maxit) */
int  hfem_batch_solve_dev(hfem_batch *bt, const double *db, double *dx, int *iters, int *updates, double *resid);
/* time every persistent-kernel shape on `sample` median blocks at or below
 * the split (40 iterations each), keep the fastest */
int  hfem_batch_autotune(hfem_batch *bt, const double *db, double *dx, int sample, double *ms, int *best);
/* bulk-phase shapes (lanes per row of its SpMV) and the one to use */
int  hfem_batch_nbulk_shapes(void);
int  hfem_batch_bulk_shape_info(int shape, int *lanes_per_row);
int  hfem_batch_set_bulk_shape(hfem_batch *bt, int shape);
/* bulk products: 1 (the default) the 3x3 node-block form over the CSR
 * values, 2 the same over a block-major copy of the FP32 values padded to
 * 12 per block (three 16-byte loads per block), 0 CSR rows; the node-block
 * forms apply when the pattern is node-blocked and every block starts on a
 * node.  hfem_batch_bulk_form reports the form in use (0, 1 or 2) */
int  hfem_batch_set_bulk_form(hfem_batch *bt, int form);
int  hfem_batch_bulk_form(const hfem_batch *bt);
/* split: blocks with more stored entries go to the bulk phase; -1 (the
 * default) = the batch's entries divided by the workgroups the persistent
 * kernel keeps resident; 0 = every block */
int  hfem_batch_set_split(hfem_batch *bt, long long split);
/* HIP-event timing of the bulk-phase products on the batch's stream (on
 * resets the counters): number of timed products, their summed duration
 * (ms) and their summed algorithmic bytes */
int  hfem_batch_set_profiling(hfem_batch *bt, int on);
int  hfem_batch_profile(const hfem_batch *bt, long long *launches, double *ms, double *bytes);
int  hfem_batch_partition_info(hfem_batch *bt, int *nbulk, long long *split_used);
/* time every bulk shape on the blocks above the split (40 iterations
 * each), keep the fastest */
int  hfem_batch_autotune_bulk(hfem_batch *bt, const double *db, double *dx, double *ms, int *best);

/* ------------------------------------------------------ MisesMat surrogate */
/* A learned stand-in for MisesMat's return mapping and consistent tangent
 * (misesmat.C:181-263, 496-544); no counterpart in the reference.  Network
 * 13 -> 3 x 128 (SiLU) -> 27, FP32 master weights (flat, hfem_mlp_nparams()
 * values: W1 128x13, b1, W2 128x128, b2, W3 128x128, b3, W4 27x128, b4).
 * Inputs (eps - epsp_n, epsp_n, kappa_n) scaled by E/sig0; outputs the
 * stress and the upper triangle of the tangent (21, row-major) as their
 * departures from the elastic trial, (D_e (eps - epsp_n) - sigma) / sig0 and
 * (D_e - D) / E (hfem_mlp_scales: input factor E/sig0, 1/sig0, 1/E).
 * mat4 = {E, nu, sig0, H}. */
typedef struct hfem_mlp hfem_mlp;
int  hfem_mlp_nparams(void);
int  hfem_mlp_create(hipfem_ctx *ctx, const double *mat4, uint64_t seed, hfem_mlp **out);
void hfem_mlp_destroy(hfem_mlp *m);
int  hfem_mlp_scales(const hfem_mlp *m, double *scales3);
int  hfem_mlp_get_params(hfem_mlp *m, float *params);
int  hfem_mlp_set_params(hfem_mlp *m, const float *params);
/* training set on the device: random strain paths (1..maxsteps Gaussian
 * increments of standard deviation stdev per Voigt component, from the
 * virgin state) through the plain return mapping; one sample per step */
int  hfem_mlp_generate(hfem_mlp *m, int64_t nsamples, uint64_t seed, double stdev, int maxsteps);
int  hfem_mlp_get_samples(hfem_mlp *m, int64_t first, int64_t count, float *X, float *Y);
/* Adam on the mean-squared error over batches drawn with replacement,
 * learning rate decayed from lr to lr_final (cosine); *final_loss: loss of
 * the last batch */
int  hfem_mlp_train(hfem_mlp *m, int steps, int batch, double lr, double lr_final, uint64_t seed,
                    double *final_loss);
/* Use the network inside MisesMat's update (no counterpart in the
 * reference): at every internal-force evaluation the Gauss points whose
 * trial state lies more than band * sigma_y(kappa_n) outside the yield
 * surface are inferred in one BF16 batch; a prediction is accepted when the
 * return-mapping equations hold at it within tol * sigma_y (elastic
 * volumetric part, deviator along the trial deviator, yield condition),
 * and the accepted point's temp state and tangent come from it; every
 * other point, and every rejected one, takes the plain return mapping.  A
 * point rejected once keeps the plain update until the next commit, so the
 * choice cannot flip back and forth inside one load step's Newton loop.
 * net = NULL switches the surrogate off.  Every MisesMat material of the
 * mesh must be the network's. */
int  hfem_model_set_surrogate(hfem_model *md, hfem_mlp *net, double band, double tol);
/* counts since the last reset: {queued, accepted, rejected} */
int  hfem_model_surrogate_stats(hfem_model *md, int64_t *counts3, int reset);
/* FP32 forward of n rows (scaled inputs / outputs, host arrays) */
int  hfem_mlp_forward(hfem_mlp *m, int n, const float *X, float *Y);
/* gradient of the mean-squared error of n rows w.r.t. the flat parameters */
int  hfem_mlp_gradients(hfem_mlp *m, int n, const float *X, const float *Y, float *grad, double *loss);
/* BF16-operand MFMA inference with FP32 accumulation (host arrays) */
int  hfem_mlp_infer_bf16(hfem_mlp *m, int n, const float *X, float *Y);

#ifdef __cplusplus
}
#endif
#endif
