/*
 * fem_oracle.h -- CPU restatement of OOFEM's assembly-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product (oofem_amd/) never
 * links, loads or calls anything in oracle/.
 *
 * Conventions: 0-based arrays, row-major dense matrices, Voigt order
 * (xx, yy, zz, yz, xz, xy) with engineering shear strains, exactly as the
 * reference's Structural3DElement::computeBmatrixAt (structural3delement.C:64).
 *
 * Every function names the reference routine it restates; the .c file
 * carries the file:line citations.
 */
#ifndef FEM_ORACLE_H
#define FEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* element types (reference class names LSpace, LTRSpace, QSpace) */
#define ORC_LSPACE   1
#define ORC_LTRSPACE 2
#define ORC_QSPACE   3

/* material table row: [model, E, nu, sig0, H, 0, 0, 0] */
#define ORC_MATP     8
#define ORC_MAT_ISOLE 0
#define ORC_MAT_MISES 1

/* per Gauss point state block (doubles), see orc_state_* accessors:
 *  [0..5]  committed plastic strain      (MisesMatStatus::plasticStrain)
 *  [6]     committed kappa               (MisesMatStatus::kappa)
 *  [7..12] temp plastic strain           (tempPlasticStrain)
 *  [13]    temp kappa                    (tempKappa)
 *  [14..19] trial deviatoric stress      (trialStressD)
 *  [20..25] temp stress                  (tempStressVector)
 *  [26..31] temp strain                  (tempStrainVector)
 *  [32..37] committed stress             (stressVector)            */
#define ORC_GPSTATE 38

typedef struct {
    int nnode;
    const double *coords;   /* nnode*3 */
    int nelem;
    int etype;              /* one element type per mesh */
    const int *conn;        /* nelem*nen, 0-based node ids, reference node order */
    const int *emat;        /* nelem, index into mat */
    int nmat;
    const double *mat;      /* nmat*ORC_MATP */
    const int *eqn;         /* nnode*3: >0 free equation (1-based), <0 -(prescribed eq), 0 none */
    int neq;
    int npeq;
} orc_mesh;

int    orc_nen(int etype);
int    orc_ngp(int etype);
int    orc_gauss(int etype, double *xi, double *w);
int    orc_dndxi(int etype, const double *xi, double *dN);
double orc_dndx(int etype, const double *X, const double *xi, double *dNdx, int *err);
double orc_detj(int etype, const double *X, const double *xi);
void   orc_bmatrix(int nen, const double *dNdx, double *B);
void   orc_isole_D(double E, double nu, double *D);

int orc_element_stiffness(int etype, const double *X, const double *Dgp, int dstride,
                          int symmetric, double *Ke);

void orc_mises_return(const double *mp, const double *eps, const double *epsp_n, double kappa_n,
                      double *sig, double *epsp_t, double *kappa_t, double *sdev_trial);
void orc_mises_tangent(const double *mp, double kappa_n, double kappa_t,
                       const double *sdev_trial, double *D);

/* structure + assembly */
long long orc_csr_count(const orc_mesh *m);
int  orc_csr_fill(const orc_mesh *m, long long *rowptr, int *colind);
int  orc_assemble_stiffness(const orc_mesh *m, const long long *rowptr, const int *colind,
                            double *val, const double *state, int tangent);
void orc_spmv(int n, const long long *rowptr, const int *colind, const double *val,
              const double *x, double *y);

void orc_spmv_symcompcol(int n, const long long *rowptr, const int *colind, const double *val,
                         const double *x, double *y);

/* IML CG; precond bit 0 = Jacobi (DiagPreconditioner), bit 4 = SymCompCol SpMV order */
int orc_pcg(int n, const long long *rowptr, const int *colind, const double *val,
            const double *b, double *x, double tol, int maxit, int precond,
            int *iters, double *resid);
/* the same, storing ||r_k||/||b|| for k = 0..min(iters, cap-1) in hist */
int orc_pcg_hist(int n, const long long *rowptr, const int *colind, const double *val,
                 const double *b, double *x, double tol, int maxit, int precond,
                 int *iters, double *resid, double *hist, int cap);

/* element loops over a displacement field */
int orc_internal_forces(const orc_mesh *m, const double *u_free, const double *u_presc,
                        double *state, int update_state,
                        double *f_free, double *f_presc, double *ebe_norm3);
void orc_state_init(const orc_mesh *m, double *state);
void orc_state_commit(const orc_mesh *m, double *state);
void orc_state_restart(const orc_mesh *m, double *state);

/* OpenMP element loops as the reference's USE_OPENMP_PARALLEL build runs
 * them (parallel elements, critical-section scatter): CPU timing baselines */
int orc_omp_threads(void);
int orc_internal_forces_omp(const orc_mesh *m, const double *u_free, const double *u_presc, double *state,
                            int update_state, double *f_free, double *f_presc, double *ebe_norm3);
int orc_assemble_stiffness_omp(const orc_mesh *m, const long long *rowptr, const int *colind, double *val,
                               const double *state, int tangent);

/* surface traction (ConstantSurfaceLoad, global c.s.) on (element, surface) pairs */
int orc_surface_load(const orc_mesh *m, int nface, const int *elem, const int *surf,
                     const double *traction, double *f_free, double *f_presc);

/* DeadWeight body load on the elements elem[0..nel) with densities dens and
 * force per unit mass b (3), added into f_free / f_presc */
int orc_body_load(const orc_mesh *m, int nel, const int *elem, const double *dens, const double *b,
                  double *f_free, double *f_presc);

/* f = sum_e K_e(tangent) du_e over prescribed increments (free rows) */
int orc_extrapolated_forces(const orc_mesh *m, const double *state, const double *du_presc, double *f);

/* StaticStructural step (optional tangent initial guess) with NRSolver
 * (per-dof-id force criterion) */
int orc_nonlinear_step(const orc_mesh *m, double *state, double *X, const double *u_presc,
                       const double *u_presc_prev, int initial_guess, const double *RT,
                       double rtolf, int maxiter, int manrmsteps, double lstol, int *out, double *err);

/* NRSolver extras (nrsolver.C:106-200, 316-357; linesearch.C:42-245):
 * linesearch != 0 runs LineSearchNM after every correction but the first
 * (ls_tol, ls_amp, ls_maxeta as lsearchtol / lsearchamp / lsearchmaxeta,
 * clamped as in LineSearchNM::initializeFrom); maxinc caps the largest
 * component of a correction (maxIncAllowed).  out[3] receives the number
 * of line searches that ended with eta != 1.  stiffmode: NonLinearStatic's
 * stiffness (nlinearstatic.h:62-64): 0 the tangent, 2 the elastic stiffness
 * for every Newton matrix (the initial guess keeps the tangent product).
 * nddm > 0: NRSolver's direct displacement control (ddm / ddv,
 * nrsolver.C:458-568): the diagonal of every Newton matrix at the free
 * equations ddm_eq (1-based) is multiplied by 1e6, and the residual there is
 * K_ii * ddm_val * ddm_factor in the first iteration, 0 afterwards
 * (ddm_factor: the step's increment of the ddltf function). */
int orc_nonlinear_step_ex(const orc_mesh *m, double *state, double *X, const double *u_presc,
                          const double *u_presc_prev, int initial_guess, const double *RT, double rtolf, int maxiter,
                          int manrmsteps, double lstol, int linesearch, double ls_tol, double ls_amp, double ls_maxeta,
                          double maxinc, int stiffmode, int nddm, const int *ddm_eq, const double *ddm_val,
                          double ddm_factor, int cnr_miniter, int *out, double *err);

/* NRSolver / StaticStructural step options (nrsolver.C:100-200 keywords and
 * defaults; orc_nr_opts_default sets them) */
typedef struct {
    int initial_guess;     /* StaticStructural IG_Tangent over the prescribed increments */
    double rtolf, rtold;   /* force / displacement criteria (<= 0: off) */
    int maxiter, miniter, manrmsteps;
    double lstol;          /* relative tolerance of every linear solve */
    int linesearch;        /* LineSearchNM after every correction but the first */
    double ls_tol, ls_amp, ls_maxeta;
    int ls_keep_point;     /* 0: the reference's order (search, cap, X += ddX:
                            * the search's step is passed twice); 1: cap
                            * before the search and keep its point */
    double maxinc;         /* maxIncAllowed */
    int stiffmode;         /* 0 tangent, 1 secant, 2 elastic once, 3 secant once */
    int nddm;
    const int *ddm_eq;
    const double *ddm_val;
    double ddm_factor;
    int cnr_miniter;       /* constrainednrminiter */
} orc_nr_opts;

void orc_nr_opts_default(orc_nr_opts *o);

/* One step with options o.  out[4] = {Newton iterations, PCG iterations,
 * status (1 converged, -1 maxiter, -2 out of range), line searches with
 * eta != 1}; *err = the largest force error of the last check.  errhist
 * (may be NULL) receives 6 values per iteration up to errcap iterations:
 * the force errors of the u, v, w groups and their displacement errors, -1
 * for a group with no free equation or a criterion that is off. */
int orc_nonlinear_step_opt(const orc_mesh *m, double *state, double *X, const double *u_presc,
                           const double *u_presc_prev, const double *RT, const orc_nr_opts *o, int *out, double *err,
                           double *errhist, int errcap);

/* Skyline U^T D U factorization + solve of a full symmetric CSR (direct
 * solver of the reference; CPU baseline only).  *nwk_out: profile size. */
int orc_skyline_solve(int n, const long long *rowptr, const int *colind, const double *val, const double *b,
                      double *x, long long *nwk_out);

/* LinearStatic::solveYourselfAt + StructuralEngngModel::computeReaction */
int orc_linear_static(const orc_mesh *m, const double *f_ext, const double *f_ext_presc,
                      const double *u_presc, double tol, int maxit,
                      double *u, double *reactions, int *iters, double *resid);

#ifdef __cplusplus
}
#endif
#endif
