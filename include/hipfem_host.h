/*
 * hipfem_host.h -- C entry points of the C++ host layer (libhipfem_host.so).
 *
 * The host layer holds the reference-shaped C++ classes (HipCSRSparseMtrx,
 * HipPCGSolver, the class factory and a StructuralProblem with the
 * LinearStatic and StaticStructural + NRSolver drivers).  These functions
 * expose one problem to callers that are not C++ (Python tests and the
 * benchmark).  They have no single counterpart in the reference; each names
 * the driver it runs.  Return 0 on success, -1 on failure with the message in
 * hfem_problem_last_error().
 */
#ifndef HIPFEM_HOST_H
#define HIPFEM_HOST_H

#include "hipfem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfem_problem hfem_problem;

const char *hfem_problem_last_error(void);

/* mesh as in hfem_model_set_mesh (include/hipfem.h) */
int  hfem_problem_create(int nnode, const double *coords, int etype, int nelem, const int *conn, const int *emat,
                         int nmat, const double *mat, const int *eqn, int neq, int npeq, hfem_problem **out);
void hfem_problem_destroy(hfem_problem *p);

/* smtype / lstype are the deck integers (SMT_HipCSR = 11, ST_HipPCG = 9);
 * lstol, lsiter, lsprecond: IMLSolver keywords (imlsolver.h:48-52);
 * mixed, irtol: hippcg precision keywords; rtolf, maxiter, manrmsteps:
 * NRSolver keywords (nrsolver.h:54-73). */
int hfem_problem_set_options(hfem_problem *p, int smtype, int lstype, double lstol, int lsiter, int lsprecond,
                             int mixed, double irtol, double rtolf, int maxiter, int manrmsteps);
/* StaticStructural initial guess (_IFT_EngngModel_initialGuess,
 * staticstructural.C:130,247-297): 0 = IG_None, 1 = IG_Tangent (default) */
int hfem_problem_set_initial_guess(hfem_problem *p, int type);
/* Inexact Newton (no counterpart in the reference: NRSolver::solve,
 * nrsolver.C:205-404, solves every correction to lstol).  on = 1: each
 * Newton correction is solved to a forcing term between lstol and eta_max
 * (Eisenstat-Walker choice 2 from the force-error history, never tighter
 * than 0.1 rtolf / error); the convergence check is unchanged. */
int hfem_problem_set_forcing(hfem_problem *p, int on, double eta_max);
/* NRSolver keywords `linesearch`, `lsearchtol`, `lsearchamp`,
 * `lsearchmaxeta` (nrsolver.h:63, linesearch.h:45-47; defaults 0.8, 2.5,
 * 4.0) and `maxinc` (nrsolver.h:70; 1e20 = off).  After the search, the
 * correction is capped and added to X as in NRSolver::solve
 * (nrsolver.C:318-350), unless hfem_problem_set_nr_criteria chose
 * ls_keep_point. */
int hfem_problem_set_nr_extras(hfem_problem *p, int linesearch, double ls_tol, double ls_amp, double ls_maxeta,
                               double maxinc);
/* NRSolver `rtold` (displacement criterion, nrsolver.C:138-148, 786-806; 0
 * = off, the default) and `miniter` (nrsolver.C:113); ls_keep_point = 1
 * (no counterpart in the reference) caps a correction before the line
 * search and keeps the search's point, where the reference's NRSolver
 * caps after it and adds the searched step to X once more
 * (nrsolver.C:318-350), the default here too. */
int hfem_problem_set_nr_criteria(hfem_problem *p, double rtold, int miniter, int ls_keep_point);
/* convergence checks of the last nonlinear step (NRSolver::checkConvergence,
 * nrsolver.C:594-885): 6 values per check -- the force errors of the u, v,
 * w dof groups and their displacement errors, -1 for a group without free
 * equations or a criterion that is off; *n = checks (up to cap copied) */
int hfem_problem_nr_errors(hfem_problem *p, double *out, int cap, int *n);
/* StructuralEngngModel::updateInternalRHS (structengngmodel.C:215-233) called
 * through the EngngModel base: internal forces of the current solution
 * (neq) and the per-dof-id EBE norms (3) */
int hfem_problem_update_internal_rhs(hfem_problem *p, double *f, double *enorm3);
/* EngngModel::assemble (engngm.h:957-958) called through the EngngModel base
 * with TangentAssembler(response_mode) into the problem's stiffness matrix:
 * 0 TangentStiffness, 1 SecantStiffness, 2 ElasticStiffness */
int hfem_problem_assemble(hfem_problem *p, int response_mode);
/* NRSolver direct displacement control (`ddm` / `ddv`, nrsolver.C:165-176,
 * 458-568): the free equations eqs (1-based; the deck's node / dof pairs
 * mapped to equations) are driven by a 1e6 penalty on the diagonal of every
 * Newton matrix, their residual in the first iteration of a step being
 * K_ii * value * factor; set_ddm_factor sets that step's factor (the
 * increment of the `ddltf` function over the step, its value at the first
 * step) */
int hfem_problem_set_ddm(hfem_problem *p, int n, const int *eqs, const double *values);
int hfem_problem_set_ddm_factor(hfem_problem *p, double factor);
/* NRSolver `constrainednrminiter` (nrsolver.C:183-185, 320-326): non-zero
 * halves a Newton correction, from iteration miniter + 1 on, whenever the
 * squared norm of the per-dof-group force errors grew since the previous
 * iteration (the reference's constrainedNRalpha 0.5, which no keyword sets);
 * not applied to corrections that go through the line search */
int hfem_problem_set_constrained_nr(hfem_problem *p, int miniter);
/* NonLinearStatic `stiffmode` (nlinearstatic.h:61-65): 0 = the tangent
 * stiffness for every Newton matrix, 1 = the secant stiffness for every
 * one, 2 = the initial elastic stiffness, assembled once and kept, 3 = the
 * secant stiffness assembled once.  IsoLE and MisesMat answer the secant
 * request with the elastic matrix (misesmat.C:496-504). */
int hfem_problem_set_stiffness_mode(hfem_problem *p, int mode);
/* StaticStructural time stepping from the current time to final_time
 * (EngngModel::solveYourself engngm.C:573-662): prescribed values and
 * external load at time t are up_ref / f_free / f_presc times the
 * piecewise-linear load-time function through (ltf_t, ltf_v) (nltf = 0:
 * t / final_time; f_free NULL: no external load).  reduction = 0 is
 * NoReductionStrategy; 1 is SimpleReductionStrategy
 * (timestepreductionstrategy.C:70-110: dtmin, dtmax, ncrf, nmaxrestarts,
 * minrequirediter, maxrequirediter): a step whose Newton loop fails is
 * restarted from the last converged state (temp Gauss-point values reset)
 * with dt * ncrf, and dt adapts to each converged step's Newton iterations.
 * Per accepted step (up to max_records): rec_time, rec_dt (the step's
 * increment) and rec_int[6] = {newton, pcg, status, restarts,
 * line searches, 0}; *nrec = steps taken. */
int hfem_problem_run_steps(hfem_problem *p, const double *up_ref, const double *f_free, const double *f_presc,
                           int nltf, const double *ltf_t, const double *ltf_v, double final_time, double dt,
                           int reduction, double dtmin, double dtmax, double ncrf, int nmaxrestarts,
                           int min_required_iter, int max_required_iter, int max_records, double *rec_time,
                           double *rec_dt, int *rec_int, int *nrec);
/* prescribed values (npeq) and total external load (neq, npeq) of the next step */
int hfem_problem_set_prescribed(hfem_problem *p, const double *up);
int hfem_problem_set_load(hfem_problem *p, const double *f_free, const double *f_presc);

/* LinearStatic::solveYourselfAt (linearstatic.C:177-252) */
int hfem_problem_linear_static(hfem_problem *p, int *iterations, double *residual);
/* StaticStructural::solveYourselfAt + NRSolver::solve (staticstructural.C:224-358,
 * nrsolver.C:205-404): out3 = {Newton iterations, PCG iterations, ConvergedReason} */
int hfem_problem_nonlinear_step(hfem_problem *p, int *out3, double *force_error);
/* NonLinearStatic with indirect control (controlmode 0): CylindricalALM
 * (calmls.C:60-101, 681-897) with its keywords psi, maxiter, maxrestarts,
 * minsteplength, steplength, initialsteplength (<= 0: steplength),
 * reqiterations, miniter, manrmsteps, rtolf, rtold, rootselectiontype (0
 * cosine, 1 dot); hyperplane control through
 * hfem_problem_set_arc_length_hpc; the CALM line search and dof-group
 * norms are not part of this build.  Called again, it keeps the arc length
 * reached unless that exceeds the new steplength. */
int hfem_problem_set_arc_length(hfem_problem *p, double psi, int maxiter, int maxrestarts, double minsteplength,
                                double steplength, double initialsteplength, int reqiterations, int miniterations,
                                int manrmsteps, double rtolf, double rtold, int rootselectiontype);
/* CylindricalALM hyperplane control (hpcmode, hpc, hpcw; calmls.C:754-797,
 * 900-957): 1 = the constraint over the controlled equations only, 2 = the
 * linearized control sum_i w_i dX_i = steplength; eqs are the controlled
 * dofs' free equations (1-based) in the domain's dof-manager order, 0 for a
 * prescribed dof (skipped, as there); 0 switches the control off.  Call it
 * after hfem_problem_set_arc_length. */
int hfem_problem_set_arc_length_hpc(hfem_problem *p, int hpcmode, int n, const int *eqs, const double *weights);
/* the reference load R (NonLinearStatic::assembleIncrementalReferenceLoadVectors,
 * refloadmode 0: the total loads of the first step), free and prescribed parts */
int hfem_problem_set_reference_load(hfem_problem *p, const double *f_free, const double *f_presc);
/* one step of CylindricalALM::solve (calmls.C:108-481): out4 = {iterations,
 * PCG iterations, restarts, ConvergedReason}, out4d = {load level reached,
 * arc length of the next step, force error, displacement error}.  After
 * maxrestarts restarts the step is accepted as it stands (as the reference
 * does, with a warning) and the reason is CR_DIVERGED_ITS.  The reactions
 * then use the reached load level times R. */
int hfem_problem_arc_length_step(hfem_problem *p, int *out4, double *out4d);
/* steps driven by the caller with a time-step reduction: save_converged
 * keeps the current solution; restart_step returns to it and resets the
 * temp Gauss-point values (StaticStructural::restartYourself and
 * EngngModel::initStepIncrements, engngm.C:620-633) */
int hfem_problem_save_converged(hfem_problem *p);
int hfem_problem_restart_step(hfem_problem *p);
/* EngngModel::updateYourself: commit the Gauss-point state */
int hfem_problem_update(hfem_problem *p);

int hfem_problem_get_solution(hfem_problem *p, double *u);
/* StructuralEngngModel::computeReaction (structengngmodel.C:183-201) */
int hfem_problem_get_reactions(hfem_problem *p, double *r);
int hfem_problem_get_state(hfem_problem *p, double *s);
/* HIP-event timing of the PCG SpMV launches, summed over the solves since
 * profiling was switched on (instrumentation; no counterpart in the reference) */
int hfem_problem_set_profiling(hfem_problem *p, int on);
int hfem_problem_profile(hfem_problem *p, int *n64, double *ms64, int *n32, double *ms32);
/* last linear solve of the factory-made hippcg: outer iterations (mixed 1)
 * or reliable updates (mixed 2), and whether it applied the stiffness
 * element by element (instrumentation; no counterpart in the reference) */
int hfem_problem_solver_info(hfem_problem *p, int *outer, int *matrix_free);
/* the GPU context the problem runs on (to create a surrogate network on it) */
int hfem_problem_context(hfem_problem *p, hipfem_ctx **ctx);
/* hfem_model_set_surrogate / hfem_model_surrogate_stats on the problem's
 * batched evaluator (no counterpart in the reference) */
int hfem_problem_set_surrogate(hfem_problem *p, hfem_mlp *net, double band, double tol);
int hfem_problem_surrogate_stats(hfem_problem *p, long long *counts3, int reset);
/* SpMV layout of the factory-made matrix: form (1 node-block, 2 row-group,
 * 0 scalar CSR), node rows or row groups, 3x3 blocks or column chunks */
int hfem_problem_matrix_layout(hfem_problem *p, int *form, long long *groups, long long *blocks);
/* SparseMtrx::at / giveClassName of the factory-made stiffness matrix */
int hfem_problem_matrix_at(hfem_problem *p, int i, int j, double *v);
int hfem_problem_matrix_class(hfem_problem *p, char *buf, int len);
int hfem_problem_matrix_nnz(hfem_problem *p, long long *nnz);

#ifdef __cplusplus
}
#endif
#endif
