// StructuralProblem: a small engineering model for meshes of one element type
// (LSpace, LTRSpace or QSpace; IsoLE / MisesMat), with the two solution
// procedures the configurations use:
//   solveLinearStatic()   LinearStatic::solveYourselfAt (linearstatic.C:177-252)
//                         + StructuralEngngModel::computeReaction
//                         (structengngmodel.C:183-201)
//   solveNonlinearStep()  StaticStructural::solveYourselfAt
//                         (staticstructural.C:224-358: tangent initial guess
//                         over the prescribed increments) driving NRSolver::solve
//                         (nrsolver.C:205-404) with checkConvergence
//                         (nrsolver.C:594-885, per-dof-id force error)
// The sparse matrix and the linear solver come from the class factory by
// type (smtype / lstype), so the problem runs on hipcsr / hippcg exactly as a
// deck selecting them would.  Assembly goes through the batched evaluator
// (the EngngModel::assemble hook, engngm.C:889-1037).
#pragma once

#include <memory>
#include <vector>

#include "dynamicinputrecord.h"
#include "hippcg.h"
#include "structengngmodel.h"

namespace oofem {

// FP64 vector resident on the problem's GPU
class DeviceVector {
    hipfem_ctx *ctx = nullptr;
    double *p = nullptr;
    int64_t n = 0;

public:
    DeviceVector() {}
    DeviceVector(hipfem_ctx *c, int64_t len);
    ~DeviceVector();
    DeviceVector(const DeviceVector &) = delete;
    DeviceVector &operator=(const DeviceVector &) = delete;
    void allocate(hipfem_ctx *c, int64_t len);
    double *ptr() { return p; }
    const double *ptr() const { return p; }
    int64_t size() const { return n; }
    void upload(const double *h);
    void download(double *h) const;
    void axpby(double a, const DeviceVector *x, double b);   // this = a x + b this
    void zero() { axpby(0.0, nullptr, 0.0); }
    double dot(const DeviceVector &o) const;
    double amax() const;   // max |x_i|
};

struct NRSolverOptions {
    double rtolf = 1.e-3;      // nrsolver.C:137 default
    double rtold = 0.0;        // displacement criterion, off by default (nrsolver.C:138-141)
    int maxiter = 100000000;   // nsmax default 1e8 (nrsolver.C:106)
    int minIterations = 0;
    int manrmsteps = 0;        // 0: modified NR (default), k: tangent every k iterations
    bool calcStiffBeforeRes = true;
    // inexact Newton (no counterpart in NRSolver, which solves every
    // correction to lstol): the correction of iteration k is solved to the
    // forcing term eta_k = min(forcingMax, max(0.9 (e_k / e_{k-1})^2,
    // 0.9 eta_{k-1}^2 when that exceeds 0.1)), at least
    // 0.1 rtolf / e_k (no tighter than the step's criterion can use) and
    // at least lstol; e_k is the force error of the convergence check.
    // The convergence check itself is unchanged.
    bool forcing = false;
    double forcingMax = 0.1;
    // line search after every correction but the first (NRSolver
    // `linesearch`, nrsolver.C:183-188, 318-324; LineSearchNM keywords
    // lsearchtol / lsearchamp / lsearchmaxeta and defaults, linesearch.C:44-51,
    // clamped as in LineSearchNM::initializeFrom, linesearch.C:247-283)
    bool lineSearch = false;
    double lsTol = 0.8, lsAmp = 2.5, lsMaxEta = 4.0;
    // false: the reference's order after the search (nrsolver.C:318-350):
    // the cap, then X += ddX, so the search's step is passed twice; true
    // (no counterpart in the reference): the cap before the search, which
    // then runs along the capped direction, and X stays at the search's point
    bool lsKeepPoint = false;
    // cap on the largest component of a correction (NRSolver `maxinc`,
    // nrsolver.C:340-353)
    double maxInc = 1.0e20;
    // stiffness of the Newton matrix (NonLinearStatic `stiffmode`,
    // nlinearstatic.h:62-64): 0 the tangent, 2 the initial elastic
    // stiffness, assembled once and kept (nlinearstatic.C:635-644)
    int stiffMode = 0;
    // direct displacement control (NRSolver `ddm` / `ddv` / `ddltf`,
    // nrsolver.C:165-176, 458-568): free equations (1-based) held by a 1e6
    // diagonal penalty, their values, and the step's increment factor of
    // the ddltf function
    std::vector<int> ddmEq;
    std::vector<double> ddmVal;
    double ddmFactor = 0.0;
    int constrainedMinIter = 0;   // constrainednrminiter (0 = off): halve a correction when the force errors grew
};

// SimpleReductionStrategy (timestepreductionstrategy.C:40-110) of the
// metastep: on == false is NoReductionStrategy (a step that does not
// converge is kept, the increment never changes)
struct ReductionStrategy {
    bool on = false;
    double dtmin = 0.0, dtmax = 1.0e300;   // dtmin / dtmax (required with on)
    double ncrf = 0.5;                     // noConvergenceReductionFactor
    int nmaxrestarts = 5;                  // numberOfMaxTimeStepReductions
    int minRequiredIter = 3, maxRequiredIter = 8;
};

// one accepted time step of solveTimeSteps
struct StepRecord {
    double time = 0.0, dt = 0.0;
    int newton = 0, pcg = 0, status = 0, restarts = 0, lineSearches = 0;
};

// StaticStructural initial guess (_IFT_EngngModel_initialGuess): IG_None = 0,
// IG_Tangent = 1 (the default, staticstructural.C:130)
enum InitialGuess { IG_None = 0, IG_Tangent = 1 };

// CylindricalALM input (calmls.C:60-101 defaults, initializeFrom 681-897):
// hyperplane control off, no CALM line search, norms over all equations
struct CALMOptions {
    double psi = 1.0;
    int maxiter = 30;              // nsmax, at least 30
    int maxRestarts = 4;
    double minStepLength = 0.0;
    double maxStepLength = 0.0;    // `steplength` (required)
    double initialStepLength = 0.0;   // default: steplength
    int reqIterations = 3;         // clamped to [3, 1000]
    int minIterations = 0;
    int manrmsteps = 0;            // > 0: tangent every manrmsteps iterations (accelerated NRM)
    double rtolf = 1.e-3, rtold = 1.e-3;
    int rootSelection = 0;         // 0 cosine of the increments (RST_Cos), 1 dot product (RST_Dot)
    int hpcMode = 0;               // 0 off, 1 constraint on the hpc equations only, 2 linearized (weighted)
    std::vector<int> hpcEq;        // controlled free equations (1-based; 0 = prescribed, skipped)
    std::vector<double> hpcW;      // weights of the linearized control
};

struct CALMStepReport {
    int iterations = 0, linearIterations = 0, restarts = 0;
    double lambda = 0.0;           // load level reached
    double stepLength = 0.0;       // arc length for the next step
    double forceError = 0.0, dispError = 0.0;
    ConvergedReason status = CR_UNKNOWN;
};

struct NRStepReport {
    int iterations = 0;        // Newton iterations
    int linearIterations = 0;  // PCG iterations summed over the step
    double forceError = 0.0;   // final max force error over dof groups
    int lineSearches = 0;      // line searches that ended with eta != 1
    // every convergence check: force errors of the u, v, w groups, then their
    // displacement errors (-1: group without free equations, or criterion off)
    std::vector<double> errors;
    ConvergedReason status = CR_UNKNOWN;
};

class StructuralProblem : public StructuralEngngModel, public BatchedEvaluatorProvider {
    hipfem_ctx *ctx;
    int nnode, etype, nelem, nmat, neq, npeq;
    std::vector<int> conn, eqn;
    std::vector<int8_t> eqDofId;   // dof id (0 u, 1 v, 2 w) of every free equation
    hfem_model *ev = nullptr;
    SparseMtrxType smtype = SMT_HipCSR;
    LinSystSolverType lstype = ST_HipPCG;
    DynamicInputRecord solverRecord{"hippcg"};
    std::unique_ptr<SparseMtrx> K;
    std::unique_ptr<SparseLinearSystemNM> linSolver;
    // state on the device
    DeviceVector X, up, upPrev, dup, RT, RTp, F, Fp, rhs, ddX;
    DeviceVector rb, gls, Xc;   // line-search start and residual, last converged solution
    // arc-length state kept between steps (CylindricalALM members)
    DeviceVector Rref, Rpref, deltaXt, deltaXb, dXs, dXm1, XInit, oldDX;
    CALMOptions calm;
    bool calmSet = false, haveOldDX = false;
    double calmDeltaL = -1.0, reachedLambda = 0.0;
    int calmMode = 0, calmOldMode = 0, calmTick = -1;   // 0 modified, 1 accelerated, 2 full NRM
    bool applyInitialGuess(int &linearIterations, bool &tangentCurrent);
    double currentTime = 0.0;   // target time of the last accepted step of solveTimeSteps
    int kMode = -1;             // stiffness the values of K hold: 0 elastic, 1 tangent, -1 none
    long smConstraintVersion = -1;   // K version after the last ddm penalty
    TimeStep currentStep{1, 0.0};
    void assembleStiffness(SparseMtrx &answer, int mode);
    void assembleNewtonMatrix();
    void applyConstraintsToStiffness();
    void applyConstraintsToLoadIncrement(int nite, double sums6[6]);
    bool structureBuilt = false;
    int lastLinearIterations = 0;
    std::vector<double> lastNRErrors;   // NRStepReport::errors of the last step
    bool profiling = false;
    int profN64 = 0, profN32 = 0;
    double profMs64 = 0.0, profMs32 = 0.0;

    void ensureStructure();
    void internalForces(const DeviceVector &x, bool updateState, double *ebe3);
    ConvergedReason linearSolve(const DeviceVector &b, DeviceVector &x);
    double lineSearch(int &lineSearches, bool &accepted);

public:
    NRSolverOptions nr;
    InitialGuess initialGuessType = IG_Tangent;

    StructuralProblem(int nnode, const double *coords, int etype, int nelem, const int *conn, const int *emat,
                      int nmat, const double *mat, const int *eqn, int neq, int npeq);
    ~StructuralProblem() override;

    // EngngModel hooks used by the plug-ins
    int giveNumberOfDomainEquations(int di, const UnknownNumberingScheme &num) override { return neq; }
    hfem_model *giveBatchedEvaluator(int di) override { return ev; }
    // EngngModel::assemble (engngm.h:957-958): one batched launch into a
    // HipCSR matrix instead of the element loop; the TangentAssembler's
    // response mode picks the consistent tangent (TangentStiffness) or the
    // elastic matrix (ElasticStiffness; SecantStiffness answers with it too,
    // misesmat.C:496-504)
    void assemble(SparseMtrx &answer, TimeStep *tStep, const MatrixAssembler &ma, const UnknownNumberingScheme &s,
                  Domain *domain) override;
    // StructuralEngngModel::updateInternalRHS (structengngmodel.C:215-233):
    // internal forces of the current solution (default numbering) through
    // the batched evaluator, the Gauss-point temp state updated, and the
    // per-dof-id EBE norms (u, v, w) in eNorm
    void updateInternalRHS(FloatArray &answer, TimeStep *tStep, Domain *d, FloatArray *eNorm) override;
    TimeStep *giveCurrentStep() { return &currentStep; }

    void setSparseMtrxType(int t) { smtype = (SparseMtrxType)t; }
    void setLinSystSolverType(int t) { lstype = (LinSystSolverType)t; }
    DynamicInputRecord &giveSolverRecord() { return solverRecord; }

    void setPrescribedValues(const double *up_host);               // npeq values for the step
    void setExternalLoad(const double *f_free, const double *f_presc);  // total external load


    ConvergedReason solveLinearStatic(int &iterations, double &residual);
    NRStepReport solveNonlinearStep();
    // StaticStructural time stepping (EngngModel::solveYourself engngm.C:573-662
    // with the metastep's reduction strategy): from the current time to
    // finalTime with initial increment dt; prescribed values and external
    // load at time t are the reference ones times ltf(t), the piecewise-linear
    // function through (ltf_t, ltf_v) (empty: t / finalTime).  f_free /
    // f_presc may be null (no external load).  Every accepted step is
    // committed (updateYourself).
    std::vector<StepRecord> solveTimeSteps(const double *up_ref, const double *f_free, const double *f_presc,
                                           const std::vector<double> &ltf_t, const std::vector<double> &ltf_v,
                                           double finalTime, double dt, const ReductionStrategy &rs);
    // NonLinearStatic with indirect control (controlmode 0): one step of
    // CylindricalALM::solve, calmls.C:108-481, on the reference load set by
    // setReferenceLoad; the load level persists between steps
    void setArcLength(const CALMOptions &o);
    void setReferenceLoad(const double *f_free, const double *f_presc);
    // hyperplane control (hpcmode 1 or 2; 0 switches it off): controlled
    // free equations, 0 for a prescribed dof, and the weights of mode 2
    void setArcLengthHpc(int mode, const std::vector<int> &eqs, const std::vector<double> &w);
    CALMStepReport solveArcLengthStep();
    double giveLoadLevel() const { return reachedLambda; }
    void updateYourself();   // commit Gauss-point state (EngngModel::updateYourself)
    // a step driven from outside: keep the converged solution, and go back
    // to it with the temp Gauss-point values reset
    // (StaticStructural::restartYourself, EngngModel::initStepIncrements)
    void saveConverged() { Xc.axpby(1.0, &X, 0.0); }
    void restartStep();

    void giveSolution(double *u) const { X.download(u); }
    void giveReactions(double *r);
    void giveGaussPointState(double *s) { hipfemCheck(hfem_model_get_state(ev, s), "get_state"); }
    int giveNeq() const { return neq; }
    int giveNpeq() const { return npeq; }
    int giveLastLinearIterations() const { return lastLinearIterations; }
    const std::vector<double> &giveLastNRErrors() const { return lastNRErrors; }
    // SpMV launch timing of the hippcg solver, summed over every solve since
    // setProfiling(true)
    void setProfiling(bool on);
    void giveProfile(int *n64, double *ms64, int *n32, double *ms32) const;
    SparseMtrx *giveStiffnessMatrix() { return K.get(); }
    SparseLinearSystemNM *giveLinearSolver() { return linSolver.get(); }
    hipfem_ctx *giveContext() { return ctx; }
};

}  // namespace oofem
