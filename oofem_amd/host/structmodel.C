#include "structmodel.h"

#include <cmath>

#include "classfactory.h"

namespace oofem {

// ------------------------------------------------------------ DeviceVector
DeviceVector::DeviceVector(hipfem_ctx *c, int64_t len) { allocate(c, len); }

DeviceVector::~DeviceVector()
{
    if (p) hipfem_dev_free(ctx, p);
}

void DeviceVector::allocate(hipfem_ctx *c, int64_t len)
{
    if (p) hipfem_dev_free(ctx, p);
    ctx = c;
    n = len;
    hipfemCheck(hipfem_dev_alloc(ctx, std::max<int64_t>(1, n), &p), "dev_alloc");
    zero();
}

void DeviceVector::upload(const double *h) { hipfemCheck(hipfem_dev_upload(ctx, p, h, n), "dev_upload"); }
void DeviceVector::download(double *h) const { hipfemCheck(hipfem_dev_download(ctx, h, p, n), "dev_download"); }

void DeviceVector::axpby(double a, const DeviceVector *x, double b)
{
    hipfemCheck(hipfem_dev_axpby(ctx, n, a, x ? x->p : nullptr, b, p), "dev_axpby");
}

double DeviceVector::dot(const DeviceVector &o) const
{
    double r = 0.0;
    hipfemCheck(hipfem_dev_dot(ctx, n, p, o.p, &r), "dev_dot");
    return r;
}

double DeviceVector::amax() const
{
    double r = 0.0;
    hipfemCheck(hipfem_dev_amax(ctx, n, p, &r), "dev_amax");
    return r;
}

// ------------------------------------------------------- StructuralProblem
StructuralProblem::StructuralProblem(int nnode_, const double *coords, int etype_, int nelem_, const int *conn_,
                                     const int *emat, int nmat_, const double *mat, const int *eqn_, int neq_,
                                     int npeq_) :
    ctx(hipfemDefaultContext()), nnode(nnode_), etype(etype_), nelem(nelem_), nmat(nmat_), neq(neq_), npeq(npeq_)
{
    const int nen = etype == HIPFEM_LSPACE ? 8 : etype == HIPFEM_LTRSPACE ? 4 : 20;
    conn.assign(conn_, conn_ + (size_t)nelem * nen);
    eqn.assign(eqn_, eqn_ + (size_t)nnode * 3);
    hipfemCheck(hfem_model_create(ctx, &ev), "hfem_model_create");
    hipfemCheck(hfem_model_set_mesh(ev, nnode, coords, etype, nelem, conn_, emat, nmat, mat, eqn_, neq, npeq),
                "set_mesh");
    for (DeviceVector *v : {&X, &RT, &F, &rhs, &ddX, &rb, &gls, &Xc}) v->allocate(ctx, neq);
    eqDofId.assign((size_t)std::max(neq, 0), 0);
    for (size_t i = 0; i < eqn.size(); i++)
        if (eqn[i] > 0) eqDofId[eqn[i] - 1] = (int8_t)(i % 3);
    for (DeviceVector *v : {&up, &upPrev, &dup, &RTp, &Fp}) v->allocate(ctx, npeq);
    // one domain whose elements answer their location arrays from conn / eqn
    domainList.push_back(std::make_unique<Domain>(1));
    domainList[0]->setElements(nelem, [this, nen](int e, IntArray &loc, IntArray *dofIds) {
        loc.resize(3 * nen);
        if (dofIds) dofIds->resize(3 * nen);
        for (int a = 0; a < nen; a++)
            for (int d = 0; d < 3; d++) {
                const int q = eqn[3 * (size_t)conn[(size_t)(e - 1) * nen + a] + d];
                loc.at(3 * a + d + 1) = q > 0 ? q : 0;
                if (dofIds) dofIds->at(3 * a + d + 1) = d + 1;
            }
    });
    // IMLSolver defaults unless the caller overrides them
    solverRecord.setField(1, _IFT_IMLSolver_lsprecond);
}

StructuralProblem::~StructuralProblem()
{
    linSolver.reset();
    K.reset();
    if (ev) hfem_model_destroy(ev);
}

void StructuralProblem::setPrescribedValues(const double *h) { up.upload(h); }

void StructuralProblem::setExternalLoad(const double *f_free, const double *f_presc)
{
    if (f_free) RT.upload(f_free); else RT.zero();
    if (f_presc) RTp.upload(f_presc); else RTp.zero();
}

void StructuralProblem::ensureStructure()
{
    if (structureBuilt) return;
    K = GiveClassFactory().createSparseMtrx(smtype);
    if (!K) OOFEM_ERROR("sparse matrix creation failed (smtype %d is not registered)", (int)smtype);
    K->buildInternalStructure(this, 1, EModelDefaultEquationNumbering());
    linSolver = GiveClassFactory().createSparseLinSolver(lstype, nullptr, this);
    if (!linSolver) OOFEM_ERROR("linear solver creation failed (lstype %d is not registered)", (int)lstype);
    linSolver->initializeFrom(solverRecord);
    structureBuilt = true;
}

void StructuralProblem::assembleStiffness(SparseMtrx &answer, int mode)
{
    auto *H = dynamic_cast<HipCSRSparseMtrx *>(&answer);
    if (!H) OOFEM_ERROR("the batched evaluator assembles into a HipCSRSparseMtrx only (got %s)", answer.giveClassName());
    hipfemCheck(hfem_model_assemble_stiffness(ev, H->handle(), mode), "assemble");
    if (&answer == K.get()) kMode = mode;
}

namespace {
// reads TangentAssembler's protected response mode through a member pointer
// named in a derived class (no change to the assembler's declaration)
struct TangentModeOf : TangentAssembler {
    static MatResponseMode get(const TangentAssembler &a) { return a.*(&TangentModeOf::rmode); }
};
}  // namespace

void StructuralProblem::assemble(SparseMtrx &answer, TimeStep *tStep, const MatrixAssembler &ma,
                                 const UnknownNumberingScheme &s, Domain *domain)
{
    auto *ta = dynamic_cast<const TangentAssembler *>(&ma);
    if (!ta) OOFEM_ERROR("the batched evaluator assembles stiffness matrices (TangentAssembler) only");
    if (domain != giveDomain(1)) OOFEM_ERROR("unknown domain");
    const MatResponseMode rm = TangentModeOf::get(*ta);
    assembleStiffness(answer, rm == TangentStiffness ? 1 : 0);
}

void StructuralProblem::updateInternalRHS(FloatArray &answer, TimeStep *tStep, Domain *d, FloatArray *eNorm)
{
    if (d != giveDomain(1)) OOFEM_ERROR("unknown domain");
    if (tStep) tStep->incrementStateCounter();
    double ebe[3];
    internalForces(X, true, ebe);
    answer.resize(neq);
    F.download(answer.givePointer());
    if (eNorm) {
        eNorm->resize(3);
        for (int k = 0; k < 3; k++) eNorm->at(k + 1) = ebe[k];
    }
}

void StructuralProblem::internalForces(const DeviceVector &x, bool updateState, double *ebe3)
{
    hipfemCheck(hfem_model_internal_forces_dev(ev, x.ptr(), npeq ? up.ptr() : nullptr, F.ptr(),
                                               npeq ? Fp.ptr() : nullptr, ebe3, updateState ? 1 : 0),
                "internal forces");
}

ConvergedReason StructuralProblem::linearSolve(const DeviceVector &b, DeviceVector &x)
{
    auto *H = dynamic_cast<HipCSRSparseMtrx *>(K.get());
    auto *S = dynamic_cast<HipPCGSolver *>(linSolver.get());
    if (H && S) {
        hipfemCheck(hippcg_set_profiling(S->handle(), profiling ? 1 : 0), "set_profiling");
        ConvergedReason r = S->solveDevice(*H, b.ptr(), x.ptr());
        lastLinearIterations = S->giveLastIterations();
        if (profiling) {
            int n64 = 0, n32 = 0;
            double m64 = 0.0, m32 = 0.0;
            hipfemCheck(hippcg_profile(S->handle(), &n64, &m64, &n32, &m32), "profile");
            profN64 += n64;
            profN32 += n32;
            profMs64 += m64;
            profMs32 += m32;
        }
        return r;
    }
    // generic SparseLinearSystemNM path on host vectors
    FloatArray bh(neq), xh(neq);
    b.download(bh.givePointer());
    x.download(xh.givePointer());
    ConvergedReason r = linSolver->solve(*K, bh, xh);
    x.upload(xh.givePointer());
    lastLinearIterations = 0;
    return r;
}

// LinearStatic::solveYourselfAt linearstatic.C:177-252
ConvergedReason StructuralProblem::solveLinearStatic(int &iterations, double &residual)
{
    ensureStructure();
    // LinearStatic assembles the stiffness with TangentAssembler(TangentStiffness)
    // (linearstatic.C:199); for IsoLE that is the elastic matrix
    assembleStiffness(*K, 0);
    X.zero();
    internalForces(X, false, nullptr);            // f_int of the prescribed values
    double sums[6];
    hipfemCheck(hfem_model_residual(ev, RT.ptr(), F.ptr(), rhs.ptr(), sums), "residual");   // rhs = f_ext - f_int
    ConvergedReason s = linearSolve(rhs, X);
    if (s != CR_CONVERGED) OOFEM_ERROR("No success in solving system.");
    auto *S = dynamic_cast<HipPCGSolver *>(linSolver.get());
    iterations = S ? S->giveLastIterations() : 0;
    residual = S ? S->giveLastResidual() : 0.0;
    internalForces(X, true, nullptr);             // stresses and reaction forces
    return s;
}

// the Newton matrix: the tangent, or the elastic stiffness assembled once
// (NonLinearStatic::updateComponent NonLinearLhs, nlinearstatic.C:616-645)
// NonLinearStatic::updateComponent(NonLinearLhs) nlinearstatic.C:615-650:
// 0 tangent every time; 1 secant every time; 2 elastic once; 3 secant
// once.  IsoLE and MisesMat answer SecantStiffness with the elastic matrix
// (misesmat.C:496-504), so 1 reassembles the elastic matrix and 3 is 2.
void StructuralProblem::assembleNewtonMatrix()
{
    // this->assemble(K, tStep, TangentAssembler(mode), numbering, domain) as
    // NonLinearStatic::updateComponent does (nlinearstatic.C:625-641)
    const EModelDefaultEquationNumbering dn;
    if (nr.stiffMode == 2 || nr.stiffMode == 3) {
        if (kMode != 0) this->assemble(*K, &currentStep, TangentAssembler(ElasticStiffness), dn, giveDomain(1));
    } else if (nr.stiffMode == 1) {
        this->assemble(*K, &currentStep, TangentAssembler(SecantStiffness), dn, giveDomain(1));
    } else {
        this->assemble(*K, &currentStep, TangentAssembler(TangentStiffness), dn, giveDomain(1));
    }
}

// NRSolver::applyConstraintsToStiffness nrsolver.C:458-500: K_ii *= 1e6 at
// the directly controlled equations, once per matrix version
void StructuralProblem::applyConstraintsToStiffness()
{
    if (nr.ddmEq.empty()) return;
    auto *H = dynamic_cast<HipCSRSparseMtrx *>(K.get());
    if (!H) OOFEM_ERROR("direct displacement control needs the HipCSR matrix");
    if (hipcsr_version(H->handle()) == smConstraintVersion) return;
    for (int q : nr.ddmEq) {
        double v = 0.0;
        hipfemCheck(hipcsr_get(H->handle(), q, q, &v), "ddm diagonal");
        hipfemCheck(hipcsr_set(H->handle(), q, q, v * 1.e6), "ddm diagonal");
    }
    smConstraintVersion = hipcsr_version(H->handle());
}

// NRSolver::applyConstraintsToLoadIncrement nrsolver.C:502-568: the residual
// at a directly controlled equation is K_ii * value * factor in the first
// iteration and 0 afterwards; the per-dof-id squared sums follow
void StructuralProblem::applyConstraintsToLoadIncrement(int nite, double sums6[6])
{
    if (nr.ddmEq.empty()) return;
    auto *H = dynamic_cast<HipCSRSparseMtrx *>(K.get());
    for (size_t k = 0; k < nr.ddmEq.size(); k++) {
        const int q = nr.ddmEq[k];
        double old = 0.0, kii = 0.0;
        hipfemCheck(hipfem_dev_download(ctx, &old, rhs.ptr() + q - 1, 1), "ddm residual");
        if (nite == 0) hipfemCheck(hipcsr_get(H->handle(), q, q, &kii), "ddm diagonal");
        const double v = nite == 0 ? kii * nr.ddmVal[k] * nr.ddmFactor : 0.0;
        hipfemCheck(hipfem_dev_upload(ctx, rhs.ptr() + q - 1, &v, 1), "ddm residual");
        double &sr = sums6[eqDofId[q - 1]];
        sr = std::max(0.0, sr + v * v - old * old);
    }
}

// staticstructural.C:247-297 / nlinearstatic.C:523-539 (IG_Tangent):
// extrapolated forces -sum_e K_e du_e of the prescribed increments (free
// entries of du_e are zero at step start), then K inc = f from zero and
// X += inc.  tangentCurrent tells whether K now holds the tangent of the
// current state.  Returns false when the linear solve fails.
bool StructuralProblem::applyInitialGuess(int &linearIterations, bool &tangentCurrent)
{
    tangentCurrent = false;
    if (initialGuessType != IG_Tangent || npeq == 0) return true;
    dup.axpby(1.0, &up, 0.0);
    dup.axpby(-1.0, &upPrev, 1.0);
    hipfemCheck(hfem_model_tangent_product_dev(ev, 1, nullptr, dup.ptr(), rhs.ptr(), nullptr), "extrapolate");
    rhs.axpby(0.0, nullptr, -1.0);
    if (rhs.dot(rhs) > 0.0) {
        assembleNewtonMatrix();                 // updateMatrix
        tangentCurrent = true;
        ddX.zero();
        ConvergedReason ls = linearSolve(rhs, ddX);
        linearIterations += lastLinearIterations;
        if (ls != CR_CONVERGED) return false;
        X.axpby(1.0, &ddX, 1.0);
    }
    return true;
}

// StaticStructural::solveYourselfAt staticstructural.C:224-358 and
// NRSolver::solve nrsolver.C:205-404.
NRStepReport StructuralProblem::solveNonlinearStep()
{
    ensureStructure();
    NRStepReport rep;
    // idsInUse per dof id (nrsolver.C:648-660)
    int idsInUse[3] = {0, 0, 0};
    for (size_t i = 0; i < eqn.size(); i++)
        if (eqn[i] > 0) idsInUse[i % 3]++;
    bool tangentCurrent = false;
    if (!applyInitialGuess(rep.linearIterations, tangentCurrent)) {
        rep.status = CR_FAILED;
        return rep;
    }
    ddX.zero();
    // stiffness before the first residual (mCalcStiffBeforeRes, nrsolver.C:245-256);
    // the matrix of the initial guess is that same tangent (no Gauss-point
    // state changed since), so it is reused
    if (!tangentCurrent) assembleNewtonMatrix();
    applyConstraintsToStiffness();
    auto *pcg = dynamic_cast<HipPCGSolver *>(linSolver.get());
    const double lstol = pcg ? pcg->giveTolerance() : 0.0;
    double prevErr = -1.0, prevEta = -1.0;
    // per-dof-group force errors of this and the previous check (NRSolver
    // forceErrVec / forceErrVecOld, nrsolver.C:612-623) for constrained Newton
    double feNew[3] = {0.0, 0.0, 0.0}, feOld[3] = {0.0, 0.0, 0.0};
    for (int nite = 0;; ++nite) {
        double ebe[3], sums[6], dsums[6] = {0, 0, 0, 0, 0, 0};
        internalForces(X, true, ebe);               // updateComponent(InternalRhs)
        hipfemCheck(hfem_model_residual(ev, RT.ptr(), F.ptr(), rhs.ptr(), sums), "residual");
        applyConstraintsToLoadIncrement(nite, sums);
        if (nr.rtold > 0.0) hipfemCheck(hfem_model_group_sums(ev, ddX.ptr(), X.ptr(), dsums), "group sums");
        // checkConvergence nrsolver.C:594-885, grouped by dof id
        bool converged = true, outOfRange = false;
        double maxErr = 0.0;
        if (nr.constrainedMinIter)
            for (int dg = 0; dg < 3; dg++) {
                feOld[dg] = feNew[dg];
                feNew[dg] = 0.0;
            }
        const size_t e0 = rep.errors.size();
        rep.errors.resize(e0 + 6, -1.0);
        for (int dg = 0; dg < 3; dg++) {
            if (!idsInUse[dg]) continue;
            if (nr.rtolf > 0.0) {
                const double denom = sums[3 + dg] + ebe[dg];
                double forceErr = denom >= 1.e-6 ? std::sqrt(sums[dg] / denom) : std::sqrt(sums[dg]);
                feNew[dg] = forceErr;
                if (std::isnan(forceErr) || forceErr > nr.rtolf * 1.e20) outOfRange = true;
                if (forceErr > nr.rtolf) converged = false;
                if (forceErr > 0.0 && nite == 0) converged = false;
                maxErr = std::max(maxErr, forceErr);
                rep.errors[e0 + dg] = forceErr;
            }
            if (nr.rtold > 0.0) {
                // displacement error (nrsolver.C:786-806): last correction over the total
                const double dispErr = dsums[3 + dg] > 1.e-6 ? std::sqrt(dsums[dg] / dsums[3 + dg])
                                                             : std::sqrt(dsums[dg]);
                if (std::isnan(dispErr) || dispErr > nr.rtold * 1.e20) outOfRange = true;
                if (dispErr > nr.rtold) converged = false;
                rep.errors[e0 + 3 + dg] = dispErr;
            }
        }
        rep.forceError = maxErr;
        rep.iterations = nite;
        if (outOfRange) { rep.status = CR_DIVERGED_TOL; break; }
        if (converged && nite >= nr.minIterations) { rep.status = CR_CONVERGED; break; }
        if (nite >= nr.maxiter) { rep.status = CR_DIVERGED_ITS; break; }
        if (nite > 0 || !nr.calcStiffBeforeRes) {
            if (nr.manrmsteps > 0 && nite % nr.manrmsteps == 0) {
                assembleNewtonMatrix();
                applyConstraintsToStiffness();
            }
        }
        if (nr.forcing && pcg) {
            double eta = nr.forcingMax;
            if (prevErr > 0.0) {
                eta = 0.9 * (maxErr / prevErr) * (maxErr / prevErr);
                const double keep = 0.9 * prevEta * prevEta;
                if (keep > 0.1) eta = std::max(eta, keep);
                eta = std::min(eta, nr.forcingMax);
            }
            if (maxErr > 0.0) eta = std::max(eta, 0.1 * nr.rtolf / maxErr);
            eta = std::min(std::max(eta, lstol), std::max(nr.forcingMax, lstol));
            pcg->setTolerance(eta);
            prevEta = eta;
        }
        prevErr = maxErr;
        ConvergedReason ls = linearSolve(rhs, ddX);  // ddX carries the previous correction
        rep.linearIterations += lastLinearIterations;
        if (ls != CR_CONVERGED) { rep.status = CR_FAILED; break; }
        const bool doLineSearch = nr.lineSearch && nite > 0;
        // constrained Newton (nrsolver.C:320-326): from iteration
        // constrainednrminiter + 1 on, halve the correction when the squared
        // norm of the force errors grew; not with the line search
        if (nr.constrainedMinIter && !doLineSearch && nite > nr.constrainedMinIter) {
            const double a = feNew[0] * feNew[0] + feNew[1] * feNew[1] + feNew[2] * feNew[2];
            const double b = feOld[0] * feOld[0] + feOld[1] * feOld[1] + feOld[2] * feOld[2];
            if (a > b) ddX.axpby(0.0, nullptr, 0.5);
        }
        // increment cap (nrsolver.C:332-345): after the line search in the
        // reference's order, before it when the search keeps its point
        auto capIncrement = [&]() {
            if (nr.maxInc < 1.0e20) {
                const double mx = ddX.amax();
                if (mx > nr.maxInc) ddX.axpby(0.0, nullptr, nr.maxInc / mx);
            }
        };
        if (!doLineSearch || nr.lsKeepPoint) capIncrement();
        if (doLineSearch) {
            // LineSearchNM::solve moves X to X_start + eta ddX (ddX scaled by
            // eta) when it meets its tolerance, else to X_start + ddX
            bool accepted = false;
            lineSearch(rep.lineSearches, accepted);
            if (!nr.lsKeepPoint) {
                // NRSolver::solve then caps the correction and adds it to X
                // (nrsolver.C:332-350): the search's step is passed twice
                capIncrement();
                X.axpby(1.0, &ddX, 1.0);
            }
        } else {
            X.axpby(1.0, &ddX, 1.0);
        }
    }
    if (nr.forcing && pcg) pcg->setTolerance(lstol);
    lastNRErrors = rep.errors;
    return rep;
}

// LineSearchNM::search linesearch.C:163-245 (1-based arrays as there)
static void lsSearch(int istep, double *prod, double *eta, double amp, double maxeta, double mineta, int &ico)
{
    int ineg = 0;
    double etaneg = 1.0, etamax = 0.0;
    for (int i = 1; i <= istep; i++) {
        etamax = std::max(etamax, eta[i]);
        if (prod[i] >= 0.0 || eta[i] >= etaneg) continue;
        etaneg = eta[i];
        ineg = i;
    }
    if (ineg) {
        int ipos = 1;
        for (int i = 1; i <= istep; i++) {
            if (prod[i] <= 0.0 || eta[i] > eta[ineg] || eta[i] < eta[ipos]) continue;
            ipos = i;
        }
        double etaint = (prod[ineg] * eta[ipos] - prod[ipos] * eta[ineg]) / (prod[ineg] - prod[ipos]);
        etaint = std::max(etaint, eta[ipos] + 0.2 * (eta[ineg] - eta[ipos]));
        if (etaint < mineta) {
            etaint = mineta;
            ico = ico == 1 ? 2 : 1;
        }
        eta[istep + 1] = etaint;
        return;
    }
    const double etamaxstep = amp * etamax;
    const double etaextrap = (prod[istep] * eta[istep - 1] - prod[istep - 1] * eta[istep]) / (prod[istep] - prod[istep - 1]);
    eta[istep + 1] = (etaextrap <= 0.0 || etaextrap > etamaxstep) ? etamaxstep : etaextrap;
    if (eta[istep + 1] > maxeta && ico == 1) {
        ico = 2;
        return;
    }
    if (eta[istep + 1] > maxeta) {
        ico = 1;
        eta[istep + 1] = maxeta;
    }
}

// LineSearchNM::solve linesearch.C:54-160 along ddX from the current X (whose
// residual is in rhs), with the directly controlled equations masked out of
// the products (eqnmask): X ends at X_start + eta ddX and ddX is scaled by
// eta when the search meets its tolerance (accepted), else X ends at
// X_start + ddX.  Returns eta.
double StructuralProblem::lineSearch(int &lineSearches, bool &accepted)
{
    const double tol = std::min(std::max(nr.lsTol, 0.6), 0.95);
    const double amp = std::min(std::max(nr.lsAmp, 1.0), 10.0);
    const double maxeta = std::min(std::max(nr.lsMaxEta, 1.5), 15.0);
    const double s0 = -rhs.dot(ddX);
    double etav = 1.0;
    bool ok = false;
    if (s0 < 0.0) {
        rb.axpby(1.0, &X, 0.0);
        constexpr int maxIter = 10;
        double prod[maxIter + 2], eta[maxIter + 2];
        prod[1] = 1.0;
        eta[1] = 0.0;
        eta[2] = 1.0;
        int ico = 0;
        for (int ils = 2; ils <= maxIter; ils++) {
            X.axpby(1.0, &rb, 0.0);
            X.axpby(eta[ils], &ddX, 1.0);
            internalForces(X, true, nullptr);
            double sums[6];
            hipfemCheck(hfem_model_residual(ev, RT.ptr(), F.ptr(), gls.ptr(), sums), "residual");
            // the eqnmask equations (the directly controlled ones) leave the product
            const double zero = 0.0;
            for (int q : nr.ddmEq) hipfemCheck(hipfem_dev_upload(ctx, gls.ptr() + q - 1, &zero, 1), "ls mask");
            const double si = -gls.dot(ddX) / s0;
            prod[ils] = si;
            if (std::fabs(si) < tol) {
                etav = eta[ils];
                ok = true;
                break;
            }
            lsSearch(ils, prod, eta, amp, maxeta, 0.2, ico);
            if (ico == 2) break;
        }
        if (!ok) etav = 1.0;
        X.axpby(1.0, &rb, 0.0);
    }
    X.axpby(etav, &ddX, 1.0);
    if (etav != 1.0) {
        ddX.axpby(0.0, nullptr, etav);
        lineSearches++;
    }
    accepted = ok;
    return etav;
}

std::vector<StepRecord> StructuralProblem::solveTimeSteps(const double *up_ref, const double *f_free,
                                                          const double *f_presc, const std::vector<double> &ltf_t,
                                                          const std::vector<double> &ltf_v, double finalTime,
                                                          double dt, const ReductionStrategy &rs)
{
    if (!(finalTime > 0.0) || !(dt > 0.0)) OOFEM_ERROR("time stepping needs a positive final time and increment");
    if (ltf_t.size() != ltf_v.size()) OOFEM_ERROR("load time function: %d times, %d values", (int)ltf_t.size(),
                                                  (int)ltf_v.size());
    auto ltf = [&](double t) {
        if (ltf_t.empty()) return t / finalTime;
        if (t <= ltf_t.front()) return ltf_v.front();
        for (size_t i = 1; i < ltf_t.size(); i++)
            if (t <= ltf_t[i])
                return ltf_v[i - 1] + (ltf_v[i] - ltf_v[i - 1]) * (t - ltf_t[i - 1]) / (ltf_t[i] - ltf_t[i - 1]);
        return ltf_v.back();
    };
    std::vector<double> upv(npeq), ff(f_free ? neq : 0), fp(f_free && f_presc ? npeq : 0);
    std::vector<StepRecord> out;
    double t = currentTime;
    while (t < finalTime * (1.0 - 1.0e-12)) {
        Xc.axpby(1.0, &X, 0.0);
        int restarts = 0;
        double target = 0.0;
        NRStepReport rep;
        for (;;) {
            target = t + dt;
            const double f = ltf(target);
            for (int i = 0; i < npeq; i++) upv[i] = up_ref[i] * f;
            setPrescribedValues(upv.data());
            if (f_free) {
                for (int i = 0; i < neq; i++) ff[i] = f_free[i] * f;
                for (size_t i = 0; i < fp.size(); i++) fp[i] = f_presc[i] * f;
                setExternalLoad(ff.data(), fp.empty() ? nullptr : fp.data());
            }
            rep = solveNonlinearStep();
            if (rep.status == CR_CONVERGED || !rs.on) break;
            // ConvergenceException: restart with a reduced increment
            // (engngm.C:620-633, MetaStep::reduceTimeStep metastep.C:138-146)
            const double dt2 = dt * rs.ncrf;
            if (dt2 < rs.dtmin) OOFEM_ERROR("Minimum time increment reached");
            if (++restarts > rs.nmaxrestarts) OOFEM_ERROR("Maximum number of time step reductions has been reached.");
            dt = dt2;
            X.axpby(1.0, &Xc, 0.0);                       // StaticStructural::restartYourself
            hipfemCheck(hfem_model_restart(ev), "restart");  // EngngModel::initStepIncrements
        }
        StepRecord r;
        r.time = target;
        r.dt = dt;
        r.newton = rep.iterations;
        r.pcg = rep.linearIterations;
        r.status = (int)rep.status;
        r.restarts = restarts;
        r.lineSearches = rep.lineSearches;
        out.push_back(r);
        if (rs.on) {
            // MetaStep::adaptTimeStep metastep.C:149-183 with
            // SimpleReductionStrategy::giveReqIterTimeIncrementAdaptationFactor
            const int n = rep.iterations;
            double fac = 1.0;
            if (n != 0 && n < rs.minRequiredIter) fac = 2.0;
            else if (n != 0 && n > rs.maxRequiredIter) fac = 0.5;
            dt *= fac < 1.0 ? std::min(1.0, fac) : std::max(1.0, fac);
            dt = std::min(std::max(dt, rs.dtmin), rs.dtmax);
        }
        if (target + dt > finalTime) dt = finalTime - target;
        updateYourself();
        t = target;
        currentTime = t;
    }
    return out;
}

void StructuralProblem::setArcLength(const CALMOptions &o)
{
    // CylindricalALM::initializeFrom calmls.C:681-897
    if (!(o.maxStepLength > 0.0)) OOFEM_ERROR("calm: steplength must be positive, got %g", o.maxStepLength);
    if (o.manrmsteps < 0 || o.maxRestarts < 0 || o.minIterations < 0) OOFEM_ERROR("calm: negative count");
    if (o.rootSelection != 0 && o.rootSelection != 1) OOFEM_ERROR("calm: rootselectiontype must be 0 or 1");
    if (o.hpcMode < 0 || o.hpcMode > 2) OOFEM_ERROR("calm: hpcmode must be 0, 1 or 2");
    if (o.hpcMode && o.hpcEq.empty()) OOFEM_ERROR("HPC Map must be specified");
    if (o.hpcMode == 2 && o.hpcW.size() != o.hpcEq.size()) OOFEM_ERROR("HPC map size and weight array size mismatch");
    for (int q : o.hpcEq)
        if (q < 0 || q > neq) OOFEM_ERROR("calm: hpc equation %d out of range", q);
    calm = o;
    if (calm.psi < 0.0) calm.psi = 1.0;
    calm.maxiter = std::max(calm.maxiter, 30);
    if (calm.initialStepLength <= 0.0) calm.initialStepLength = calm.maxStepLength;
    calm.reqIterations = std::min(std::max(calm.reqIterations, 3), 1000);
    if (calmDeltaL <= 0.0 || calmDeltaL > calm.maxStepLength) calmDeltaL = calm.initialStepLength;
    calmMode = calm.manrmsteps > 0 ? 1 : 0;
    calmOldMode = calmMode;
    if (!calmSet) {
        for (DeviceVector *v : {&Rref, &deltaXt, &deltaXb, &dXs, &dXm1, &XInit, &oldDX}) v->allocate(ctx, neq);
        Rpref.allocate(ctx, npeq);
    }
    calmSet = true;
}

void StructuralProblem::setArcLengthHpc(int mode, const std::vector<int> &eqs, const std::vector<double> &w)
{
    if (!calmSet) OOFEM_ERROR("arc-length options must be set before the hyperplane control");
    CALMOptions o = calm;
    o.hpcMode = mode;
    o.hpcEq = eqs;
    o.hpcW = w;
    if (mode == 0) {
        o.hpcEq.clear();
        o.hpcW.clear();
    }
    const double keep = calmDeltaL;
    setArcLength(o);
    calmDeltaL = keep;
}

void StructuralProblem::setReferenceLoad(const double *f_free, const double *f_presc)
{
    if (!calmSet) OOFEM_ERROR("arc-length options must be set before the reference load");
    Rref.upload(f_free);
    if (npeq) Rpref.upload(f_presc);
}

// CylindricalALM::solve calmls.C:108-481 with computeDeltaLambda 995-1146
// and checkConvergence 484-677 (hyperplane control off, no line search,
// norms over all equations, no initial load vector R0).  rhs, ddX, X and the
// Gauss-point state are the driver's; the solver's own vectors persist in
// members as the reference's do.
CALMStepReport StructuralProblem::solveArcLengthStep()
{
    if (!calmSet) OOFEM_ERROR("arc-length control needs its options (steplength) first");
    ensureStructure();
    CALMStepReport rep;
    if (calmTick == 0) calmMode = calmOldMode;
    if (calmTick > 0) calmTick--;
    // nlinearstatic.C:523-539: the initial guess precedes the numerical method
    bool tangentCurrent = false;
    if (!applyInitialGuess(rep.linearIterations, tangentCurrent)) {
        rep.status = CR_FAILED;
        return rep;
    }
    auto solve = [&](const DeviceVector &b, DeviceVector &x) {
        const ConvergedReason r = linearSolve(b, x);   // x is the initial guess, as in IMLSolver
        rep.linearIterations += lastLinearIterations;
        return r == CR_CONVERGED;
    };
    XInit.axpby(1.0, &X, 0.0);
    deltaXt.zero();
    deltaXb.zero();
    const double RR = Rref.dot(Rref);
    const double psi2 = calm.psi * calm.psi;
    // hyperplane control: the controlled free equations (prescribed ones,
    // equation 0, are skipped as calmls.C does)
    std::vector<int> hpcOn;
    std::vector<double> hpcWOn, hX, hD, hB, hR;
    for (size_t i = 0; i < calm.hpcEq.size(); i++)
        if (calm.hpcEq[i] > 0) {
            hpcOn.push_back(calm.hpcEq[i]);
            hpcWOn.push_back(calm.hpcMode == 2 ? calm.hpcW[i] : 1.0);
        }
    auto gatherHpc = [&](const DeviceVector &v, std::vector<double> &out) {
        out.resize(hpcOn.size());
        for (size_t i = 0; i < hpcOn.size(); i++)
            hipfemCheck(hipfem_dev_download(ctx, &out[i], v.ptr() + hpcOn[i] - 1, 1), "hpc entry");
    };
    if (calm.hpcMode) gatherHpc(Rref, hR);
    double Lambda = reachedLambda, DeltaLambda = 0.0, deltaLambda = 0.0;
    int irest = 0, nite = 0;
    double ebe[3];
    for (;;) {
        // A: predictor from the tangent of the step's start
        dXs.zero();
        assembleNewtonMatrix();
        if (!solve(Rref, deltaXt)) { rep.status = CR_FAILED; return rep; }
        double p;
        if (calm.hpcMode == 0) {
            p = std::sqrt(deltaXt.dot(deltaXt) + psi2 * RR);
        } else if (calm.hpcMode == 1) {
            gatherHpc(deltaXt, hX);
            double xx = 0.0, rr = 0.0;
            for (size_t i = 0; i < hpcOn.size(); i++) {
                xx += hX[i] * hX[i];
                rr += hR[i] * hR[i];
            }
            p = std::sqrt(xx + psi2 * rr);
        } else {
            gatherHpc(deltaXt, hX);
            p = 0.0;
            for (size_t i = 0; i < hpcOn.size(); i++) p += hX[i] * hpcWOn[i];
        }
        Lambda = reachedLambda;
        const double XR = (calm.rootSelection == 0 || !haveOldDX) ? deltaXt.dot(Rref) : deltaXt.dot(oldDX);
        DeltaLambda = deltaLambda = (XR < 0.0 ? -1.0 : 1.0) * calmDeltaL / p;
        Lambda += DeltaLambda;
        rhs.axpby(DeltaLambda, &Rref, 0.0);
        if (!solve(rhs, dXs)) { rep.status = CR_FAILED; return rep; }
        X.axpby(1.0, &dXs, 1.0);
        nite = 0;
        internalForces(X, true, ebe);
        bool restart = false, converged = false;
        do {
            // B: corrector iterations
            nite++;
            dXm1.axpby(1.0, &dXs, 0.0);
            const double DeltaLambdam1 = DeltaLambda;
            if (calmMode == 2 || (calmMode == 1 && nite % calm.manrmsteps == 0)) {
                assembleNewtonMatrix();
                if (!solve(Rref, deltaXt)) { rep.status = CR_FAILED; return rep; }
            }
            rhs.axpby(Lambda, &Rref, 0.0);
            rhs.axpby(-1.0, &F, 1.0);
            if (!solve(rhs, deltaXb)) { rep.status = CR_FAILED; return rep; }
            // B.3: the cylindrical constraint |dX + ddX|^2 + psi^2 RR (DeltaLambda + dl)^2 = deltaL^2
            if (calm.hpcMode == 2) {
                // linearized control over the weighted hpc equations (calmls.C:1123-1143)
                gatherHpc(dXs, hD);
                gatherHpc(deltaXb, hB);
                gatherHpc(deltaXt, hX);
                double nom = 0.0, denom = 0.0;
                for (size_t i = 0; i < hpcOn.size(); i++) {
                    nom += (hD[i] + hB[i]) * hpcWOn[i];
                    denom += hX[i] * hpcWOn[i];
                }
                if (std::fabs(denom) < 1.e-20) OOFEM_ERROR("zero denominator in linearized control");
                deltaLambda = (calmDeltaL - nom) / denom;
            } else if (calm.hpcMode == 1) {
                // the constraint over the hpc equations only (calmls.C:1036-1058, 1082-1099)
                gatherHpc(dXs, hD);
                gatherHpc(deltaXb, hB);
                gatherHpc(deltaXt, hX);
                double rr = 0.0, RRh = 0.0, b2 = 0.0, b3 = 0.0, a4 = 0.0, a5 = 0.0;
                for (size_t i = 0; i < hpcOn.size(); i++) {
                    const double pr = hD[i] + hB[i];
                    rr += hX[i] * hX[i];
                    RRh += hR[i] * hR[i];
                    b2 += hX[i] * pr;
                    b3 += pr * pr;
                    a4 += hD[i] * hB[i] + hD[i] * hD[i];
                    a5 += hD[i] * hX[i];
                }
                const double a1 = rr + psi2 * RRh;
                const double a2 = RRh * psi2 * DeltaLambda * 2.0 + 2.0 * b2;
                const double a3 = b3 - calmDeltaL * calmDeltaL + DeltaLambda * DeltaLambda * RRh * psi2;
                double discr = a2 * a2 - 4.0 * a1 * a3;
                if (discr < 0.0) {
                    deltaLambda = 0.0;
                } else {
                    discr = std::sqrt(discr);
                    const double lam1 = (-a2 + discr) / 2. / a1, lam2 = (-a2 - discr) / 2. / a1;
                    if (calm.rootSelection == 0) {
                        const double cos1 = (a4 + a5 * lam1) / calmDeltaL / calmDeltaL;
                        const double cos2 = (a4 + a5 * lam2) / calmDeltaL / calmDeltaL;
                        deltaLambda = cos1 > cos2 ? lam1 : lam2;
                    } else {
                        // RST_Dot takes the full vectors, as calmls.C:1104-1118 does
                        const double XXt = dXs.dot(deltaXt), dXdX = dXs.dot(dXs), dXX_ = dXs.dot(deltaXb);
                        const double s = DeltaLambda + deltaLambda;
                        const double c1 = (dXdX + dXX_ + lam1 * XXt) + (psi2 * (s * s + s * lam1));
                        const double c2 = (dXdX + dXX_ + lam2 * XXt) + (psi2 * (s * s + s * lam2));
                        deltaLambda = c1 > c2 ? lam1 : lam2;
                    }
                }
            } else {
                const double XX = deltaXt.dot(deltaXt), XXt = dXs.dot(deltaXt), X_Xt = deltaXb.dot(deltaXt);
                const double dXdX = dXs.dot(dXs), dXX_ = dXs.dot(deltaXb), X_X_ = deltaXb.dot(deltaXb);
                const double a1 = XX + psi2 * RR;
                const double a2 = RR * psi2 * DeltaLambda * 2.0 + 2.0 * XXt + 2.0 * X_Xt;
                const double a3 = dXdX + 2.0 * dXX_ + X_X_ + DeltaLambda * DeltaLambda * RR * psi2 - calmDeltaL * calmDeltaL;
                double discr = a2 * a2 - 4.0 * a1 * a3;
                if (discr < 0.0) {
                    // calmls.C:1060-1064: a warning, and no load change in this iteration
                    deltaLambda = 0.0;
                } else {
                    discr = std::sqrt(discr);
                    const double lam1 = (-a2 + discr) / 2. / a1, lam2 = (-a2 - discr) / 2. / a1;
                    if (calm.rootSelection == 0) {
                        const double a4 = dXX_ + dXdX, a5 = XXt;
                        const double cos1 = (a4 + a5 * lam1) / calmDeltaL / calmDeltaL;
                        const double cos2 = (a4 + a5 * lam2) / calmDeltaL / calmDeltaL;
                        deltaLambda = cos1 > cos2 ? lam1 : lam2;
                    } else {
                        // RST_Dot uses the previous deltaLambda, as calmls.C:1108-1118 does
                        const double s = DeltaLambda + deltaLambda;
                        const double c1 = (dXdX + dXX_ + lam1 * XXt) + (psi2 * (s * s + s * lam1));
                        const double c2 = (dXdX + dXX_ + lam2 * XXt) + (psi2 * (s * s + s * lam2));
                        deltaLambda = c1 > c2 ? lam1 : lam2;
                    }
                }
            }
            // B.4-B.6
            ddX.axpby(deltaLambda, &deltaXt, 0.0);
            ddX.axpby(1.0, &deltaXb, 1.0);
            dXs.axpby(1.0, &dXm1, 0.0);
            dXs.axpby(1.0, &ddX, 1.0);
            X.axpby(1.0, &XInit, 0.0);
            X.axpby(1.0, &dXs, 1.0);
            const double drProduct = ddX.dot(ddX);
            DeltaLambda = DeltaLambdam1 + deltaLambda;
            Lambda = reachedLambda + DeltaLambda;
            internalForces(X, true, ebe);
            // B.7: convergence check
            rhs.axpby(Lambda, &Rref, 0.0);
            rhs.axpby(-1.0, &F, 1.0);
            double forceErr = rhs.dot(rhs);
            const double dXX = X.dot(X), eNorm = ebe[0] + ebe[1] + ebe[2];
            if (RR * Lambda * Lambda > 1.e-6)
                forceErr = std::sqrt(forceErr / (RR * Lambda * Lambda));
            else if (eNorm > 1.e-6)
                forceErr = std::sqrt(forceErr / eNorm);
            else
                forceErr = std::sqrt(forceErr);
            const double dispErr = dXX < 1.e-6 ? drProduct : std::sqrt(drProduct / dXX);
            const bool outOfRange = std::fabs(forceErr) > calm.rtolf * 1.e10 || std::fabs(dispErr) > calm.rtold * 1.e10 ||
                                    std::isnan(forceErr) || std::isnan(dispErr);
            converged = !(std::fabs(forceErr) > calm.rtolf || std::fabs(dispErr) > calm.rtold);
            rep.forceError = forceErr;
            rep.dispError = dispErr;
            if (nite >= calm.maxiter || outOfRange) {
                irest++;
                if (irest <= calm.maxRestarts) {
                    // restart the step with a quarter of the arc length and full Newton for two steps
                    calmDeltaL = std::max(calmDeltaL * 0.25, calm.minStepLength);
                    X.axpby(1.0, &XInit, 0.0);
                    hipfemCheck(hfem_model_restart(ev), "restart");   // EngngModel::initStepIncrements
                    calmOldMode = calmMode;
                    calmMode = 2;
                    calmTick = 2;
                    restart = true;
                    break;
                }
                // calmls.C:418-424: a warning; the step is then accepted as it stands
                rep.status = CR_DIVERGED_ITS;
                break;
            }
        } while (!converged || nite < calm.minIterations);
        if (!restart) break;
    }
    // arc length for the next step (calmls.C:450-470)
    if (nite > calm.reqIterations)
        calmDeltaL = calmDeltaL * calm.reqIterations / nite;
    else
        calmDeltaL = calmDeltaL * std::sqrt(std::sqrt((double)calm.reqIterations / (double)nite));
    if (calmDeltaL > calm.maxStepLength) calmDeltaL = calm.maxStepLength;
    if (calmDeltaL < calm.minStepLength) calmDeltaL = calm.minStepLength;
    if (rep.status == CR_UNKNOWN) rep.status = CR_CONVERGED;
    reachedLambda = Lambda;
    oldDX.axpby(1.0, &dXs, 0.0);
    haveOldDX = true;
    // external loads of the reached level (NonLinearStatic::computeExternalLoadReactionContribution)
    RT.axpby(Lambda, &Rref, 0.0);
    if (npeq) RTp.axpby(Lambda, &Rpref, 0.0);
    rep.iterations = nite;
    rep.restarts = irest;
    rep.lambda = Lambda;
    rep.stepLength = calmDeltaL;
    return rep;
}

void StructuralProblem::restartStep()
{
    X.axpby(1.0, &Xc, 0.0);
    hipfemCheck(hfem_model_restart(ev), "restart");
}

void StructuralProblem::updateYourself()
{
    hipfemCheck(hfem_model_commit(ev), "commit");
    upPrev.axpby(1.0, &up, 0.0);   // prescribed values of the equilibrated step
}

void StructuralProblem::setProfiling(bool on)
{
    profiling = on;
    profN64 = profN32 = 0;
    profMs64 = profMs32 = 0.0;
}

void StructuralProblem::giveProfile(int *n64, double *ms64, int *n32, double *ms32) const
{
    *n64 = profN64;
    *ms64 = profMs64;
    *n32 = profN32;
    *ms32 = profMs32;
}

void StructuralProblem::giveReactions(double *r)
{
    // StructuralEngngModel::computeReaction: f_int (prescribed numbering) - f_ext (prescribed numbering)
    DeviceVector R(ctx, npeq);
    R.axpby(1.0, &Fp, 0.0);
    R.axpby(-1.0, &RTp, 1.0);
    R.download(r);
}

}  // namespace oofem
