#include "hippcg.h"

#include "classfactory.h"

namespace oofem {

REGISTER_SparseLinSolver(HipPCGSolver, ST_HipPCG)

HipPCGSolver::HipPCGSolver(Domain *d, EngngModel *m) : SparseLinearSystemNM(d, m), ctx(hipfemDefaultContext())
{
    hipfemCheck(hippcg_create(ctx, &s), "hippcg_create");
    applyOptions();
}

HipPCGSolver::~HipPCGSolver()
{
    if (s) hippcg_destroy(s);
}

void HipPCGSolver::applyOptions()
{
    hipfemCheck(hippcg_set_options(s, tol, maxite, precond), "hippcg_set_options");
    hipfemCheck(hippcg_set_precision(s, mixed, irtol, 0), "hippcg_set_precision");
    hipfemCheck(hippcg_set_reliable(s, reldelta), "hippcg_set_reliable");
}

void HipPCGSolver::setTolerance(double t)
{
    if (!(t > 0.0)) OOFEM_ERROR("tolerance must be positive, got %g", t);
    tol = t;
    hipfemCheck(hippcg_set_options(s, tol, maxite, precond), "hippcg_set_options");
}

// IMLSolver::initializeFrom imlsolver.C:67-97: the same keywords, defaults
// and reading macros.  IMLSolver's stype 1 (GMRES) and its ILU / IC
// preconditioners (lsprecond 2-4) are not part of hippcg and are refused.
void HipPCGSolver::initializeFrom(InputRecord &ir)
{
    int val = 0;
    IR_GIVE_OPTIONAL_FIELD(ir, val, _IFT_IMLSolver_stype);
    if (val != 0) throw ValueInputException(ir, _IFT_IMLSolver_stype, "hippcg implements the CG solver (stype 0) only");

    tol = 1.e-5;
    IR_GIVE_OPTIONAL_FIELD(ir, tol, _IFT_IMLSolver_lstol);
    maxite = 200;
    IR_GIVE_OPTIONAL_FIELD(ir, maxite, _IFT_IMLSolver_lsiter);
    val = 0;
    IR_GIVE_OPTIONAL_FIELD(ir, val, _IFT_IMLSolver_lsprecond);
    if (val != 0 && val != 1) throw ValueInputException(ir, _IFT_IMLSolver_lsprecond, "unknown preconditioner type");
    precond = val;

    mixed = 0;
    IR_GIVE_OPTIONAL_FIELD(ir, mixed, _IFT_HipPCGSolver_mixed);
    if (mixed < 0 || mixed > 2) throw ValueInputException(ir, _IFT_HipPCGSolver_mixed, "mixed is 0, 1 or 2");
    irtol = 1e-4;
    IR_GIVE_OPTIONAL_FIELD(ir, irtol, _IFT_HipPCGSolver_irtol);
    reldelta = 0.1;
    IR_GIVE_OPTIONAL_FIELD(ir, reldelta, _IFT_HipPCGSolver_reldelta);
    applyOptions();
}

ConvergedReason HipPCGSolver::solve(SparseMtrx &A, FloatArray &b, FloatArray &x)
{
    if (x.giveSize() != b.giveSize()) OOFEM_ERROR("size mismatch");
    auto *H = dynamic_cast<HipCSRSparseMtrx *>(&A);
    if (!H) OOFEM_ERROR("hippcg needs a HipCSRSparseMtrx (smtype %d), got %s", (int)SMT_HipCSR, A.giveClassName());
    int reason = 0;
    hipfemCheck(hippcg_solve(s, H->handle(), b.givePointer(), x.givePointer(), &reason, &lastIterations, &lastResidual),
                "hippcg_solve");
    return reason == HIPFEM_CR_CONVERGED ? CR_CONVERGED : CR_DIVERGED_ITS;
}

// SparseLinearSystemNM::solve(A, B, X) sparselinsystemnm.C:47-70
ConvergedReason SparseLinearSystemNM::solve(SparseMtrx &A, FloatMatrix &B, FloatMatrix &X)
{
    const int neq = A.giveNumberOfRows(), nrhs = B.giveNumberOfColumns();
    if (A.giveNumberOfRows() != B.giveNumberOfRows()) OOFEM_ERROR("A and B matrix mismatch");
    if (X.giveNumberOfRows() != neq || X.giveNumberOfColumns() != nrhs) X.resize(neq, nrhs);
    ConvergedReason status = CR_UNKNOWN;
    FloatArray bi(neq), xi(neq);
    for (int j = 1; j <= nrhs; ++j) {
        for (int i = 1; i <= neq; ++i) {
            bi.at(i) = B.at(i, j);
            xi.at(i) = X.at(i, j);
        }
        status = this->solve(A, bi, xi);
        if (status != CR_CONVERGED) return status;
        for (int i = 1; i <= neq; ++i) X.at(i, j) = xi.at(i);
    }
    return status;
}

// several right-hand sides at once on the device (hippcg_solve_multi: six
// columns per batched launch sequence with precision 2, else column by column)
ConvergedReason HipPCGSolver::solve(SparseMtrx &A, FloatMatrix &B, FloatMatrix &X)
{
    auto *H = dynamic_cast<HipCSRSparseMtrx *>(&A);
    if (!H) OOFEM_ERROR("hippcg needs a HipCSRSparseMtrx (smtype %d), got %s", (int)SMT_HipCSR, A.giveClassName());
    const int neq = A.giveNumberOfRows(), nrhs = B.giveNumberOfColumns();
    if (B.giveNumberOfRows() != neq) OOFEM_ERROR("A and B matrix mismatch");
    if (X.giveNumberOfRows() != neq || X.giveNumberOfColumns() != nrhs) X.resize(neq, nrhs);
    std::vector<double> b((size_t)neq * nrhs), x((size_t)neq * nrhs);
    for (int j = 1; j <= nrhs; ++j)
        for (int i = 1; i <= neq; ++i) {
            b[(size_t)(j - 1) * neq + i - 1] = B.at(i, j);
            x[(size_t)(j - 1) * neq + i - 1] = X.at(i, j);
        }
    int reason = 0;
    hipfemCheck(hippcg_solve_multi(s, H->handle(), nrhs, b.data(), x.data(), &reason, &lastIterations, &lastResidual),
                "hippcg_solve_multi");
    for (int j = 1; j <= nrhs; ++j)
        for (int i = 1; i <= neq; ++i) X.at(i, j) = x[(size_t)(j - 1) * neq + i - 1];
    return reason == HIPFEM_CR_CONVERGED ? CR_CONVERGED : CR_DIVERGED_ITS;
}

ConvergedReason HipPCGSolver::solveDevice(HipCSRSparseMtrx &A, const double *db, double *dx)
{
    int reason = 0;
    hipfemCheck(hippcg_solve_dev(s, A.handle(), db, dx, &reason, &lastIterations, &lastResidual), "hippcg_solve_dev");
    return reason == HIPFEM_CR_CONVERGED ? CR_CONVERGED : CR_DIVERGED_ITS;
}

}  // namespace oofem
