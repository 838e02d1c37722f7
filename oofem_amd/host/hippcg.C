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

// IMLSolver::initializeFrom imlsolver.C:67-101 keywords and defaults
void HipPCGSolver::initializeFrom(InputRecord &ir)
{
    if (ir.has("stype") && (int)ir.get("stype", 0) != 0)
        OOFEM_ERROR("hippcg implements the CG solver type (stype 0) only");
    tol = ir.get("lstol", 1.e-5);
    maxite = (int)ir.get("lsiter", 200);
    precond = (int)ir.get("lsprecond", 0);
    if (precond != 0 && precond != 1) OOFEM_ERROR("unknown preconditioner type %d", precond);
    mixed = (int)ir.get("mixed", 0);
    irtol = ir.get("irtol", 1e-4);
    reldelta = ir.get("reldelta", 0.1);
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
