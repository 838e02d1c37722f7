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

ConvergedReason HipPCGSolver::solveDevice(HipCSRSparseMtrx &A, const double *db, double *dx)
{
    int reason = 0;
    hipfemCheck(hippcg_solve_dev(s, A.handle(), db, dx, &reason, &lastIterations, &lastResidual), "hippcg_solve_dev");
    return reason == HIPFEM_CR_CONVERGED ? CR_CONVERGED : CR_DIVERGED_ITS;
}

}  // namespace oofem
