// HipPCGSolver: the SparseLinearSystemNM of type ST_HipPCG (reference
// interface sparselinsystemnm.h:56-97).  Reads the IMLSolver keywords with
// IML semantics (imlsolver.C:67-101: lstol default 1e-5, lsiter default 200,
// lsprecond 0 none / 1 diagonal) plus `mixed` (0 FP64 PCG, 1 FP32 inner PCG
// in FP64 iterative refinement, 2 FP32 PCG with FP64 reliable updates),
// `irtol` (inner relative tolerance of 1) and `reldelta` (reliable-update
// factor of 2, default 0.1).
#pragma once

#include "hipcsr.h"
#include "imlsolver.h"

// hippcg's own keywords (no counterpart in IMLSolver)
#define _IFT_HipPCGSolver_mixed "mixed"
#define _IFT_HipPCGSolver_irtol "irtol"
#define _IFT_HipPCGSolver_reldelta "reldelta"

namespace oofem {

class HipPCGSolver : public SparseLinearSystemNM {
    hipfem_ctx *ctx;
    hippcg_solver *s = nullptr;
    double tol = 1.e-5;
    int maxite = 200;
    int precond = 0;
    int mixed = 0;
    double irtol = 1e-4;
    double reldelta = 0.1;
    int lastIterations = 0;
    double lastResidual = 0.0;

    void applyOptions();

public:
    HipPCGSolver(Domain *d, EngngModel *m);
    ~HipPCGSolver() override;
    void initializeFrom(InputRecord &ir) override;
    ConvergedReason solve(SparseMtrx &A, FloatArray &b, FloatArray &x) override;
    ConvergedReason solve(SparseMtrx &A, FloatMatrix &B, FloatMatrix &X) override;
    // device-resident form: b and x are device vectors of the matrix's context
    ConvergedReason solveDevice(HipCSRSparseMtrx &A, const double *db, double *dx);
    LinSystSolverType giveLinSystSolverType() const override { return ST_HipPCG; }
    SparseMtrxType giveRecommendedMatrix(bool symmetric) const override { return SMT_HipCSR; }
    const char *giveClassName() const override { return "HipPCGSolver"; }

    // relative residual tolerance of the following solves (lstol)
    void setTolerance(double t);
    double giveTolerance() const { return tol; }
    int giveLastIterations() const { return lastIterations; }
    double giveLastResidual() const { return lastResidual; }
    hippcg_solver *handle() { return s; }
};

}  // namespace oofem
