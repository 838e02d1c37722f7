"""Problem drivers over the C ABI, following the reference's engineering models.

``linear_static`` restates LinearStatic::solveYourselfAt (linearstatic.C:177-252)
and StructuralEngngModel::computeReaction (structengngmodel.C:183-201):
build the K pattern once, assemble K through the batched evaluator,
rhs = f_ext - f_int(prescribed values), solve with hippcg from x = 0,
reactions = f_int(prescribed numbering) - f_ext(prescribed numbering).
"""
from __future__ import annotations

import time

import numpy as np

from .hipfem import BatchedEvaluator, Context, HipCSRSparseMtrx, HipPCGSolver, CR_CONVERGED


class LinearStaticProblem:
    """One deck on one GPU: evaluator, K (hipcsr) and solver (hippcg)."""

    def __init__(self, ctx: Context, deck, lstol=1e-10, lsiter=100000, precision=0, inner_tol=1e-4):
        self.ctx = ctx
        self.deck = deck
        self.K = HipCSRSparseMtrx(ctx)
        self.ev = BatchedEvaluator(ctx, deck)
        self.solver = HipPCGSolver(ctx, lstol=lstol, lsiter=lsiter, lsprecond=1, precision=precision,
                                   inner_tol=inner_tol)
        t = time.perf_counter()
        self.ev.build_structure(self.K)
        self.t_structure = time.perf_counter() - t
        self.f_ext, self.f_ext_presc = deck.external_load()
        self.u_presc = deck.prescribed_values()

    def solve(self):
        d = self.deck
        timings = {}
        t0 = time.perf_counter()
        self.ev.assemble_stiffness(self.K)
        self.ctx.synchronize()
        timings["assemble"] = time.perf_counter() - t0
        t1 = time.perf_counter()
        fint, _, _ = self.ev.internal_forces(np.zeros(d.neq), self.u_presc)
        rhs = self.f_ext - fint
        u = np.zeros(d.neq)
        reason, iters, res = self.solver.solve(self.K, rhs, u)
        timings["solve"] = time.perf_counter() - t1
        t2 = time.perf_counter()
        _, fp, _ = self.ev.internal_forces(u, self.u_presc, update_state=True)
        reactions = fp - self.f_ext_presc
        timings["reactions"] = time.perf_counter() - t2
        if reason != CR_CONVERGED:
            raise RuntimeError(f"No success in solving system (iterations {iters}, residual {res:.3e})")
        return dict(u=u, reactions=reactions, iters=iters, resid=res, timings=timings,
                    outer=self.solver.last_outer())


def linear_static(ctx, deck, **kw):
    return LinearStaticProblem(ctx, deck, **kw).solve()


def reaction_sums(deck, reactions):
    """Sum of reactions per direction (u, v, w)."""
    out = np.zeros(3)
    m = deck.eqn < 0
    dirs = np.nonzero(m)[1]
    peq = -deck.eqn[m] - 1
    for k in range(3):
        out[k] = reactions[peq[dirs == k]].sum()
    return out
