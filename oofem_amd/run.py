"""Run one of the reference's input decks end to end on the GPU path: the
deck reader, the C++ host drivers behind the hipcsr / hippcg plug-ins, the
.out result file and the deck's errorcheck block.

The engineering-model record chooses the driver:
- ``LinearStatic``: one linear solve per step with that step's prescribed
  values and loads (LinearStatic::solveYourselfAt, linearstatic.C:177-252);
- ``StaticStructural``: one NRSolver step per time step
  (StaticStructural::solveYourselfAt, staticstructural.C:224-358);
- ``NonLinearStatic`` with direct control (``controlmode 1``): one NRSolver
  step per time step on the total loads of that time
  (NonLinearStatic::proceedStep, nlinearstatic.C:453-567; refloadmode 0),
  no initial guess unless ``initialguess 1``, and the Newton matrix from
  ``stiffmode`` (0 tangent, 2 initial elastic).  Arc-length control
  (controlmode 0, CylindricalALM), incremental reference loads
  (refloadmode 1) and the secant stiffness are refused.  Solver keywords may
  sit on the first metastep record.
NRSolver's direct displacement control (``ddm`` node/dof pairs, ``ddv``
values, ``ddltf`` function) is applied in both nonlinear drivers.

Its solver keywords are read as the reference reads them: smtype, lstype,
lstol, lsiter, lsprecond, rtolf, maxiter, manrmsteps, initialguess.
smtype and lstype select the plug-ins exactly as the class factory does, so
a deck runs on this path when it says ``smtype 11 lstype 9``; any other type
is rejected by the factory with an error.  ``keywords`` adds or replaces
engineering-model keywords as an edit of the deck's record would (the
"only the keywords change" switch, e.g. smtype 11 lstype 9 lstol 1e-12 for a
deck written for the reference's default direct solver).
"""
from __future__ import annotations

import time

import numpy as np

from .deck import evaluate_checks, read_deck
from .meshgen import NGP
from .outfile import write_header, write_step


def _collect(prob, P, u, up):
    d = prob.deck
    u_node = np.zeros((d.nnode, 3))
    m = d.eqn > 0
    u_node[m] = u[d.eqn[m] - 1]
    m = d.eqn < 0
    u_node[m] = up[-d.eqn[m] - 1]
    R = P.reactions()
    reac = {(int(n), int(k) + 1): float(R[-d.eqn[n, k] - 1]) for n, k in zip(*np.nonzero(d.eqn < 0))}
    st = P.state().reshape(d.nelem, NGP[d.etype], -1)
    return dict(u_node=u_node, reactions=reac, gp_sig=st[:, :, 20:26], gp_eps=st[:, :, 26:32], u=u, state=st)


def run_deck(path, out_path=None, keywords=None, log=None):
    """Solve every step of the deck at `path`; write the .out file to
    out_path when given.  keywords: engineering-model keywords that replace
    or extend the deck's.  Returns dict(results {step: collected
    quantities}, checks [(check, value, ok)], iterations {step: (newton,
    pcg)})."""
    from .problem import StructuralProblem
    prob = read_deck(path)
    e = prob.engng
    model = prob.model
    if model not in ("linearstatic", "staticstructural", "nonlinearstatic"):
        raise ValueError(f"engineering model '{model}' is not part of this build")
    over = {k.lower(): v for k, v in (keywords or {}).items()}
    ms = prob.metasteps[0] if getattr(prob, "metasteps", None) else None

    def kwd(key, default, cast=float):
        if key in over:
            return cast(over[key])
        if ms is not None and ms.has(key):
            return ms.scalar(key, default, cast)
        return e.scalar(key, default, cast)

    # IMLSolver / NRSolver / EngngModel defaults (imlsolver.C:67-101,
    # nrsolver.C:106-200, staticstructural.C:130)
    kw = dict(smtype=kwd("smtype", 0, int), lstype=kwd("lstype", 0, int), lstol=kwd("lstol", 1e-5),
              lsiter=kwd("lsiter", 200, int), lsprecond=kwd("lsprecond", 0, int))
    if model == "staticstructural":
        kw.update(rtolf=kwd("rtolf", 1e-3), maxiter=kwd("maxiter", 100000000, int),
                  manrmsteps=kwd("manrmsteps", 0, int), initial_guess=kwd("initialguess", 1, int))
    elif model == "nonlinearstatic":
        if kwd("controlmode", 0, int) != 1:
            raise ValueError("NonLinearStatic arc-length control (controlmode 0, CylindricalALM) is not part of this "
                             "build; use controlmode 1")
        if kwd("refloadmode", 0, int) != 0:
            raise ValueError("NonLinearStatic incremental reference loads (refloadmode 1) are not part of this build")
        kw.update(rtolf=kwd("rtolf", 1e-3), maxiter=kwd("maxiter", 100000000, int),
                  manrmsteps=kwd("manrmsteps", 0, int), initial_guess=kwd("initialguess", 0, int),
                  stiff_mode=kwd("stiffmode", 0, int))
    d = prob.deck
    P = StructuralProblem(d, **kw)
    # NRSolver direct displacement control: ddm = (node label, dof id) pairs,
    # ddv = values, ddltf = load-time function of the values
    ddm = None
    if model != "linearstatic" and (ms is not None and ms.has("ddm") or e.has("ddm")):
        rec = ms if ms is not None and ms.has("ddm") else e
        pairs = rec.array("ddm", int)
        vals = rec.array("ddv")
        if len(pairs) != 2 * len(vals):
            P.close()
            raise ValueError("direct displacement mask size mismatch")
        nid = {int(n): i for i, n in enumerate(prob.node_numbers)}
        eqs = [int(d.eqn[nid[pairs[2 * k]], pairs[2 * k + 1] - 1]) for k in range(len(vals))]
        if min(eqs) <= 0:
            P.close()
            raise ValueError("ddm names a dof that is prescribed by a boundary condition")
        P.set_ddm(eqs, vals)
        ddm = prob.ltf[rec.scalar("ddltf", cast=int)]
    results, iters = {}, {}
    f = open(out_path, "w") if out_path else None
    try:
        if f:
            write_header(f, prob)
        for k in range(1, prob.nsteps + 1):
            t0 = time.perf_counter()
            up, (ff, fp) = prob.step_loading(k)
            P.set_prescribed(up)
            P.set_load(ff, fp)
            if ddm is not None:
                # NRSolver::applyConstraintsToLoadIncrement: the increment of
                # the function over the step (its value at the first step)
                t = prob.time_of_step(k)
                P.set_ddm_factor(ddm(t) - (ddm(prob.time_of_step(k - 1)) if k > 1 else 0.0))
            if model == "linearstatic":
                it, _ = P.linear_static()
                iters[k] = (0, it)
            else:
                rep = P.nonlinear_step()
                iters[k] = (rep["newton"], rep["pcg"])
                if rep["status"] != 1:
                    raise RuntimeError(f"step {k}: Newton loop did not converge: {rep}")
            results[k] = _collect(prob, P, P.solution(), up)
            if model != "linearstatic":
                P.update()
            if f:
                write_step(f, prob, k, prob.time_of_step(k), results[k], time.perf_counter() - t0)
            if log:
                log(f"[deck] step {k}: {iters[k]}")
    finally:
        if f:
            f.close()
        P.close()
    return dict(results=results, checks=evaluate_checks(prob, results), iterations=iters)
