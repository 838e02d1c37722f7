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
  ``stiffmode`` (0 tangent, 1 secant, 2 initial elastic, 3 initial
  secant); the step's load is its
  reference load (total loads at its time, refloadmode 0, or their
  increment over the step, refloadmode 1) plus the sum of the earlier
  steps' reference loads, as NonLinearStatic::updateLoadVectors keeps it;
- ``NonLinearStatic`` with indirect control (``controlmode 0``, the
  reference's default): one CylindricalALM step per time step on the
  reference load of the first step (calmls.C:108-481), keywords psi,
  steplength, initialsteplength, minsteplength, maxiter, maxrestarts,
  reqiterations, miniter, manrmsteps, rtolv / rtolf / rtold,
  rootselectiontype, and the hyperplane control hpcmode / hpc / hpcw; the
  CALM line search and dof-group norms are refused, and so are incremental
  reference loads with it.
  Solver keywords may sit on the first metastep record.
StaticStructural with ``treductiontype simplereduction`` (dtmin, dtmax,
ncrf, nmaxrestarts, minrequirediter, maxrequirediter) steps adaptively:
a step whose Newton loop fails restarts from the converged state with a
reduced increment, and errorcheck step numbers count accepted steps.
NRSolver's direct displacement control (``ddm`` node/dof pairs, ``ddv``
values, ``ddltf`` function) is applied in both nonlinear drivers.

Its solver keywords are read as the reference reads them: smtype, lstype,
lstol, lsiter, lsprecond, rtolv, rtolf, rtold, miniter, maxiter, manrmsteps,
initialguess, and NRSolver's linesearch, lsearchtol, lsearchamp,
lsearchmaxeta, maxinc and constrainednrminiter.  The line search runs in the
reference's order (search, cap, then the correction added to X once more);
``lskeeppoint 1``, a keyword of this build only, keeps the search's point.
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

    def _has(key):
        return key in over or (ms is not None and ms.has(key)) or e.has(key)

    # IMLSolver / NRSolver / EngngModel defaults (imlsolver.C:67-101,
    # nrsolver.C:106-200, staticstructural.C:130)
    kw = dict(smtype=kwd("smtype", 0, int), lstype=kwd("lstype", 0, int), lstol=kwd("lstol", 1e-5),
              lsiter=kwd("lsiter", 200, int), lsprecond=kwd("lsprecond", 0, int))
    # NRSolver::initializeFrom nrsolver.C:133-148: rtolv sets both rtolf and
    # rtold, and rtolf / rtold given on their own override it
    rtolv = kwd("rtolv", None, float) if _has("rtolv") else None
    nr_rtolf = kwd("rtolf", 1e-3 if rtolv is None else rtolv)
    nr_rtold = kwd("rtold", 0.0 if rtolv is None else rtolv)
    if model == "staticstructural":
        kw.update(rtolf=nr_rtolf, maxiter=kwd("maxiter", 100000000, int),
                  manrmsteps=kwd("manrmsteps", 0, int), initial_guess=kwd("initialguess", 1, int))
    calm = None
    if model == "nonlinearstatic":
        # "controllmode" is the reference's older spelling (nlinearstatic.C:106)
        cmode = kwd("controlmode", kwd("controllmode", 0, int), int)
        if cmode not in (0, 1):
            raise ValueError(f"unsupported NonLinearStatic controlmode {cmode}")
        refload = kwd("refloadmode", 0, int)
        if refload not in (0, 1):
            raise ValueError(f"unknown NonLinearStatic refloadmode {refload}")
        if refload == 1 and cmode == 0:
            raise ValueError("incremental reference loads (refloadmode 1) with arc-length control are not part "
                             "of this build")
        kw.update(initial_guess=kwd("initialguess", 0, int), stiff_mode=kwd("stiffmode", 0, int))
        if cmode == 1:
            kw.update(rtolf=nr_rtolf, maxiter=kwd("maxiter", 100000000, int),
                      manrmsteps=kwd("manrmsteps", 0, int))
        else:
            # CylindricalALM::initializeFrom calmls.C:681-897
            if "nccdg" in over or (ms is not None and ms.has("nccdg")) or e.has("nccdg"):
                raise ValueError("CylindricalALM dof-group norms (nccdg) are not part of this build")
            if kwd("linesearch", 0, int):
                raise ValueError("the CylindricalALM line search is not part of this build")
            if not (("steplength" in over) or (ms is not None and ms.has("steplength")) or e.has("steplength")):
                raise ValueError("CylindricalALM needs steplength")
            rtol = kwd("rtolv", 1e-3)
            calm = dict(steplength=kwd("steplength", 0.0), psi=kwd("psi", 1.0), maxiter=kwd("maxiter", 30, int),
                        maxrestarts=kwd("maxrestarts", 4, int), minsteplength=kwd("minsteplength", 0.0),
                        initialsteplength=kwd("initialsteplength", 0.0), reqiterations=kwd("reqiterations", 3, int),
                        miniter=kwd("miniter", 0, int), manrmsteps=kwd("manrmsteps", 0, int),
                        rtolf=kwd("rtolf", rtol), rtold=kwd("rtold", rtol),
                        rootselectiontype=kwd("rootselectiontype", 0, int))
    d = prob.deck
    P = StructuralProblem(d, **kw)
    if model != "linearstatic" and calm is None:
        # NRSolver::initializeFrom nrsolver.C:106-200 and LineSearchNM's
        # defaults (linesearch.C:44-52; its initializeFrom clamps them)
        P.set_nr_extras(linesearch=bool(kwd("linesearch", 0, int)), ls_tol=kwd("lsearchtol", 0.8),
                        ls_amp=kwd("lsearchamp", 2.5), ls_maxeta=kwd("lsearchmaxeta", 4.0),
                        maxinc=kwd("maxinc", 1.0e20))
        if kwd("constrainednrminiter", 0, int):
            P.set_constrained_nr(kwd("constrainednrminiter", 0, int))
        # rtold, miniter (nrsolver.C:113, 138-148); lskeeppoint is this
        # build's keyword for the line search that keeps its own point
        P.set_nr_criteria(rtold=nr_rtold, miniter=kwd("miniter", 0, int),
                          ls_keep_point=bool(kwd("lskeeppoint", 0, int)))
    if calm is not None:
        P.set_arc_length(**calm)
        # hyperplane control (calmls.C:754-797): the hpc array holds (node
        # label, dof id) pairs; given without hpcmode it means mode 1
        rec = ms if ms is not None and ms.has("hpc") else e
        hmode = kwd("hpcmode", 0, int)
        if rec.has("hpc"):
            pairs = rec.array("hpc", int)
            if len(pairs) % 2:
                P.close()
                raise ValueError("HPC Map size must be even number, it contains pairs <node, nodeDof>")
            hmode = 2 if hmode == 2 else 1
            wsrc = (ms if ms is not None and ms.has("hpcw") else e).array("hpcw", float, default=[])
            if hmode == 2 and len(wsrc) != len(pairs) // 2:
                P.close()
                raise ValueError("HPC map size and weight array size mismatch")
            # convertHPCMap calmls.C:900-957: the dof-manager order of the domain
            eqs, ws = [], []
            for n, label in enumerate(prob.node_numbers):
                for k in range(len(pairs) // 2):
                    if pairs[2 * k] == int(label):
                        q = int(d.eqn[n, pairs[2 * k + 1] - 1])
                        eqs.append(q if q > 0 else 0)
                        ws.append(wsrc[k] if hmode == 2 else 1.0)
            P.set_arc_length_hpc(hmode, eqs, ws if hmode == 2 else None)
        elif hmode:
            P.close()
            raise ValueError("HPC Map must be specified")
    # NRSolver direct displacement control: ddm = (node label, dof id) pairs,
    # ddv = values, ddltf = load-time function of the values
    ddm = None
    if model != "linearstatic" and (ms is not None and ms.has("ddm") or e.has("ddm")):
        if calm is not None:
            P.close()
            raise ValueError("direct displacement control (ddm) belongs to NRSolver, not to CylindricalALM")
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
    # MetaStep treductiontype (metastep.C:56-61): SimpleReductionStrategy
    # (timestepreductionstrategy.C:95-110) drives StaticStructural's steps
    red = None
    if kwd("treductiontype", "noreduction", str).lower() == "simplereduction":
        if model != "staticstructural" or ddm is not None:
            P.close()
            raise ValueError("time-step reduction is built for StaticStructural without ddm")
        try:
            red = dict(dtmin=kwd("dtmin", None), dtmax=kwd("dtmax", None))
        except Exception:
            red = None
        if red is None or red["dtmin"] is None or red["dtmax"] is None:
            P.close()
            raise ValueError("SimpleReductionStrategy needs dtmin and dtmax")
        red.update(ncrf=kwd("ncrf", 0.5), nmaxrestarts=kwd("nmaxrestarts", 5, int),
                   minreq=kwd("minrequirediter", 3, int), maxreq=kwd("maxrequirediter", 8, int))
    R0 = (np.zeros(d.neq), np.zeros(d.npeq))   # NonLinearStatic's initialLoadVector (free, prescribed)
    results, iters = {}, {}
    f = open(out_path, "w") if out_path else None
    try:
        if f:
            write_header(f, prob)
        if red is not None:
            # EngngModel::solveYourself engngm.C:573-662: a failed Newton loop
            # restarts the step from the converged state with dt * ncrf; the
            # next increment follows the Newton iterations (metastep.C:138-183)
            dt = prob.time_of_step(1)
            final = prob.time_of_step(prob.nsteps)
            tcur, k = 0.0, 0
            while tcur < final * (1.0 - 1.0e-12):
                t0 = time.perf_counter()
                P.save_converged()
                restarts = 0
                while True:
                    target = tcur + dt
                    up, (ff, fp) = prob.loading_at(target)
                    P.set_prescribed(up)
                    P.set_load(ff, fp)
                    rep = P.nonlinear_step()
                    if rep["status"] == 1:
                        break
                    dt2 = dt * red["ncrf"]
                    if dt2 < red["dtmin"]:
                        raise RuntimeError("Minimum time increment reached")
                    restarts += 1
                    if restarts > red["nmaxrestarts"]:
                        raise RuntimeError("Maximum number of time step reductions has been reached.")
                    dt = dt2
                    P.restart_step()
                k += 1
                iters[k] = (rep["newton"], rep["pcg"])
                results[k] = _collect(prob, P, P.solution(), up)
                results[k].update(time=target, restarts=restarts)
                n = rep["newton"]
                fac = 2.0 if (n != 0 and n < red["minreq"]) else 0.5 if (n != 0 and n > red["maxreq"]) else 1.0
                dt = min(max(dt * fac, red["dtmin"]), red["dtmax"])
                if target + dt > final:
                    dt = final - target
                P.update()
                if f:
                    write_step(f, prob, k, target, results[k], time.perf_counter() - t0)
                if log:
                    log(f"[deck] step {k} (t = {target:g}, {restarts} reductions): {iters[k]}")
                tcur = target
        for k in range(1, prob.nsteps + 1 if red is None else 1):
            t0 = time.perf_counter()
            up, (ff, fp) = prob.step_loading(k)
            P.set_prescribed(up)
            if model == "nonlinearstatic" and calm is None:
                # direct control (nlinearstatic.C:396-430, 496-507, 552-556):
                # the reference load R of the step is the total loads at its
                # time (refloadmode 0) or their increment over the step
                # (refloadmode 1, Function::evaluate VM_Incremental), and
                # the solver's total load is R + R0, where R0 gathers
                # loadLevel (1 under NRSolver) times every earlier R
                if refload == 1:
                    f0, p0 = prob.loads_at(prob.time_of_step(k - 1))   # t - dt
                    ff, fp = ff - f0, fp - p0
                P.set_load(ff + R0[0], fp + R0[1])
                R0 = (R0[0] + ff, R0[1] + fp)
            elif calm is None:
                P.set_load(ff, fp)
            elif k == 1:
                # indirect control assembles the reference load once
                # (nlinearstatic.C:496-507, loadInitFlag): the total loads of the first step
                P.set_reference_load(ff, fp)
            if ddm is not None:
                # NRSolver::applyConstraintsToLoadIncrement nrsolver.C:512-520:
                # with total reference loads the increment of the function
                # over the step (its value at the first step), with
                # incremental ones (NonLinearStatic refloadmode 1) its value
                t = prob.time_of_step(k)
                if model == "nonlinearstatic" and refload == 1:
                    P.set_ddm_factor(ddm(t))
                else:
                    P.set_ddm_factor(ddm(t) - (ddm(prob.time_of_step(k - 1)) if k > 1 else 0.0))
            level = None
            if model == "linearstatic":
                it, _ = P.linear_static()
                iters[k] = (0, it)
            elif calm is not None:
                rep = P.arc_length_step()
                iters[k] = (rep["iterations"], rep["pcg"])
                level = rep["load_level"]
                if rep["status"] != 1 and log:
                    # the reference warns and accepts the step (calmls.C:418-424)
                    log(f"[deck] step {k}: CylindricalALM did not converge after {rep['restarts']} restarts")
            else:
                rep = P.nonlinear_step()
                iters[k] = (rep["newton"], rep["pcg"])
                if rep["status"] != 1:
                    raise RuntimeError(f"step {k}: Newton loop did not converge: {rep}")
            results[k] = _collect(prob, P, P.solution(), up)
            if level is not None:
                results[k]["load_level"] = level
                results[k]["iterations"] = iters[k][0]
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
