"""Python front end of the C++ host layer (libhipfem_host.so).

``StructuralProblem`` wraps the host engineering model: the sparse matrix
and the linear solver are created by the class factory from deck integers
(``smtype`` 11 = hipcsr, ``lstype`` 9 = hippcg), assembly goes through the
batched evaluator, and the two drivers are the C++ restatements of
LinearStatic::solveYourselfAt and StaticStructural::solveYourselfAt +
NRSolver::solve.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build
from .hipfem import lib as _devlib

SMT_HipCSR = 11
ST_HipPCG = 9
CR_CONVERGED = 1

_hl = None


class ProblemError(RuntimeError):
    pass


def hostlib():
    global _hl
    if _hl is None:
        if not os.path.exists(_build.HOST_LIB):
            raise ProblemError(f"{_build.HOST_LIB} is missing: build it with __graft_entry__.build()")
        _devlib()  # libhipfem.so first (the host library links it)
        L = C.CDLL(_build.HOST_LIB)
        L.hfem_problem_last_error.restype = C.c_char_p
        L.hfem_problem_destroy.restype = None
        L.hfem_problem_destroy.argtypes = [C.c_void_p]
        L.hfem_problem_set_options.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                               C.c_double, C.c_double, C.c_int, C.c_int]
        L.hfem_problem_set_forcing.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.hfem_problem_set_stiffness_mode.argtypes = [C.c_void_p, C.c_int]
        L.hfem_problem_set_ddm.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.hfem_problem_set_ddm_factor.argtypes = [C.c_void_p, C.c_double]
        L.hfem_problem_set_constrained_nr.argtypes = [C.c_void_p, C.c_int]
        L.hfem_problem_save_converged.argtypes = [C.c_void_p]
        L.hfem_problem_restart_step.argtypes = [C.c_void_p]
        L.hfem_problem_set_nr_extras.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.hfem_problem_set_nr_criteria.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int]
        L.hfem_problem_nr_errors.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.hfem_problem_update_internal_rhs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.hfem_problem_assemble.argtypes = [C.c_void_p, C.c_int]
        L.hfem_problem_set_arc_length.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double,
                                                  C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                                  C.c_int]
        L.hfem_problem_set_reference_load.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.hfem_problem_set_arc_length_hpc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.hfem_problem_arc_length_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.hfem_problem_run_steps.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                             C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p]
        _hl = L
    return _hl


def _check(rc, what):
    if rc != 0:
        raise ProblemError(f"{what}: {hostlib().hfem_problem_last_error().decode(errors='replace')}")


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class StructuralProblem:
    def __init__(self, deck, smtype=SMT_HipCSR, lstype=ST_HipPCG, lstol=1e-10, lsiter=100000, lsprecond=1,
                 mixed=0, irtol=1e-4, rtolf=1e-8, maxiter=50, manrmsteps=1, initial_guess=1, forcing=0.0,
                 stiff_mode=0):
        """forcing > 0: inexact Newton, corrections solved to forcing terms
        of at most this value (hfem_problem_set_forcing); 0: every
        correction to lstol, as the reference's NRSolver.  stiff_mode:
        NonLinearStatic's stiffmode (0 tangent, 1 secant, 2 initial elastic,
        3 initial secant)."""
        self.deck = d = deck
        self._keep = [np.ascontiguousarray(x) for x in (d.coords, d.conn, d.emat, d.mat, d.eqn)]
        c, cn, em, mt, eq = self._keep
        self.h = C.c_void_p()
        _check(hostlib().hfem_problem_create(d.nnode, _p(c), d.etype, d.nelem, _p(cn), _p(em), mt.shape[0], _p(mt),
                                             _p(eq), d.neq, d.npeq, C.byref(self.h)), "create")
        _check(hostlib().hfem_problem_set_options(self.h, smtype, lstype, lstol, lsiter, lsprecond, mixed, irtol,
                                                  rtolf, maxiter, manrmsteps), "set_options")
        _check(hostlib().hfem_problem_set_initial_guess(self.h, int(initial_guess)), "set_initial_guess")
        _check(hostlib().hfem_problem_set_forcing(self.h, int(forcing > 0), float(forcing) if forcing > 0 else 0.1),
               "set_forcing")
        _check(hostlib().hfem_problem_set_stiffness_mode(self.h, int(stiff_mode)), "set_stiffness_mode")

    def set_ddm(self, eqs, values):
        """NRSolver direct displacement control on free equations (1-based)."""
        e = np.ascontiguousarray(eqs, dtype=np.int32)
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(hostlib().hfem_problem_set_ddm(self.h, len(e), _p(e) if len(e) else None, _p(v) if len(v) else None),
               "set_ddm")

    def set_constrained_nr(self, miniter):
        """NRSolver constrainednrminiter (0 = off)."""
        _check(hostlib().hfem_problem_set_constrained_nr(self.h, int(miniter)), "set_constrained_nr")

    def set_ddm_factor(self, factor):
        _check(hostlib().hfem_problem_set_ddm_factor(self.h, float(factor)), "set_ddm_factor")

    def set_nr_extras(self, linesearch=False, ls_tol=0.8, ls_amp=2.5, ls_maxeta=4.0, maxinc=1e20):
        """NRSolver line search (linesearch / lsearchtol / lsearchamp /
        lsearchmaxeta) and increment cap (maxinc)."""
        _check(hostlib().hfem_problem_set_nr_extras(self.h, int(bool(linesearch)), ls_tol, ls_amp, ls_maxeta, maxinc),
               "set_nr_extras")

    def set_nr_criteria(self, rtold=0.0, miniter=0, ls_keep_point=False):
        """NRSolver rtold and miniter; ls_keep_point: cap before the line
        search and keep its point (the reference caps after it and adds the
        searched step once more)."""
        _check(hostlib().hfem_problem_set_nr_criteria(self.h, float(rtold), int(miniter), int(bool(ls_keep_point))),
               "set_nr_criteria")

    def nr_errors(self, cap=100000):
        """Every convergence check of the last nonlinear step: rows of the
        u, v, w force errors and displacement errors (-1: group unused or
        criterion off)."""
        n = C.c_int(0)
        out = np.zeros((cap, 6))
        _check(hostlib().hfem_problem_nr_errors(self.h, _p(out), cap, C.byref(n)), "nr_errors")
        return out[:min(cap, n.value)].copy()

    def update_internal_rhs(self):
        """EngngModel::updateInternalRHS through the base-class hook: (f, eNorm)."""
        f = np.zeros(self.deck.neq)
        e = np.zeros(3)
        _check(hostlib().hfem_problem_update_internal_rhs(self.h, _p(f), _p(e)), "update_internal_rhs")
        return f, e

    def assemble(self, response_mode=0):
        """EngngModel::assemble through the base-class hook with
        TangentAssembler(response_mode): 0 tangent, 1 secant, 2 elastic."""
        _check(hostlib().hfem_problem_assemble(self.h, int(response_mode)), "assemble")

    def run_steps(self, up_ref, final_time, dt, reduction=None, f_ref=None, ltf=None, max_records=10000):
        """StaticStructural time stepping to final_time (hfem_problem_run_steps).
        reduction: None or dict(dtmin, dtmax, ncrf=0.5, nmaxrestarts=5,
        minrequirediter=3, maxrequirediter=8); f_ref: (f_free, f_presc) or
        None; ltf: (times, values) or None (t / final_time).  Returns a list
        of dict(time, dt, newton, pcg, status, restarts, line_searches)."""
        d = self.deck
        up = np.ascontiguousarray(up_ref, dtype=np.float64) if d.npeq else None
        ff = fp = None
        if f_ref is not None:
            ff = np.ascontiguousarray(f_ref[0], dtype=np.float64)
            fp = np.ascontiguousarray(f_ref[1], dtype=np.float64) if f_ref[1] is not None and d.npeq else None
        tt = vv = None
        nl = 0
        if ltf is not None:
            tt = np.ascontiguousarray(ltf[0], dtype=np.float64)
            vv = np.ascontiguousarray(ltf[1], dtype=np.float64)
            nl = len(tt)
        red = dict(dtmin=0.0, dtmax=1e300, ncrf=0.5, nmaxrestarts=5, minrequirediter=3, maxrequirediter=8)
        if reduction is not None:
            red.update(reduction)
        rt = np.zeros(max_records)
        rd = np.zeros(max_records)
        ri = np.zeros((max_records, 6), dtype=np.int32)
        n = C.c_int(0)
        _check(hostlib().hfem_problem_run_steps(self.h, _p(up), _p(ff), _p(fp), nl, _p(tt), _p(vv), float(final_time),
                                                float(dt), int(reduction is not None), float(red["dtmin"]),
                                                float(red["dtmax"]), float(red["ncrf"]), int(red["nmaxrestarts"]),
                                                int(red["minrequirediter"]), int(red["maxrequirediter"]), max_records,
                                                _p(rt), _p(rd), _p(ri), C.byref(n)), "run_steps")
        return [dict(time=float(rt[i]), dt=float(rd[i]), newton=int(ri[i, 0]), pcg=int(ri[i, 1]),
                     status=int(ri[i, 2]), restarts=int(ri[i, 3]), line_searches=int(ri[i, 4]))
                for i in range(min(n.value, max_records))]

    def close(self):
        if self.h:
            hostlib().hfem_problem_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_prescribed(self, up):
        up = np.ascontiguousarray(up, dtype=np.float64)
        _check(hostlib().hfem_problem_set_prescribed(self.h, _p(up) if up.size else None), "set_prescribed")

    def set_load(self, f_free, f_presc=None):
        f = np.ascontiguousarray(f_free, dtype=np.float64) if f_free is not None else None
        fp = np.ascontiguousarray(f_presc, dtype=np.float64) if f_presc is not None and len(f_presc) else None
        _check(hostlib().hfem_problem_set_load(self.h, _p(f), _p(fp)), "set_load")

    def linear_static(self):
        it, res = C.c_int(0), C.c_double(0)
        _check(hostlib().hfem_problem_linear_static(self.h, C.byref(it), C.byref(res)), "linear_static")
        return it.value, res.value

    def nonlinear_step(self):
        out = (C.c_int * 3)()
        err = C.c_double(0)
        _check(hostlib().hfem_problem_nonlinear_step(self.h, out, C.byref(err)), "nonlinear_step")
        return dict(newton=out[0], pcg=out[1], status=out[2], force_error=err.value)

    def set_arc_length(self, steplength, psi=1.0, maxiter=30, maxrestarts=4, minsteplength=0.0,
                       initialsteplength=0.0, reqiterations=3, miniter=0, manrmsteps=0, rtolf=1e-3, rtold=1e-3,
                       rootselectiontype=0):
        """NonLinearStatic indirect control (controlmode 0): CylindricalALM
        options under the deck's keyword names (calmls.h:51-75)."""
        _check(hostlib().hfem_problem_set_arc_length(self.h, float(psi), int(maxiter), int(maxrestarts),
                                                     float(minsteplength), float(steplength),
                                                     float(initialsteplength), int(reqiterations), int(miniter),
                                                     int(manrmsteps), float(rtolf), float(rtold),
                                                     int(rootselectiontype)), "set_arc_length")

    def set_arc_length_hpc(self, hpcmode, eqs, weights=None):
        """CylindricalALM hyperplane control: hpcmode 1 (constraint over the
        controlled free equations eqs, 1-based, 0 = prescribed) or 2
        (linearized, sum of weights times the increments = steplength)."""
        e = np.ascontiguousarray(eqs, dtype=np.int32)
        w = np.ascontiguousarray(weights if weights is not None else np.zeros(len(e)), dtype=np.float64)
        _check(hostlib().hfem_problem_set_arc_length_hpc(self.h, int(hpcmode), len(e), _p(e) if len(e) else None,
                                                         _p(w) if len(w) else None), "set_arc_length_hpc")

    def set_reference_load(self, f_free, f_presc=None):
        f = np.ascontiguousarray(f_free, dtype=np.float64)
        fp = np.ascontiguousarray(f_presc if f_presc is not None else np.zeros(self.deck.npeq), dtype=np.float64)
        _check(hostlib().hfem_problem_set_reference_load(self.h, _p(f), _p(fp) if fp.size else None),
               "set_reference_load")

    def arc_length_step(self):
        """One CylindricalALM step: iterations, PCG iterations, restarts,
        status, the load level reached and the next arc length."""
        out = (C.c_int * 4)()
        outd = (C.c_double * 4)()
        _check(hostlib().hfem_problem_arc_length_step(self.h, out, outd), "arc_length_step")
        return dict(iterations=out[0], pcg=out[1], restarts=out[2], status=out[3], load_level=outd[0],
                    step_length=outd[1], force_error=outd[2], disp_error=outd[3])

    def save_converged(self):
        _check(hostlib().hfem_problem_save_converged(self.h), "save_converged")

    def restart_step(self):
        """Back to the saved solution with the temp Gauss-point values reset."""
        _check(hostlib().hfem_problem_restart_step(self.h), "restart_step")

    def update(self):
        _check(hostlib().hfem_problem_update(self.h), "update")

    def solution(self):
        u = np.zeros(self.deck.neq)
        _check(hostlib().hfem_problem_get_solution(self.h, _p(u)), "get_solution")
        return u

    def reactions(self):
        r = np.zeros(max(self.deck.npeq, 1))
        _check(hostlib().hfem_problem_get_reactions(self.h, _p(r)), "get_reactions")
        return r[: self.deck.npeq]

    def state(self):
        n = self.deck.nelem * {1: 8, 2: 1, 3: 27}[self.deck.etype]
        s = np.zeros(n * 38)
        _check(hostlib().hfem_problem_get_state(self.h, _p(s)), "get_state")
        return s.reshape(n, 38)

    def set_profiling(self, on=True):
        _check(hostlib().hfem_problem_set_profiling(self.h, int(bool(on))), "set_profiling")

    def profile(self):
        n64, n32 = C.c_int(0), C.c_int(0)
        m64, m32 = C.c_double(0), C.c_double(0)
        _check(hostlib().hfem_problem_profile(self.h, C.byref(n64), C.byref(m64), C.byref(n32), C.byref(m32)),
               "profile")
        return dict(n64=n64.value, ms64=m64.value, n32=n32.value, ms32=m32.value)

    def context(self):
        """The problem's GPU context (borrowed), e.g. to create a surrogate on it."""
        from .hipfem import Context
        h = C.c_void_p()
        _check(hostlib().hfem_problem_context(self.h, C.byref(h)), "context")
        return Context(handle=h.value)

    def set_surrogate(self, net, band=0.05, tol=1e-4):
        _check(hostlib().hfem_problem_set_surrogate(self.h, net.h if net is not None else None, C.c_double(band),
                                                    C.c_double(tol)), "set_surrogate")

    def surrogate_stats(self, reset=False):
        c = (C.c_longlong * 3)()
        _check(hostlib().hfem_problem_surrogate_stats(self.h, c, int(bool(reset))), "surrogate_stats")
        return dict(queued=c[0], accepted=c[1], rejected=c[2])

    def solver_info(self):
        o, mf = C.c_int(0), C.c_int(0)
        _check(hostlib().hfem_problem_solver_info(self.h, C.byref(o), C.byref(mf)), "solver_info")
        return dict(outer=o.value, matrix_free=bool(mf.value))

    def matrix_at(self, i, j):
        v = C.c_double(0)
        _check(hostlib().hfem_problem_matrix_at(self.h, int(i), int(j), C.byref(v)), "matrix_at")
        return v.value

    def matrix_layout(self):
        f = C.c_int(0)
        g, b = C.c_longlong(0), C.c_longlong(0)
        _check(hostlib().hfem_problem_matrix_layout(self.h, C.byref(f), C.byref(g), C.byref(b)), "matrix_layout")
        return dict(form=f.value, groups=g.value, blocks=b.value)

    def matrix_nnz(self):
        v = C.c_int64(0)
        _check(hostlib().hfem_problem_matrix_nnz(self.h, C.byref(v)), "matrix_nnz")
        return v.value

    def matrix_class(self):
        buf = C.create_string_buffer(64)
        _check(hostlib().hfem_problem_matrix_class(self.h, buf, 64), "matrix_class")
        return buf.value.decode()
