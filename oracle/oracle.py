"""ctypes front end of the CPU oracle (oracle/_build/libfem_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  The C sources
(fem_oracle.c) cite the reference routines they restate.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfem_oracle.so")
_lib = None

GPSTATE = 38
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class OrcMesh(C.Structure):
    _fields_ = [("nnode", C.c_int), ("coords", C.c_void_p), ("nelem", C.c_int), ("etype", C.c_int),
                ("conn", C.c_void_p), ("emat", C.c_void_p), ("nmat", C.c_int), ("mat", C.c_void_p),
                ("eqn", C.c_void_p), ("neq", C.c_int), ("npeq", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_dndx.restype = C.c_double
        L.orc_detj.restype = C.c_double
        L.orc_csr_count.restype = C.c_longlong
        _lib = L
    return _lib


class Mesh:
    """Keeps the numpy arrays alive behind an orc_mesh struct."""

    def __init__(self, deck):
        self.deck = deck
        self._keep = [deck.coords, deck.conn, deck.emat, deck.mat, deck.eqn]
        self.s = OrcMesh(deck.nnode, deck.coords.ctypes.data, deck.nelem, deck.etype,
                         deck.conn.ctypes.data, deck.emat.ctypes.data, deck.mat.shape[0],
                         deck.mat.ctypes.data, deck.eqn.ctypes.data, deck.neq, deck.npeq)

    @property
    def ref(self):
        return C.byref(self.s)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def gauss(etype):
    xi = np.zeros(81)
    w = np.zeros(27)
    n = lib().orc_gauss(etype, xi.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p))
    return xi[: 3 * n].reshape(n, 3), w[:n]


def dndx(etype, X, xi):
    nen = {1: 8, 2: 4, 3: 20}[etype]
    out = np.zeros(3 * nen)
    err = C.c_int(0)
    det = lib().orc_dndx(etype, _d(X).ctypes.data_as(C.c_void_p), _d(xi).ctypes.data_as(C.c_void_p),
                         out.ctypes.data_as(C.c_void_p), C.byref(err))
    return det, out.reshape(nen, 3), err.value


def isole_D(E, nu):
    D = np.zeros(36)
    lib().orc_isole_D(C.c_double(E), C.c_double(nu), D.ctypes.data_as(C.c_void_p))
    return D.reshape(6, 6)


def element_stiffness(etype, X, D, symmetric=True):
    nen = {1: 8, 2: 4, 3: 20}[etype]
    D = _d(D)
    stride = 0 if D.size == 36 else 36
    Ke = np.zeros((3 * nen) ** 2)
    rc = lib().orc_element_stiffness(etype, _d(X).ctypes.data_as(C.c_void_p), D.ctypes.data_as(C.c_void_p),
                                     stride, int(symmetric), Ke.ctypes.data_as(C.c_void_p))
    if rc:
        raise RuntimeError(f"orc_element_stiffness rc={rc}")
    return Ke.reshape(3 * nen, 3 * nen)


def mises_return(mp, eps, epsp_n, kappa_n):
    sig = np.zeros(6)
    epsp = np.zeros(6)
    sdev = np.zeros(6)
    kap = C.c_double(0)
    lib().orc_mises_return(_d(mp).ctypes.data_as(C.c_void_p), _d(eps).ctypes.data_as(C.c_void_p),
                           _d(epsp_n).ctypes.data_as(C.c_void_p), C.c_double(kappa_n),
                           sig.ctypes.data_as(C.c_void_p), epsp.ctypes.data_as(C.c_void_p), C.byref(kap),
                           sdev.ctypes.data_as(C.c_void_p))
    return sig, epsp, kap.value, sdev


def mises_tangent(mp, kappa_n, kappa_t, sdev):
    D = np.zeros(36)
    lib().orc_mises_tangent(_d(mp).ctypes.data_as(C.c_void_p), C.c_double(kappa_n), C.c_double(kappa_t),
                            _d(sdev).ctypes.data_as(C.c_void_p), D.ctypes.data_as(C.c_void_p))
    return D.reshape(6, 6)


def csr(deck, state=None, tangent=False):
    """Assembled free-free stiffness (rowptr int64, colind int32, val) in the
    reference's element order."""
    m = Mesh(deck)
    nnz = lib().orc_csr_count(m.ref)
    rowptr = np.zeros(deck.neq + 1, dtype=np.int64)
    colind = np.zeros(nnz, dtype=np.int32)
    val = np.zeros(nnz)
    lib().orc_csr_fill(m.ref, rowptr.ctypes.data_as(C.c_void_p), colind.ctypes.data_as(C.c_void_p))
    st = state.ctypes.data_as(C.c_void_p) if state is not None else None
    rc = lib().orc_assemble_stiffness(m.ref, rowptr.ctypes.data_as(C.c_void_p), colind.ctypes.data_as(C.c_void_p),
                                      val.ctypes.data_as(C.c_void_p), st, int(tangent))
    if rc:
        raise RuntimeError(f"orc_assemble_stiffness rc={rc}")
    return rowptr, colind, val


def spmv(rowptr, colind, val, x):
    y = np.zeros(rowptr.shape[0] - 1)
    lib().orc_spmv(len(y), rowptr.ctypes.data_as(C.c_void_p), colind.ctypes.data_as(C.c_void_p),
                   val.ctypes.data_as(C.c_void_p), _d(x).ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p))
    return y


def pcg(rowptr, colind, val, b, x0=None, tol=1e-10, maxit=100000, precond=1):
    n = rowptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else _d(x0).copy()
    it = C.c_int(0)
    res = C.c_double(0)
    rc = lib().orc_pcg(n, rowptr.ctypes.data_as(C.c_void_p), colind.ctypes.data_as(C.c_void_p),
                       val.ctypes.data_as(C.c_void_p), _d(b).ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                       C.c_double(tol), maxit, precond, C.byref(it), C.byref(res))
    return x, it.value, res.value, rc


def skyline_solve(rowptr, colind, val, b):
    """Profile U^T D U direct solve (Skyline, skyline.C:248-290,429-497).
    Returns (x, nwk)."""
    n = rowptr.shape[0] - 1
    x = np.zeros(n)
    nwk = C.c_longlong(0)
    rc = lib().orc_skyline_solve(n, rowptr.ctypes.data_as(C.c_void_p), colind.ctypes.data_as(C.c_void_p),
                                 val.ctypes.data_as(C.c_void_p), _d(b).ctypes.data_as(C.c_void_p),
                                 x.ctypes.data_as(C.c_void_p), C.byref(nwk))
    if rc:
        raise RuntimeError(f"orc_skyline_solve rc={rc}")
    return x, nwk.value


def state_new(deck):
    return np.zeros(GPSTATE * deck.nelem * {1: 8, 2: 1, 3: 27}[deck.etype])


def state_commit(deck, state):
    m = Mesh(deck)
    lib().orc_state_commit(m.ref, state.ctypes.data_as(C.c_void_p))


def state_restart(deck, state):
    """Element::initForNewStep on every element (temp := committed)."""
    m = Mesh(deck)
    lib().orc_state_restart(m.ref, state.ctypes.data_as(C.c_void_p))


def internal_forces(deck, u_free, u_presc=None, state=None, update_state=False):
    m = Mesh(deck)
    f = np.zeros(deck.neq)
    fp = np.zeros(deck.npeq)
    ebe = np.zeros(3)
    up = _d(u_presc) if u_presc is not None else np.zeros(deck.npeq)
    st = state.ctypes.data_as(C.c_void_p) if state is not None else None
    rc = lib().orc_internal_forces(m.ref, _d(u_free).ctypes.data_as(C.c_void_p), up.ctypes.data_as(C.c_void_p),
                                   st, int(update_state), f.ctypes.data_as(C.c_void_p),
                                   fp.ctypes.data_as(C.c_void_p), ebe.ctypes.data_as(C.c_void_p))
    if rc:
        raise RuntimeError(f"orc_internal_forces rc={rc}")
    return f, fp, ebe


def surface_load(deck, factor=1.0):
    m = Mesh(deck)
    f = np.zeros(deck.neq)
    fp = np.zeros(deck.npeq)
    el = np.ascontiguousarray(deck.faces[:, 0], dtype=np.int32)
    sf = np.ascontiguousarray(deck.faces[:, 1], dtype=np.int32)
    tr = _d(deck.traction * factor)
    rc = lib().orc_surface_load(m.ref, len(el), el.ctypes.data_as(C.c_void_p), sf.ctypes.data_as(C.c_void_p),
                                tr.ctypes.data_as(C.c_void_p), f.ctypes.data_as(C.c_void_p), fp.ctypes.data_as(C.c_void_p))
    if rc:
        raise RuntimeError(f"orc_surface_load rc={rc}")
    return f, fp


def body_load(deck, elems, dens, b, factor=1.0):
    """orc_body_load: DeadWeight of the elements ``elems`` (0-based) with
    densities ``dens`` and force per unit mass ``b`` times ``factor``."""
    m = Mesh(deck)
    f = np.zeros(deck.neq)
    fp = np.zeros(deck.npeq)
    el = np.ascontiguousarray(elems, dtype=np.int32)
    de = _d(np.asarray(dens, dtype=np.float64))
    bb = _d(np.asarray(b, dtype=np.float64) * factor)
    rc = lib().orc_body_load(m.ref, len(el), el.ctypes.data_as(C.c_void_p), de.ctypes.data_as(C.c_void_p),
                             bb.ctypes.data_as(C.c_void_p), f.ctypes.data_as(C.c_void_p),
                             fp.ctypes.data_as(C.c_void_p))
    if rc:
        raise RuntimeError(f"orc_body_load rc={rc}")
    return f, fp


def nonlinear_step(deck, state, X, u_presc, RT=None, rtolf=1e-8, maxiter=50, manrmsteps=1, lstol=1e-12,
                   u_presc_prev=None, initial_guess=True):
    """One StaticStructural step with NRSolver; X and state updated in place.
    initial_guess: the tangent predictor over the prescribed increments
    u_presc - u_presc_prev (IG_Tangent, the reference's default).
    Returns (newton_iterations, cg_iterations, status, force_error)."""
    m = Mesh(deck)
    out = (C.c_int * 3)()
    err = C.c_double(0)
    rt = _d(RT) if RT is not None else None
    prev = _d(u_presc_prev) if u_presc_prev is not None else None
    rc = lib().orc_nonlinear_step(m.ref, state.ctypes.data_as(C.c_void_p), X.ctypes.data_as(C.c_void_p),
                                  _d(u_presc).ctypes.data_as(C.c_void_p),
                                  prev.ctypes.data_as(C.c_void_p) if prev is not None else None,
                                  int(bool(initial_guess)),
                                  rt.ctypes.data_as(C.c_void_p) if rt is not None else None,
                                  C.c_double(rtolf), maxiter, manrmsteps, C.c_double(lstol), out, C.byref(err))
    if rc:
        raise RuntimeError(f"orc_nonlinear_step rc={rc}")
    return out[0], out[1], out[2], err.value


def nonlinear_step_ex(deck, state, X, u_presc, u_presc_prev=None, initial_guess=True, RT=None, rtolf=1e-8,
                      maxiter=50, manrmsteps=1, lstol=1e-12, linesearch=False, ls_tol=0.8, ls_amp=2.5, ls_maxeta=4.0,
                      maxinc=1e20, stiffmode=0, ddm=None, ddm_factor=0.0):
    """nonlinear_step with the NRSolver extras: line search (LineSearchNM,
    linesearch.C:54-245), the increment cap maxinc (nrsolver.C:340-353) and
    NonLinearStatic's stiffness mode (0 tangent, 2 elastic); ddm: (free
    equations 1-based, values) of NRSolver's direct displacement control with
    the step's increment factor ddm_factor.
    Returns dict(newton, pcg, status, force_error, line_searches)."""
    m = Mesh(deck)
    out = (C.c_int * 4)()
    err = C.c_double(0)
    rt = _d(RT) if RT is not None else None
    prev = _d(u_presc_prev) if u_presc_prev is not None else None
    nd = 0
    de = dv = None
    if ddm is not None:
        de = np.ascontiguousarray(ddm[0], dtype=np.int32)
        dv = _d(ddm[1])
        nd = len(de)
    rc = lib().orc_nonlinear_step_ex(m.ref, state.ctypes.data_as(C.c_void_p), X.ctypes.data_as(C.c_void_p),
                                     _d(u_presc).ctypes.data_as(C.c_void_p),
                                     prev.ctypes.data_as(C.c_void_p) if prev is not None else None,
                                     int(bool(initial_guess)),
                                     rt.ctypes.data_as(C.c_void_p) if rt is not None else None, C.c_double(rtolf),
                                     maxiter, manrmsteps, C.c_double(lstol), int(bool(linesearch)), C.c_double(ls_tol),
                                     C.c_double(ls_amp), C.c_double(ls_maxeta), C.c_double(maxinc), int(stiffmode),
                                     nd, de.ctypes.data_as(C.c_void_p) if nd else None,
                                     dv.ctypes.data_as(C.c_void_p) if nd else None, C.c_double(ddm_factor), out,
                                     C.byref(err))
    if rc:
        raise RuntimeError(f"orc_nonlinear_step_ex rc={rc}")
    return dict(newton=out[0], pcg=out[1], status=out[2], force_error=err.value, line_searches=out[3])


def static_structural(deck, up_ref, final_time, dt, ltf=None, reduction=None, **nr):
    """StaticStructural time stepping with NRSolver steps and the time-step
    reduction of EngngModel::solveYourself (engngm.C:573-662): a step that
    fails (maximum iterations or divergence, a ConvergenceException) is
    restarted from the last converged state with dt * ncrf
    (MetaStep::reduceTimeStep, metastep.C:138-146; error below dtmin or
    after nmaxrestarts restarts), and after every converged step dt is
    adapted to its Newton iterations (MetaStep::adaptTimeStep,
    metastep.C:149-183 with SimpleReductionStrategy's factors,
    timestepreductionstrategy.C:70-90: x2 below minrequirediter, x0.5 above
    maxrequirediter, clamped to [dtmin, dtmax] and to the final time).
    reduction: None (NoReductionStrategy: a failed step is kept as it is and
    dt never changes) or dict(dtmin, dtmax, ncrf=0.5, nmaxrestarts=5,
    minrequirediter=3, maxrequirediter=8).  Prescribed values at time t are
    up_ref * ltf(t) (ltf default t / final_time).  Returns (X, state, steps)
    with steps a list of dict(time, dt, newton, pcg, status, restarts)."""
    ltf = ltf or (lambda t: t / final_time)
    st = state_new(deck)
    X = np.zeros(deck.neq)
    up_prev = np.zeros(deck.npeq)
    t = 0.0
    steps = []
    red = None if reduction is None else dict(dict(ncrf=0.5, nmaxrestarts=5, minrequirediter=3, maxrequirediter=8),
                                             **reduction)
    while t < final_time - 1e-12 * final_time:
        restarts = 0
        while True:
            target = t + dt
            up = up_ref * ltf(target)
            Xs = X.copy()
            rep = nonlinear_step_ex(deck, st, X, up, u_presc_prev=up_prev, **nr)
            if rep["status"] == 1 or red is None:
                break
            dt2 = dt * red["ncrf"]
            if dt2 < red["dtmin"]:
                raise RuntimeError("Minimum time increment reached")
            restarts += 1
            if restarts > red["nmaxrestarts"]:
                raise RuntimeError("Maximum number of time step reductions has been reached.")
            dt = dt2
            X[:] = Xs
            state_restart(deck, st)           # EngngModel::initStepIncrements
        steps.append(dict(time=target, dt=dt, newton=rep["newton"], pcg=rep["pcg"], status=rep["status"],
                          restarts=restarts))
        if red is not None:
            n = rep["newton"]
            f = 1.0 if (n == 0 or red["minrequirediter"] <= n <= red["maxrequirediter"]) else (
                2.0 if n < red["minrequirediter"] else 0.5)
            dt = dt * (min(1.0, f) if f < 1 else max(1.0, f))
            dt = min(max(dt, red["dtmin"]), red["dtmax"])
        if target + dt > final_time:
            dt = final_time - target
        state_commit(deck, st)
        up_prev = up
        t = target
    return X, st, steps


def extrapolated_forces(deck, state, du_presc):
    """sum_e K_e(tangent) du_e over prescribed increments, free rows."""
    m = Mesh(deck)
    f = np.zeros(deck.neq)
    rc = lib().orc_extrapolated_forces(m.ref, state.ctypes.data_as(C.c_void_p),
                                       _d(du_presc).ctypes.data_as(C.c_void_p), f.ctypes.data_as(C.c_void_p))
    if rc:
        raise RuntimeError(f"orc_extrapolated_forces rc={rc}")
    return f


def linear_static(deck, tol=1e-10, maxit=100000, u_presc=None):
    m = Mesh(deck)
    f, fp = surface_load(deck)
    up = deck.prescribed_values() if u_presc is None else _d(u_presc)
    u = np.zeros(deck.neq)
    R = np.zeros(deck.npeq)
    it = C.c_int(0)
    res = C.c_double(0)
    rc = lib().orc_linear_static(m.ref, f.ctypes.data_as(C.c_void_p), fp.ctypes.data_as(C.c_void_p),
                                 up.ctypes.data_as(C.c_void_p), C.c_double(tol), maxit,
                                 u.ctypes.data_as(C.c_void_p), R.ctypes.data_as(C.c_void_p), C.byref(it), C.byref(res))
    return dict(u=u, reactions=R, iters=it.value, resid=res.value, rc=rc)
