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


def csr_structure(deck):
    """Free-free pattern (rowptr int64, colind int32) of the deck's elements."""
    m = Mesh(deck)
    nnz = lib().orc_csr_count(m.ref)
    rowptr = np.zeros(deck.neq + 1, dtype=np.int64)
    colind = np.zeros(nnz, dtype=np.int32)
    lib().orc_csr_fill(m.ref, rowptr.ctypes.data_as(C.c_void_p), colind.ctypes.data_as(C.c_void_p))
    return rowptr, colind


def assemble(deck, rowptr, colind, state=None, tangent=False, omp=False, val=None):
    """Stiffness values on a pattern of csr_structure: the reference's
    element order, or (omp) its OpenMP element loop with critical-section
    scatter (timing baseline)."""
    m = Mesh(deck)
    if val is None:
        val = np.zeros(colind.shape[0])
    st = state.ctypes.data_as(C.c_void_p) if state is not None else None
    fn = lib().orc_assemble_stiffness_omp if omp else lib().orc_assemble_stiffness
    rc = fn(m.ref, rowptr.ctypes.data_as(C.c_void_p), colind.ctypes.data_as(C.c_void_p),
            val.ctypes.data_as(C.c_void_p), st, int(tangent))
    if rc:
        raise RuntimeError(f"orc_assemble_stiffness rc={rc}")
    return val


def omp_threads():
    return int(lib().orc_omp_threads())


def csr(deck, state=None, tangent=False):
    """Assembled free-free stiffness (rowptr int64, colind int32, val) in the
    reference's element order."""
    rowptr, colind = csr_structure(deck)
    return rowptr, colind, assemble(deck, rowptr, colind, state=state, tangent=tangent)


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


def pcg_hist(rowptr, colind, val, b, x0=None, tol=1e-10, maxit=100000, precond=1):
    """orc_pcg with its residual history ||r_k||/||b||, k = 0..iters."""
    n = rowptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else _d(x0).copy()
    it = C.c_int(0)
    res = C.c_double(0)
    hist = np.zeros(min(maxit, 1 << 22) + 1)
    rc = lib().orc_pcg_hist(n, rowptr.ctypes.data_as(C.c_void_p), colind.ctypes.data_as(C.c_void_p),
                            val.ctypes.data_as(C.c_void_p), _d(b).ctypes.data_as(C.c_void_p),
                            x.ctypes.data_as(C.c_void_p), C.c_double(tol), maxit, precond, C.byref(it),
                            C.byref(res), hist.ctypes.data_as(C.c_void_p), hist.shape[0])
    return x, it.value, res.value, rc, hist[:it.value + 1].copy()


_REF_CG_PATH = os.path.join(_HERE, "_ref", "libref_iml_cg.so")


def have_ref_cg():
    """True when the reference's iml/cg.h has been compiled into oracle/_ref."""
    return os.path.exists(_REF_CG_PATH)


def ref_cg(rowptr, colind, val, b, x0=None, tol=1e-10, maxit=100000, storage=0, precond=1):
    """The reference's own CG template (iml/cg.h, compiled unchanged into
    oracle/_ref/libref_iml_cg.so by oracle/Makefile) on a symmetric CSR
    matrix.  storage 0: SymCompCol product order (symcompcol.C:201-218),
    1: CompCol order (compcol.C:128-142).  Returns (x, flag, iters, tol,
    hist) with hist[k] = ||r_k||/||b|| as cg.h computes it."""
    L = C.CDLL(_REF_CG_PATH)
    n = rowptr.shape[0] - 1
    x = np.zeros(n) if x0 is None else _d(x0).copy()
    it = C.c_int(0)
    t = C.c_double(0)
    cap = min(maxit, 1 << 22) + 2
    norms = np.zeros(cap)
    nn = C.c_int(0)
    flag = L.ref_iml_cg(n, np.ascontiguousarray(rowptr, np.int64).ctypes.data_as(C.c_void_p),
                        np.ascontiguousarray(colind, np.int32).ctypes.data_as(C.c_void_p),
                        _d(val).ctypes.data_as(C.c_void_p), _d(b).ctypes.data_as(C.c_void_p),
                        x.ctypes.data_as(C.c_void_p), maxit, C.c_double(tol), storage, precond, C.byref(it),
                        C.byref(t), norms.ctypes.data_as(C.c_void_p), cap, C.byref(nn))
    if flag < 0:
        raise RuntimeError(f"ref_iml_cg failed with {flag}")
    rec = norms[:nn.value]
    normb = rec[0] if rec[0] != 0.0 else 1.0
    return x, flag, it.value, t.value, rec[1:] / normb


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


def internal_forces(deck, u_free, u_presc=None, state=None, update_state=False, omp=False):
    """f_int (free, prescribed) and the per-dof-id EBE norms; omp: the
    reference's OpenMP element loop (timing baseline)."""
    m = Mesh(deck)
    f = np.zeros(deck.neq)
    fp = np.zeros(deck.npeq)
    ebe = np.zeros(3)
    up = _d(u_presc) if u_presc is not None else np.zeros(deck.npeq)
    st = state.ctypes.data_as(C.c_void_p) if state is not None else None
    fn = lib().orc_internal_forces_omp if omp else lib().orc_internal_forces
    rc = fn(m.ref, _d(u_free).ctypes.data_as(C.c_void_p), up.ctypes.data_as(C.c_void_p), st, int(update_state),
            f.ctypes.data_as(C.c_void_p), fp.ctypes.data_as(C.c_void_p), ebe.ctypes.data_as(C.c_void_p))
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


class NROpts(C.Structure):
    _fields_ = [("initial_guess", C.c_int), ("rtolf", C.c_double), ("rtold", C.c_double), ("maxiter", C.c_int),
                ("miniter", C.c_int), ("manrmsteps", C.c_int), ("lstol", C.c_double), ("linesearch", C.c_int),
                ("ls_tol", C.c_double), ("ls_amp", C.c_double), ("ls_maxeta", C.c_double),
                ("ls_keep_point", C.c_int), ("maxinc", C.c_double), ("stiffmode", C.c_int), ("nddm", C.c_int),
                ("ddm_eq", C.c_void_p), ("ddm_val", C.c_void_p), ("ddm_factor", C.c_double),
                ("cnr_miniter", C.c_int)]


def nonlinear_step_ex(deck, state, X, u_presc, u_presc_prev=None, initial_guess=True, RT=None, rtolf=1e-8,
                      maxiter=50, manrmsteps=1, lstol=1e-12, linesearch=False, ls_tol=0.8, ls_amp=2.5, ls_maxeta=4.0,
                      maxinc=1e20, stiffmode=0, ddm=None, ddm_factor=0.0, constrainednrminiter=0, rtold=0.0,
                      miniter=0, ls_keep_point=False, history=False):
    """nonlinear_step with the NRSolver extras: line search (LineSearchNM,
    linesearch.C:54-245) in the reference's order (search, cap, X += ddX,
    nrsolver.C:316-350) or, with ls_keep_point, capped before the search and
    ending at the search's point; the increment cap maxinc; the displacement
    criterion rtold and miniter; NonLinearStatic's stiffness mode (0
    tangent, 1/3 secant = elastic, 2 elastic); ddm: (free equations
    1-based, values) of NRSolver's direct displacement control with the
    step's increment factor ddm_factor.
    Returns dict(newton, pcg, status, force_error, line_searches[, errors]):
    errors (with history) is an (iterations + 1) x 6 array of the u, v, w
    force errors and displacement errors of every convergence check (-1:
    group unused or criterion off)."""
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
    o = NROpts(int(bool(initial_guess)), float(rtolf), float(rtold), int(maxiter), int(miniter), int(manrmsteps),
               float(lstol), int(bool(linesearch)), float(ls_tol), float(ls_amp), float(ls_maxeta),
               int(bool(ls_keep_point)), float(maxinc), int(stiffmode), nd,
               de.ctypes.data if nd else None, dv.ctypes.data if nd else None, float(ddm_factor),
               int(constrainednrminiter))
    cap = min(int(maxiter), 10000) + 1
    hist = np.zeros((cap, 6)) if history else None
    rc = lib().orc_nonlinear_step_opt(m.ref, state.ctypes.data_as(C.c_void_p), X.ctypes.data_as(C.c_void_p),
                                      _d(u_presc).ctypes.data_as(C.c_void_p),
                                      prev.ctypes.data_as(C.c_void_p) if prev is not None else None,
                                      rt.ctypes.data_as(C.c_void_p) if rt is not None else None, C.byref(o), out,
                                      C.byref(err), hist.ctypes.data_as(C.c_void_p) if history else None,
                                      cap if history else 0)
    if rc:
        raise RuntimeError(f"orc_nonlinear_step_opt rc={rc}")
    r = dict(newton=out[0], pcg=out[1], status=out[2], force_error=err.value, line_searches=out[3])
    if history:
        r["errors"] = hist[:out[0] + 1].copy()
    return r


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


class ArcLength:
    """CylindricalALM (calmls.C) restated over the oracle's assembly, internal
    forces and CG: hyperplane control off, no CALM line search, norms over
    all equations, no initial load vector.  Options as the deck's keywords
    (initializeFrom, calmls.C:681-897, defaults of the constructor 60-101).
    The arc length, the reached load level, the Newton mode and the previous
    increment persist between steps as the reference's members do."""

    def __init__(self, steplength, psi=1.0, maxiter=30, maxrestarts=4, minsteplength=0.0, initialsteplength=0.0,
                 reqiterations=3, miniter=0, manrmsteps=0, rtolf=1e-3, rtold=1e-3, rootselectiontype=0,
                 lstol=1e-12, stiffmode=0, hpcmode=0, hpc_eqs=(), hpc_weights=()):
        if not steplength > 0.0:
            raise ValueError("steplength must be positive")
        self.psi = 1.0 if psi < 0.0 else psi
        self.nsmax = max(int(maxiter), 30)
        self.maxrestarts = int(maxrestarts)
        self.minsl, self.maxsl = float(minsteplength), float(steplength)
        self.deltaL = initialsteplength if initialsteplength > 0.0 else steplength
        self.reqit = min(max(int(reqiterations), 3), 1000)
        self.miniter = int(miniter)
        self.manrm = int(manrmsteps)
        self.rtolf, self.rtold = float(rtolf), float(rtold)
        self.rst = int(rootselectiontype)
        self.lstol = lstol
        self.stiffmode = stiffmode
        self.mode = self.old_mode = 1 if self.manrm > 0 else 0   # 0 modified, 1 accelerated, 2 full NRM
        self.tick = -1
        self.reached = 0.0
        self.old_dX = None
        # hyperplane control (calmls.C:754-797, 900-957): controlled free
        # equations (1-based; 0 = prescribed, skipped), weights of mode 2
        self.hpc = int(hpcmode)
        keep = [i for i, q in enumerate(hpc_eqs) if q > 0]
        self.hpc_idx = np.array([hpc_eqs[i] - 1 for i in keep], dtype=np.int64)
        self.hpc_w = np.array([hpc_weights[i] for i in keep]) if self.hpc == 2 else np.ones(len(keep))

    def step(self, deck, state, X, up, R):
        """One step of CylindricalALM::solve (calmls.C:108-481); X (free
        total displacements) and the temp Gauss-point state are updated in
        place.  Returns dict(iterations, restarts, load_level, step_length,
        force_error, disp_error, converged, pcg)."""
        if self.tick == 0:
            self.mode = self.old_mode
        if self.tick > 0:
            self.tick -= 1
        n = deck.neq
        XInit = X.copy()
        deltaXt = np.zeros(n)
        deltaXb = np.zeros(n)
        RR = float(R @ R)
        psi2 = self.psi * self.psi
        pcg_total = [0]

        def matrix():
            if self.stiffmode == 2:
                return csr(deck)
            return csr(deck, state, tangent=True)

        def solve(K, b, x0):
            x, it, res, rc = pcg(*K, b, x0=x0, tol=self.lstol, maxit=100000)
            if rc:
                raise RuntimeError(f"orc_pcg rc={rc}")
            pcg_total[0] += it
            return x

        irest = 0
        converged = True
        while True:
            dX = np.zeros(n)
            K = matrix()
            deltaXt = solve(K, R, deltaXt)
            m = self.hpc_idx
            if self.hpc == 1:
                p = np.sqrt(deltaXt[m] @ deltaXt[m] + psi2 * (R[m] @ R[m]))
            elif self.hpc == 2:
                p = deltaXt[m] @ self.hpc_w
            else:
                p = np.sqrt(deltaXt @ deltaXt + psi2 * RR)
            XR = deltaXt @ R if (self.rst == 0 or self.old_dX is None) else deltaXt @ self.old_dX
            DL = dl = (-1.0 if XR < 0.0 else 1.0) * self.deltaL / p
            Lam = self.reached + DL
            dX = solve(K, DL * R, dX)
            X[:] = XInit + dX
            nite = 0
            F, _, ebe = internal_forces(deck, X, up, state=state, update_state=True)
            restart = False
            while True:
                nite += 1
                dXm1 = dX.copy()
                DLm1 = DL
                if self.mode == 2 or (self.mode == 1 and nite % self.manrm == 0):
                    K = matrix()
                    deltaXt = solve(K, R, deltaXt)
                deltaXb = solve(K, Lam * R - F, deltaXb)
                # computeDeltaLambda calmls.C:995-1146 (eta = 1)
                XX, XXt, X_Xt = deltaXt @ deltaXt, dX @ deltaXt, deltaXb @ deltaXt
                dXdX, dXX_, X_X_ = dX @ dX, dX @ deltaXb, deltaXb @ deltaXb
                a1 = XX + psi2 * RR
                a2 = RR * psi2 * DL * 2.0 + 2.0 * XXt + 2.0 * X_Xt
                a3 = dXdX + 2.0 * dXX_ + X_X_ + DL * DL * RR * psi2 - self.deltaL * self.deltaL
                a4, a5 = dXX_ + dXdX, XXt
                if self.hpc == 1:
                    pr = dX[m] + deltaXb[m]
                    Rm = R[m] @ R[m]
                    a1 = deltaXt[m] @ deltaXt[m] + psi2 * Rm
                    a2 = Rm * psi2 * DL * 2.0 + 2.0 * (deltaXt[m] @ pr)
                    a3 = pr @ pr - self.deltaL * self.deltaL + DL * DL * Rm * psi2
                    a4 = dX[m] @ deltaXb[m] + dX[m] @ dX[m]
                    a5 = dX[m] @ deltaXt[m]
                discr = a2 * a2 - 4.0 * a1 * a3
                if self.hpc == 2:
                    denom = deltaXt[m] @ self.hpc_w
                    if abs(denom) < 1e-20:
                        raise RuntimeError("zero denominator in linearized control")
                    dl = (self.deltaL - (dX[m] + deltaXb[m]) @ self.hpc_w) / denom
                elif discr < 0.0:
                    dl = 0.0
                else:
                    sq = np.sqrt(discr)
                    lam1, lam2 = (-a2 + sq) / 2.0 / a1, (-a2 - sq) / 2.0 / a1
                    if self.rst == 0:
                        cos1 = (a4 + a5 * lam1) / self.deltaL / self.deltaL
                        cos2 = (a4 + a5 * lam2) / self.deltaL / self.deltaL
                        dl = lam1 if cos1 > cos2 else lam2
                    else:
                        s = DL + dl
                        c1 = (dXdX + dXX_ + lam1 * XXt) + psi2 * (s * s + s * lam1)
                        c2 = (dXdX + dXX_ + lam2 * XXt) + psi2 * (s * s + s * lam2)
                        dl = lam1 if c1 > c2 else lam2
                ddX = dl * deltaXt + deltaXb
                dX = dXm1 + ddX
                X[:] = XInit + dX
                dr = ddX @ ddX
                DL = DLm1 + dl
                Lam = self.reached + DL
                F, _, ebe = internal_forces(deck, X, up, state=state, update_state=True)
                # checkConvergence calmls.C:484-677 (all equations)
                r = Lam * R - F
                fe = r @ r
                dXX = X @ X
                en = ebe.sum()
                if RR * Lam * Lam > 1e-6:
                    fe = np.sqrt(fe / (RR * Lam * Lam))
                elif en > 1e-6:
                    fe = np.sqrt(fe / en)
                else:
                    fe = np.sqrt(fe)
                de = dr if dXX < 1e-6 else np.sqrt(dr / dXX)
                out = abs(fe) > self.rtolf * 1e10 or abs(de) > self.rtold * 1e10 or np.isnan(fe) or np.isnan(de)
                conv = not (abs(fe) > self.rtolf or abs(de) > self.rtold)
                if nite >= self.nsmax or out:
                    irest += 1
                    if irest <= self.maxrestarts:
                        self.deltaL = max(self.deltaL * 0.25, self.minsl)
                        X[:] = XInit
                        state_restart(deck, state)
                        self.old_mode, self.mode, self.tick = self.mode, 2, 2
                        restart = True
                        break
                    converged = False
                    break
                if conv and nite >= self.miniter:
                    break
            if not restart:
                break
        if nite > self.reqit:
            self.deltaL = self.deltaL * self.reqit / nite
        else:
            self.deltaL = self.deltaL * np.sqrt(np.sqrt(self.reqit / nite))
        self.deltaL = min(self.deltaL, self.maxsl)
        self.deltaL = max(self.deltaL, self.minsl)
        self.reached = Lam
        self.old_dX = dX.copy()
        return dict(iterations=nite, restarts=irest, load_level=Lam, step_length=self.deltaL, force_error=fe,
                    disp_error=de, converged=converged, pcg=pcg_total[0])
