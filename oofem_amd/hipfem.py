"""Python front end of libhipfem (the gfx950 C ABI declared in include/hipfem.h).

Names mirror the reference's plug-in interfaces so that scripts read like
the reference: ``HipCSRSparseMtrx`` is the SparseMtrx (sparsemtrx.h),
``HipPCGSolver`` the SparseLinearSystemNM (sparselinsystemnm.h) and
``BatchedEvaluator`` the element/material evaluator behind
EngngModel::assemble (engngm.C:889-1037).  Every call goes to the HIP
library; there is no CPU fallback: if the library is missing or no GPU is
present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

CR_CONVERGED = 1
CR_DIVERGED_ITS = -1

_lib = None


class HipFemError(RuntimeError):
    pass


def lib():
    """Load libhipfem.so (never builds it: run __graft_entry__.build() first)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_build.LIB):
            raise HipFemError(f"{_build.LIB} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(_build.LIB)
        vp, ip, dp = C.c_void_p, C.c_int, C.c_double
        L.hipfem_last_error.restype = C.c_char_p
        L.hipfem_version.restype = C.c_char_p
        L.hipfem_ctx_stream.restype = vp
        L.hipcsr_nnz.restype = C.c_int64
        L.hipcsr_version.restype = C.c_long
        L.hipcsr_get.argtypes = [vp, ip, ip, C.POINTER(dp)]
        L.hipcsr_set.argtypes = [vp, ip, ip, dp]
        L.hipcsr_scale.argtypes = [vp, dp]
        L.hipcsr_add.argtypes = [vp, dp, vp]
        L.hipcsr_add_diagonal.argtypes = [vp, dp, vp]
        L.hippcg_set_options.argtypes = [vp, dp, ip, ip]
        L.hippcg_set_precision.argtypes = [vp, ip, dp, ip]
        L.hippcg_set_reliable.argtypes = [vp, dp]
        for f in ("hipfem_ctx_destroy", "hipcsr_destroy", "hippcg_destroy", "hfem_model_destroy"):
            getattr(L, f).restype = None
            getattr(L, f).argtypes = [vp]
        _lib = L
    return _lib


def _check(rc, what=""):
    if rc != 0:
        msg = lib().hipfem_last_error().decode(errors="replace")
        raise HipFemError(f"{what} failed (status {rc}): {msg}")


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Context:
    """One GPU + one HIP stream (hipfem_ctx)."""

    def __init__(self, device: int = 0, handle=None):
        self.h = C.c_void_p()
        self.owned = handle is None
        if handle is None:
            _check(lib().hipfem_ctx_create(device, C.byref(self.h)), "hipfem_ctx_create")
        else:
            self.h = C.c_void_p(handle)   # borrowed (e.g. a problem's context): never destroyed here
        self.device = device

    def synchronize(self):
        _check(lib().hipfem_ctx_synchronize(self.h), "synchronize")

    @property
    def stream(self):
        return lib().hipfem_ctx_stream(self.h)

    def close(self):
        if self.h and self.owned:
            lib().hipfem_ctx_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceVector:
    """FP64 vector resident on the context's GPU (hipfem_dev_*)."""

    def __init__(self, ctx: Context, n: int, host=None):
        self.ctx = ctx
        self.n = int(n)
        self.ptr = C.c_void_p()
        _check(lib().hipfem_dev_alloc(ctx.h, C.c_int64(self.n), C.byref(self.ptr)), "dev_alloc")
        if host is not None:
            self.upload(host)
        else:
            self.fill(0.0)

    @property
    def addr(self):
        return self.ptr.value

    def upload(self, host):
        h = _f64(host)
        if h.size != self.n:
            raise ValueError("size mismatch")
        _check(lib().hipfem_dev_upload(self.ctx.h, self.ptr, _p(h), C.c_int64(self.n)), "dev_upload")

    def download(self):
        h = np.zeros(self.n)
        _check(lib().hipfem_dev_download(self.ctx.h, _p(h), self.ptr, C.c_int64(self.n)), "dev_download")
        return h

    def axpby(self, a, x, b):
        """self = a*x + b*self"""
        _check(lib().hipfem_dev_axpby(self.ctx.h, C.c_int64(self.n), C.c_double(a), x.ptr if x is not None else None,
                                      C.c_double(b), self.ptr), "dev_axpby")

    def fill(self, v=0.0):
        _check(lib().hipfem_dev_axpby(self.ctx.h, C.c_int64(self.n), C.c_double(0.0), None, C.c_double(0.0),
                                      self.ptr), "dev_fill")
        if v != 0.0:
            raise NotImplementedError("fill with non-zero")

    def dot(self, other):
        out = C.c_double(0)
        _check(lib().hipfem_dev_dot(self.ctx.h, C.c_int64(self.n), self.ptr, other.ptr, C.byref(out)), "dev_dot")
        return out.value

    def close(self):
        if self.ptr:
            lib().hipfem_dev_free(self.ctx.h, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipCSRSparseMtrx:
    """SparseMtrx 'hipcsr' (reference sparsemtrx.h:60-281)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = C.c_void_p()
        _check(lib().hipcsr_create(ctx.h, C.byref(self.h)), "hipcsr_create")

    def close(self):
        if self.h:
            lib().hipcsr_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- structure
    def buildInternalStructure(self, neq, loc_arrays):
        """Pattern from element location arrays (nelem x ndofe, 1-based, 0 = none)."""
        loc = _i32(loc_arrays)
        if loc.ndim != 2:
            raise ValueError("loc_arrays must be 2-D (elements x dofs)")
        _check(lib().hipcsr_build_structure(self.h, int(neq), loc.shape[0], loc.shape[1], _p(loc)),
               "buildInternalStructure")

    def buildInternalStructureCSR(self, rowptr, colind):
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        ci = _i32(colind)
        _check(lib().hipcsr_build_structure_csr(self.h, len(rp) - 1, _p(rp), _p(ci)), "buildInternalStructure")

    # --- assembly
    def assemble(self, loc, mat, cloc=None):
        loc = _i32(loc)
        mat = _f64(mat)
        if cloc is None:
            _check(lib().hipcsr_assemble(self.h, len(loc), _p(loc), _p(mat)), "assemble")
        else:
            cloc = _i32(cloc)
            _check(lib().hipcsr_assemble_rc(self.h, len(loc), _p(loc), len(cloc), _p(cloc), _p(mat)), "assemble")
        return 1

    def assembleBegin(self):
        _check(lib().hipcsr_assemble_begin(self.h), "assembleBegin")
        return 1

    def assembleEnd(self):
        _check(lib().hipcsr_assemble_end(self.h), "assembleEnd")
        return 1

    def zero(self):
        _check(lib().hipcsr_zero(self.h), "zero")

    def times_scalar(self, a):
        _check(lib().hipcsr_scale(self.h, float(a)), "times")

    def add(self, a, other):
        _check(lib().hipcsr_add(self.h, float(a), other.h), "add")

    def addDiagonal(self, a, d):
        d = _f64(d)
        _check(lib().hipcsr_add_diagonal(self.h, float(a), _p(d)), "addDiagonal")

    # --- products
    def times(self, x):
        x = _f64(x)
        y = np.zeros(self.giveNumberOfRows())
        _check(lib().hipcsr_times(self.h, _p(x), _p(y)), "times")
        return y

    def timesT(self, x):
        x = _f64(x)
        y = np.zeros(self.giveNumberOfRows())
        _check(lib().hipcsr_times_t(self.h, _p(x), _p(y)), "timesT")
        return y

    def times_matrix(self, B):
        B = _f64(B)
        Y = np.zeros_like(B)
        _check(lib().hipcsr_times_mat(self.h, B.shape[1], _p(B), _p(Y)), "times(FloatMatrix)")
        return Y

    def times_dev(self, dx_ptr, dy_ptr):
        _check(lib().hipcsr_times_dev(self.h, C.c_void_p(dx_ptr), C.c_void_p(dy_ptr)), "times_dev")

    # --- element access
    def at(self, i, j):
        v = C.c_double(0.0)
        _check(lib().hipcsr_get(self.h, int(i), int(j), C.byref(v)), "at")
        return v.value

    def set_at(self, i, j, v):
        _check(lib().hipcsr_set(self.h, int(i), int(j), float(v)), "at")

    def isAllocatedAt(self, i, j):
        return bool(lib().hipcsr_is_allocated(self.h, int(i), int(j)))

    def giveNumberOfRows(self):
        return lib().hipcsr_nrows(self.h)

    def giveNumberOfColumns(self):
        return self.giveNumberOfRows()

    def giveVersion(self):
        return lib().hipcsr_version(self.h)

    def nnz(self):
        return lib().hipcsr_nnz(self.h)

    def download(self):
        n, nnz = self.giveNumberOfRows(), self.nnz()
        rp = np.zeros(n + 1, dtype=np.int64)
        ci = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz)
        _check(lib().hipcsr_download(self.h, _p(rp), _p(ci), _p(v)), "download")
        return rp, ci, v

    def upload_values(self, val):
        val = _f64(val)
        _check(lib().hipcsr_upload_values(self.h, _p(val)), "upload_values")

    def clone(self):
        out = HipCSRSparseMtrx.__new__(HipCSRSparseMtrx)
        out.ctx = self.ctx
        out.h = C.c_void_p()
        _check(lib().hipcsr_clone(self.h, C.byref(out.h)), "clone")
        return out

    def set_spmv_shape(self, lanes, unroll, nontemporal=True):
        _check(lib().hipcsr_set_spmv_shape(self.h, int(lanes), int(unroll), int(bool(nontemporal))), "set_spmv_shape")

    def set_spmv_stream(self, rows_per_block, unroll=4):
        _check(lib().hipcsr_set_spmv_stream(self.h, int(rows_per_block), int(unroll)), "set_spmv_stream")

    def set_matrix_free(self, on=True):
        _check(lib().hipcsr_set_matrix_free(self.h, int(bool(on))), "set_matrix_free")

    def is_matrix_free(self):
        return bool(lib().hipcsr_is_matrix_free(self.h))

    def times_matrix_free(self, x):
        x = _f64(x)
        y = np.zeros(self.giveNumberOfRows())
        _check(lib().hipcsr_times_matrix_free(self.h, _p(x), _p(y)), "times_matrix_free")
        return y

    def set_spmv_nodeblock(self, on=True):
        _check(lib().hipcsr_set_spmv_nodeblock(self.h, int(bool(on))), "set_spmv_nodeblock")

    def is_nodeblocked(self):
        """1: 3x3 node-block form, 2: row-group form, 0: neither"""
        return int(lib().hipcsr_is_nodeblocked(self.h))

    def spmv_shape(self):
        a, b, c = C.c_int(0), C.c_int(0), C.c_int(0)
        _check(lib().hipcsr_get_spmv_shape(self.h, C.byref(a), C.byref(b), C.byref(c)), "get_spmv_shape")
        return a.value, b.value, c.value

    def giveClassName(self):
        return "HipCSRSparseMtrx"


class HipPCGSolver:
    """SparseLinearSystemNM 'hippcg' with IMLSolver options (imlsolver.C:67-101)."""

    def __init__(self, ctx: Context, lstol=1e-5, lsiter=200, lsprecond=1, precision=0, inner_tol=1e-4, max_outer=100,
                 reldelta=0.1):
        self.ctx = ctx
        self.h = C.c_void_p()
        _check(lib().hippcg_create(ctx.h, C.byref(self.h)), "hippcg_create")
        self.initializeFrom(lstol=lstol, lsiter=lsiter, lsprecond=lsprecond, precision=precision,
                            inner_tol=inner_tol, max_outer=max_outer, reldelta=reldelta)

    def initializeFrom(self, lstol=1e-5, lsiter=200, lsprecond=1, precision=0, inner_tol=1e-4, max_outer=100,
                       reldelta=0.1):
        _check(lib().hippcg_set_options(self.h, float(lstol), int(lsiter), int(lsprecond)), "initializeFrom")
        _check(lib().hippcg_set_precision(self.h, int(precision), float(inner_tol), int(max_outer)), "initializeFrom")
        _check(lib().hippcg_set_reliable(self.h, float(reldelta)), "initializeFrom")

    def close(self):
        if self.h:
            lib().hippcg_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, A: HipCSRSparseMtrx, b, x):
        """x (numpy, modified in place) holds the initial guess.  Returns
        (ConvergedReason, iterations, achieved relative residual)."""
        b = _f64(b)
        if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
            raise TypeError("x must be a contiguous float64 numpy array (updated in place)")
        reason, it, ach = C.c_int(0), C.c_int(0), C.c_double(0)
        _check(lib().hippcg_solve(self.h, A.h, _p(b), _p(x), C.byref(reason), C.byref(it), C.byref(ach)), "solve")
        return reason.value, it.value, ach.value

    def solve_multi(self, A: HipCSRSparseMtrx, B, X):
        """Several right-hand sides (columns of B, n x nrhs); X (n x nrhs,
        float64, Fortran order, modified in place) holds the initial guesses.
        Returns (ConvergedReason, largest iterations, largest residual)."""
        B = np.asfortranarray(B, dtype=np.float64)
        if not (isinstance(X, np.ndarray) and X.dtype == np.float64 and X.flags.f_contiguous and X.shape == B.shape):
            raise TypeError("X must be a Fortran-ordered float64 array shaped like B (updated in place)")
        reason, it, ach = C.c_int(0), C.c_int(0), C.c_double(0)
        _check(lib().hippcg_solve_multi(self.h, A.h, int(B.shape[1]), C.c_void_p(B.ctypes.data),
                                        C.c_void_p(X.ctypes.data), C.byref(reason), C.byref(it), C.byref(ach)),
               "solve_multi")
        return reason.value, it.value, ach.value

    def solve_dev(self, A: HipCSRSparseMtrx, db_ptr, dx_ptr):
        reason, it, ach = C.c_int(0), C.c_int(0), C.c_double(0)
        _check(lib().hippcg_solve_dev(self.h, A.h, C.c_void_p(db_ptr), C.c_void_p(dx_ptr), C.byref(reason),
                                      C.byref(it), C.byref(ach)), "solve_dev")
        return reason.value, it.value, ach.value

    def last_outer(self):
        return lib().hippcg_last_outer(self.h)

    def last_matrix_free(self):
        return bool(lib().hippcg_last_matrix_free(self.h))

    def set_fused(self, fused=True):
        _check(lib().hippcg_set_fused(self.h, int(bool(fused))), "set_fused")

    def set_profiling(self, on=True):
        _check(lib().hippcg_set_profiling(self.h, int(bool(on))), "set_profiling")

    def profile(self):
        """Timed SpMV launches of the last solve: counts and summed ms (FP64, FP32 inner)."""
        n64, n32 = C.c_int(0), C.c_int(0)
        ms64, ms32 = C.c_double(0), C.c_double(0)
        _check(lib().hippcg_profile(self.h, C.byref(n64), C.byref(ms64), C.byref(n32), C.byref(ms32)), "profile")
        return dict(n64=n64.value, ms64=ms64.value, n32=n32.value, ms32=ms32.value)

    def set_history(self, cap):
        """Record ||r_k||/||b|| of every plain solve, up to cap values."""
        _check(lib().hippcg_set_history(self.h, int(cap)), "set_history")

    def history(self, cap=1 << 20):
        """The last solve's residual history (k = 0..iterations)."""
        n = C.c_int(0)
        _check(lib().hippcg_get_history(self.h, None, 0, C.byref(n)), "get_history")
        out = np.zeros(max(1, min(cap, n.value)))
        _check(lib().hippcg_get_history(self.h, _p(out), int(out.shape[0]), C.byref(n)), "get_history")
        return out[:min(cap, n.value)]

    def giveClassName(self):
        return "HipPCGSolver"


class BatchedEvaluator:
    """Element/material evaluator for one mesh (hfem_model)."""

    def __init__(self, ctx: Context, deck=None):
        self.ctx = ctx
        self.h = C.c_void_p()
        _check(lib().hfem_model_create(ctx.h, C.byref(self.h)), "hfem_model_create")
        self.deck = None
        if deck is not None:
            self.set_mesh(deck)

    def close(self):
        if self.h:
            lib().hfem_model_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mesh(self, d):
        self.deck = d
        self._keep = [_f64(d.coords), _i32(d.conn), _i32(d.emat), _f64(d.mat), _i32(d.eqn)]
        c, cn, em, mt, eq = self._keep
        _check(lib().hfem_model_set_mesh(self.h, d.nnode, _p(c), d.etype, d.nelem, _p(cn), _p(em), mt.shape[0],
                                         _p(mt), _p(eq), d.neq, d.npeq), "set_mesh")

    def build_structure(self, K: HipCSRSparseMtrx):
        _check(lib().hfem_model_build_structure(self.h, K.h), "build_structure")

    def assemble_stiffness(self, K: HipCSRSparseMtrx, tangent=False):
        _check(lib().hfem_model_assemble_stiffness(self.h, K.h, int(bool(tangent))), "assemble_stiffness")

    def internal_forces(self, u_free, u_presc=None, update_state=False):
        d = self.deck
        u_free = _f64(u_free)
        up = _f64(u_presc) if u_presc is not None else None
        f = np.zeros(d.neq)
        fp = np.zeros(d.npeq)
        ebe = np.zeros(3)
        _check(lib().hfem_model_internal_forces(self.h, _p(u_free), _p(up), _p(f), _p(fp), _p(ebe),
                                                int(bool(update_state))), "internal_forces")
        return f, fp, ebe

    def internal_forces_dev(self, du_free, du_presc, df_free, df_presc, update_state=False, want_ebe=False):
        ebe = np.zeros(3) if want_ebe else None
        _check(lib().hfem_model_internal_forces_dev(self.h, C.c_void_p(du_free), C.c_void_p(du_presc or 0),
                                                    C.c_void_p(df_free), C.c_void_p(df_presc or 0), _p(ebe),
                                                    int(bool(update_state))), "internal_forces_dev")
        return ebe

    def tangent_product(self, x_free=None, x_presc=None, tangent=True):
        """y = sum_e K_e x_e (free and prescribed rows) with the elastic or
        the tangent D of the temp Gauss-point state, on host arrays."""
        d = self.deck
        xf = DeviceVector(self.ctx, max(d.neq, 1), x_free) if x_free is not None else None
        xp = DeviceVector(self.ctx, max(d.npeq, 1), x_presc) if x_presc is not None and d.npeq else None
        yf = DeviceVector(self.ctx, max(d.neq, 1))
        yp = DeviceVector(self.ctx, max(d.npeq, 1)) if d.npeq else None
        _check(lib().hfem_model_tangent_product_dev(self.h, int(bool(tangent)), C.c_void_p(xf.addr if xf else 0),
                                                    C.c_void_p(xp.addr if xp else 0), C.c_void_p(yf.addr),
                                                    C.c_void_p(yp.addr if yp else 0)), "tangent_product")
        return yf.download()[: d.neq], (yp.download()[: d.npeq] if yp else np.zeros(0))

    def set_element_operator(self, form=2):
        """FP32 element operator form: 0 one launch per color, 1 gather,
        2 per-color slots (default)."""
        _check(lib().hfem_model_set_element_operator(self.h, int(form)), "set_element_operator")

    def set_surrogate(self, net, band=0.05, tol=1e-4):
        _check(lib().hfem_model_set_surrogate(self.h, net.h if net is not None else None, C.c_double(band),
                                              C.c_double(tol)), "set_surrogate")

    def surrogate_stats(self, reset=False):
        c = np.zeros(3, dtype=np.int64)
        _check(lib().hfem_model_surrogate_stats(self.h, _p(c), int(bool(reset))), "surrogate_stats")
        return dict(queued=int(c[0]), accepted=int(c[1]), rejected=int(c[2]))

    def restart(self):
        """Temp Gauss-point values := committed (EngngModel::initStepIncrements)."""
        _check(lib().hfem_model_restart(self.h), "restart")

    def commit(self):
        _check(lib().hfem_model_commit(self.h), "commit")

    def reset_state(self):
        _check(lib().hfem_model_reset_state(self.h), "reset_state")

    def set_state(self, state):
        """Gauss-point state (ngp x 38, as get_state returns it)."""
        st = np.ascontiguousarray(state, dtype=np.float64)
        if st.size != lib().hfem_model_ngp_total(self.h) * 38:
            raise ValueError("state size mismatch")
        _check(lib().hfem_model_set_state(self.h, _p(st)), "set_state")

    def get_state(self):
        n = lib().hfem_model_ngp_total(self.h)
        s = np.zeros(n * 38)
        _check(lib().hfem_model_get_state(self.h, _p(s)), "get_state")
        return s.reshape(n, 38)

    def ncolors(self):
        return lib().hfem_model_ncolors(self.h)


class MisesSurrogate:
    """Learned MisesMat surrogate (hfem_mlp): 13 -> 3 x 128 SiLU -> 27, FP32
    master weights, HIP training kernels, BF16 MFMA inference."""

    def __init__(self, ctx: Context, E, nu, sig0, H, seed=1):
        self.ctx = ctx
        self.h = C.c_void_p()
        mat = np.array([E, nu, sig0, H], dtype=np.float64)
        self.E, self.nu, self.sig0, self.H = float(E), float(nu), float(sig0), float(H)
        lib().hfem_mlp_create.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        _check(lib().hfem_mlp_create(ctx.h, _p(mat), C.c_uint64(seed), C.byref(self.h)), "hfem_mlp_create")
        self.nparams = lib().hfem_mlp_nparams()
        s = np.zeros(3)
        _check(lib().hfem_mlp_scales(self.h, _p(s)), "scales")
        self.in_scale, self.sig_scale, self.tan_scale = s

    def close(self):
        if self.h:
            lib().hfem_mlp_destroy.argtypes = [C.c_void_p]
            lib().hfem_mlp_destroy.restype = None
            lib().hfem_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().hfem_mlp_get_params(self.h, _p(p)), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().hfem_mlp_set_params(self.h, _p(p)), "set_params")

    def generate(self, nsamples, seed=7, stdev=2e-4, maxsteps=50):
        _check(lib().hfem_mlp_generate(self.h, C.c_int64(nsamples), C.c_uint64(seed), C.c_double(stdev),
                                       C.c_int(maxsteps)), "generate")
        self.nsamples = int(nsamples)

    def samples(self, first, count):
        X = np.zeros((count, 13), dtype=np.float32)
        Y = np.zeros((count, 27), dtype=np.float32)
        _check(lib().hfem_mlp_get_samples(self.h, C.c_int64(first), C.c_int64(count), _p(X), _p(Y)), "get_samples")
        return X, Y

    def train(self, steps, batch=4096, lr=1e-3, lr_final=1e-5, seed=11):
        loss = C.c_double(0)
        _check(lib().hfem_mlp_train(self.h, C.c_int(steps), C.c_int(batch), C.c_double(lr), C.c_double(lr_final),
                                    C.c_uint64(seed), C.byref(loss)), "train")
        return loss.value

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 27), dtype=np.float32)
        _check(lib().hfem_mlp_forward(self.h, C.c_int(X.shape[0]), _p(X), _p(Y)), "forward")
        return Y

    def gradients(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_double(0)
        _check(lib().hfem_mlp_gradients(self.h, C.c_int(X.shape[0]), _p(X), _p(Y), _p(g), C.byref(loss)),
               "gradients")
        return loss.value, g

    def decode(self, X, Y):
        """Physical stress (n x 6) and tangent upper triangle (n x 21) from
        scaled inputs and outputs: the outputs are departures from the
        elastic trial, sigma = D_e (eps - epsp_n) - sig0 o_sigma and
        D = D_e - E o_D."""
        E, nu = self.E, self.nu
        G, K = E / (2 * (1 + nu)), E / (3 * (1 - 2 * nu))
        De = np.zeros((6, 6))
        De[:3, :3] = K - 2 * G / 3
        De[np.arange(3), np.arange(3)] += 2 * G
        De[np.arange(3, 6), np.arange(3, 6)] = G
        el = np.asarray(X, dtype=np.float64)[:, :6] / self.in_scale
        Y = np.asarray(Y, dtype=np.float64)
        sig = el @ De.T - Y[:, :6] / self.sig_scale
        iu = np.triu_indices(6)
        D = De[iu][None, :] - Y[:, 6:] / self.tan_scale
        return sig, D

    def infer_bf16(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 27), dtype=np.float32)
        _check(lib().hfem_mlp_infer_bf16(self.h, C.c_int(X.shape[0]), _p(X), _p(Y)), "infer_bf16")
        return Y
