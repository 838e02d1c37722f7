"""hippcg_solve_dev_noalloc (the allocation-free plain PCG) against
hippcg_solve_dev through the C ABI: two calls in a row, the second on a
smaller system, each bit for bit the allocating entry point's result (same
kernels, same order), and the size limit refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from oofem_amd import meshgen
from oracle import oracle

pytestmark = pytest.mark.gpu

HIPFEM_ETOOBIG = -6
MAX_N = 1048576


def _system(ctx, d):
    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx
    K = HipCSRSparseMtrx(ctx)
    ev = BatchedEvaluator(ctx, d)
    ev.build_structure(K)
    ev.assemble_stiffness(K)
    f, _ = oracle.surface_load(d)
    return K, ev, f


def _solve(ctx, fn, K, f):
    from oofem_amd.hipfem import DeviceVector, HipPCGSolver, lib
    S = HipPCGSolver(ctx, lstol=1e-10, lsiter=100000, lsprecond=1)
    b = DeviceVector(ctx, len(f), f)
    x = DeviceVector(ctx, len(f))
    reason, it, ach = C.c_int(0), C.c_int(0), C.c_double(0)
    rc = getattr(lib(), fn)(S.h, K.h, b.ptr, x.ptr, C.byref(reason), C.byref(it), C.byref(ach))
    out = (rc, reason.value, it.value, ach.value, x.download())
    S.close()
    return out


def test_noalloc_solve_equals_allocating_solve(ctx):
    decks = [meshgen.cantilever_hex(8, 8, 24, jitter=0.2, e_sigma=0.5, seed=11, name="na_big"),
             meshgen.cantilever_hex(3, 4, 9, jitter=0.2, e_sigma=0.5, seed=12, name="na_small")]
    assert decks[1].neq < decks[0].neq
    for d in decks:
        print(f"no-alloc PCG on {d.name}: neq={d.neq}", flush=True)
        K, ev, f = _system(ctx, d)
        rc0, r0, it0, a0, x0 = _solve(ctx, "hippcg_solve_dev", K, f)
        rc1, r1, it1, a1, x1 = _solve(ctx, "hippcg_solve_dev_noalloc", K, f)
        assert rc0 == 0 and rc1 == 0
        assert r0 == r1 == 1 and it0 == it1 and a0 == a1 and a1 <= 1e-10
        np.testing.assert_array_equal(x1, x0)
        ref = oracle.linear_static(d, tol=1e-10)
        assert np.linalg.norm(x1 - ref["u"]) <= 1e-10 * np.linalg.norm(ref["u"])


def test_noalloc_refuses_systems_beyond_its_limit(ctx):
    """A diagonal pattern of MAX_N + 1 equations: the call returns
    HIPFEM_ETOOBIG before touching the GPU."""
    from oofem_amd.hipfem import HipCSRSparseMtrx, HipPCGSolver, lib
    n = MAX_N + 1
    K = HipCSRSparseMtrx(ctx)
    K.buildInternalStructureCSR(np.arange(n + 1, dtype=np.int64), np.arange(n, dtype=np.int32))
    S = HipPCGSolver(ctx, lstol=1e-10, lsiter=10, lsprecond=1)
    reason, it, ach = C.c_int(0), C.c_int(0), C.c_double(0)
    rc = lib().hippcg_solve_dev_noalloc(S.h, K.h, C.c_void_p(8), C.c_void_p(8), C.byref(reason), C.byref(it),
                                        C.byref(ach))
    assert rc == HIPFEM_ETOOBIG
    assert b"exceed" in lib().hipfem_last_error()
    S.close()
