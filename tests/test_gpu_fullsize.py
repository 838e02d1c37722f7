"""The hot path at the plasticity benchmark's full size (64x64x256 LSpace
block, 3,240,575 free equations, nnz 256,384,867), checked through
properties that do not depend on the size: symmetry of the assembled
consistent tangent, zero response to a rigid translation, the FP32 element
operator against the FP64 matrix, and PCG solutions whose true residual
meets the tolerance, bitwise repeatable (needs the GPU)."""
import numpy as np
import pytest

from oofem_amd import meshgen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def block(ctx):
    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx
    d = meshgen.config3()
    assert d.neq == 3240575 and d.nelem == 1048576
    ev = BatchedEvaluator(ctx, d)
    K = HipCSRSparseMtrx(ctx)
    ev.build_structure(K)
    # the uniaxial field at 6/20 of the end displacement after a committed
    # 5/20: every Gauss point plastic, as in the benchmark's load steps
    zmax = d.coords[:, 2].max()
    fe = d.eqn[:, 2] > 0
    for frac, commit in ((5.0 / 20.0, True), (6.0 / 20.0, False)):
        u = np.zeros(d.neq)
        u[d.eqn[fe, 2] - 1] = d.coords[fe, 2] / zmax * frac * d.meta["u_end"]
        ev.internal_forces(u, d.prescribed_values(frac), update_state=True)
        if commit:
            ev.commit()
    s = ev.get_state()
    assert (s[:, 13] > s[:, 6]).mean() > 0.99
    del s
    ev.assemble_stiffness(K, tangent=True)
    assert K.nnz() == 256384867
    yield d, ev, K
    K.close()
    ev.close()


def test_full_block_tangent_symmetry_and_rigid_translation(block):
    """x2 . (K x1) = x1 . (K x2) (the MisesMat consistent tangent is
    symmetric), and a rigid translation of every DOF, free and prescribed,
    produces no force at any free DOF."""
    d, ev, K = block
    rng = np.random.default_rng(3)
    x1, x2 = rng.standard_normal(d.neq), rng.standard_normal(d.neq)
    y1, y2 = K.times(x1), K.times(x2)
    a, b = x2 @ y1, x1 @ y2
    assert abs(a - b) <= 1e-12 * np.linalg.norm(x2) * np.linalg.norm(y1)
    for comp in range(3):
        xf = np.zeros(d.neq)
        xp = np.zeros(d.npeq)
        m = d.eqn[:, comp] > 0
        xf[d.eqn[m, comp] - 1] = 1.0
        m = d.eqn[:, comp] < 0
        xp[-d.eqn[m, comp] - 1] = 1.0
        yf, _ = ev.tangent_product(xf, xp, tangent=True)
        assert np.abs(yf).max() <= 1e-10 * np.abs(y1).max(), comp


def test_full_block_element_operator_matches_matrix(block):
    """The FP32 element operator against the FP64 matrix product, to FP32
    accuracy, on random vectors."""
    d, ev, K = block
    ev.set_element_operator(2)
    assert K.is_matrix_free()
    rng = np.random.default_rng(4)
    for _ in range(2):
        x = rng.standard_normal(d.neq)
        y64 = K.times(x)
        y32 = K.times_matrix_free(x)
        assert np.linalg.norm(y32 - y64) <= 2e-5 * np.linalg.norm(y64)


@pytest.mark.parametrize("precision", [0, 2])
def test_full_block_pcg_true_residual_and_repeatability(ctx, block, precision):
    """Plain FP64 PCG and the mixed form (FP32 element operator, FP64
    reliable updates) to 1e-8: the FP64 true residual |b - K x| / |b| meets
    the tolerance, and a second solve returns the same bits."""
    from oofem_amd.hipfem import DeviceVector, HipPCGSolver
    d, ev, K = block
    rng = np.random.default_rng(5)
    b = rng.standard_normal(d.neq)
    db = DeviceVector(ctx, d.neq, b)
    dx = DeviceVector(ctx, d.neq)
    s = HipPCGSolver(ctx, lstol=1e-8, lsiter=200000, lsprecond=1, precision=precision)
    out = []
    for _ in range(2):
        dx.fill(0.0)
        reason, it, res = s.solve_dev(K, db.addr, dx.addr)
        assert reason == 1 and res <= 1e-8, (reason, it, res)
        out.append(dx.download())
    s.close()
    x = out[0]
    assert np.array_equal(out[0], out[1])
    r = b - K.times(x)
    assert np.linalg.norm(r) <= 1.0001e-8 * np.linalg.norm(b)
    assert np.all(np.isfinite(x))
