"""NRSolver semantics of the GPU driver against closed forms, and the
reference's EngngModel hooks called through the base class (needs the GPU).

The closed forms are those of tests/nrpins.py: the per-dof-group force
criterion at the first check of a traction step and of a
prescribed-displacement step, the line search and increment cap in the
reference's order and with the search's point kept, and the displacement
criterion rtold."""
import numpy as np
import pytest

from oofem_amd import meshgen
from oracle import oracle
from tests import nrpins

pytestmark = pytest.mark.gpu


def test_force_criterion_closed_form_gpu():
    """First check of a traction step from rest: u, v, w force errors 0, 1,
    0; first check of the uniaxial layer's prescribed-displacement step: the
    EBE-normalized per-group errors of the closed form."""
    from oofem_amd.problem import StructuralProblem
    d = nrpins.traction_block()
    f, fp = d.external_load()
    P = StructuralProblem(d, lstol=1e-13, rtolf=1e-8, maxiter=5, manrmsteps=1, initial_guess=0)
    P.set_load(f, fp)
    P.set_prescribed(d.prescribed_values())
    rep = P.nonlinear_step()
    e = P.nr_errors()
    assert rep["status"] == 1 and rep["newton"] == 1 and len(e) == 2
    np.testing.assert_array_equal(e[0], [0.0, 1.0, 0.0, -1.0, -1.0, -1.0])
    P.close()
    for nx, ny in ((2, 3), (3, 2), (1, 1), (4, 1)):
        d = nrpins.uniaxial_layer(nx, ny)
        P = StructuralProblem(d, lstol=1e-13, rtolf=1e-8, maxiter=5, manrmsteps=1, initial_guess=0)
        P.set_load(np.zeros(d.neq), np.zeros(d.npeq))
        P.set_prescribed(d.prescribed_values())
        rep = P.nonlinear_step()
        assert rep["status"] == 1
        np.testing.assert_allclose(P.nr_errors()[0, :3], nrpins.uniaxial_layer_errors(nx, ny), rtol=1e-13)
        P.close()


@pytest.mark.parametrize("mode,maxiter,factor,status,newton", [
    ("reference", 2, 1.25, -1, 2), ("reference", 5, 0.75, -1, 5),
    ("keep_point", 10, 1.0, 1, 4), ("no_search", 10, 1.0, 1, 4)])
def test_line_search_and_maxinc_closed_form_gpu(mode, maxiter, factor, status, newton):
    """Linear block, maxinc = max|X*|/4: the reference's order (search, cap,
    X += ddX) ends at 1.25 X* after two corrections and 0.75 X* after five;
    with the search's point kept, or without a search, the fourth check
    converges at X*."""
    from oofem_amd.problem import StructuralProblem
    d = nrpins.linear_block()
    f, fp = d.external_load()
    rp, ci, val = oracle.csr(d)
    Xs, _, _, rc = oracle.pcg(rp, ci, val, f, tol=1e-14)
    m0 = np.abs(Xs).max()
    P = StructuralProblem(d, lstol=1e-14, rtolf=1e-8, maxiter=maxiter, manrmsteps=1, initial_guess=0)
    P.set_nr_extras(linesearch=mode != "no_search", maxinc=0.25 * m0)
    P.set_nr_criteria(ls_keep_point=mode == "keep_point")
    P.set_load(f, fp)
    P.set_prescribed(d.prescribed_values())
    rep = P.nonlinear_step()
    assert rep["status"] == status and rep["newton"] == newton, rep
    assert np.abs(P.solution() - factor * Xs).max() <= 1e-9 * m0
    P.close()


def test_displacement_criterion_closed_form_gpu():
    """rtold: displacement errors 0, 1, ~0 over the checks of a linear step,
    which then takes two Newton iterations instead of one."""
    from oofem_amd.problem import StructuralProblem
    d = nrpins.linear_block()
    f, fp = d.external_load()
    P = StructuralProblem(d, lstol=1e-14, rtolf=1e-8, maxiter=10, manrmsteps=1, initial_guess=0)
    P.set_nr_criteria(rtold=1e-6)
    P.set_load(f, fp)
    P.set_prescribed(d.prescribed_values())
    rep = P.nonlinear_step()
    e = P.nr_errors()
    assert rep["status"] == 1 and rep["newton"] == 2 and len(e) == 3
    np.testing.assert_array_equal(e[0, 3:], [0.0, 0.0, 0.0])
    np.testing.assert_allclose(e[1, 3:], [1.0, 1.0, 1.0], rtol=1e-12)
    assert (e[2, 3:] < 1e-10).all()
    P.close()


def test_engngmodel_hooks_through_the_base_class():
    """EngngModel::assemble(SparseMtrx&, TimeStep*, const MatrixAssembler&,
    const UnknownNumberingScheme&, Domain*) with TangentAssembler(mode) and
    StructuralEngngModel::updateInternalRHS(FloatArray&, TimeStep*, Domain*,
    FloatArray*), both called through the EngngModel base: the oracle's
    consistent tangent of a plastic state and its elastic matrix, and its
    internal forces and EBE norms."""
    from oofem_amd.problem import StructuralProblem
    d = meshgen.config3(3, 3, 6)
    up = d.prescribed_values(0.5)
    P = StructuralProblem(d, lstol=1e-12, rtolf=1e-8, maxiter=30, manrmsteps=1)
    P.set_load(np.zeros(d.neq), np.zeros(d.npeq))
    P.set_prescribed(up)
    assert P.nonlinear_step()["status"] == 1
    # out of balance: the converged displacements with the end moved further
    up2 = d.prescribed_values(0.7)
    P.set_prescribed(up2)
    st = oracle.state_new(d)
    X = P.solution()
    fo, _, ebeo = oracle.internal_forces(d, X, up2, state=st, update_state=True)
    f, e = P.update_internal_rhs()
    assert np.abs(f - fo).max() <= 1e-12 * np.abs(fo).max()
    np.testing.assert_allclose(e, ebeo, rtol=1e-12)
    assert (st.reshape(-1, 38)[:, 13] > 0).mean() > 0.3
    rp, ci, _ = oracle.csr(d)
    rows = (0, d.neq // 2, d.neq - 1)
    for mode, tangent in ((0, True), (2, False), (1, False)):
        P.assemble(mode)
        _, _, oval = oracle.csr(d, state=st, tangent=tangent)
        got = np.array([P.matrix_at(i + 1, ci[k] + 1) for i in rows for k in range(rp[i], rp[i + 1])])
        ref = np.array([oval[k] for i in rows for k in range(rp[i], rp[i + 1])])
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(oval).max(), mode
    P.close()
