"""HIP path against the oracle, through the C ABI (needs the GPU).

Bars: integer data (patterns) bit-exact; FP64 values within a few ulps of
the largest entry (tolerances written per test); solutions within 1e-10
relative of the oracle's, the plain FP64 path's bar against the reference
solver.
"""
import os

import numpy as np
import pytest

from oofem_amd import meshgen
from oofem_amd.deck import evaluate_checks, read_deck
from oracle import oracle
from tests.conftest import GOLDEN
from tests.deckrun import run_deck_gpu, run_deck_oracle
from tests.imlcases import CASES as IML_CASES
from tests.imlcases import fixture as _iml_fixture

IML = _iml_fixture()

pytestmark = pytest.mark.gpu


def _small_decks():
    hexd = meshgen.cantilever_hex(4, 3, 6, jitter=0.2, e_sigma=0.5, seed=11, name="hex_small")
    coords, tets = meshgen.kuhn_tets(4, 3, 5, 1.0, 0.8, 1.3, jitter=0.15, seed=3)
    bc = np.zeros((coords.shape[0], 3), bool)
    bc[: 5 * 4] = True
    mat = np.array([[meshgen.MAT_ISOLE, 210e3, 0.3, 0, 0, 0, 0, 0]])
    tetd = meshgen.Deck("tet_small", meshgen.LTRSPACE, coords, tets, np.zeros(len(tets), np.int32), mat, bc,
                        np.zeros_like(coords))
    qd = read_deck(os.path.join(GOLDEN, "cantilever_Qspace.in")).deck
    return {"lspace": hexd, "ltrspace": tetd, "qspace": qd}


DECKS = _small_decks()
SHAPE_DECKS = dict(DECKS, tetbar=meshgen.config2(nx=4, ny=3, nz=6, seed=5, jitter=0.12))


@pytest.mark.parametrize("name", list(DECKS))
def test_stiffness_matches_oracle(ctx, name):
    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx
    d = DECKS[name]
    rp, ci, val = oracle.csr(d)
    K = HipCSRSparseMtrx(ctx)
    ev = BatchedEvaluator(ctx, d)
    ev.build_structure(K)
    ev.assemble_stiffness(K)
    grp, gci, gval = K.download()
    np.testing.assert_array_equal(grp, rp)
    np.testing.assert_array_equal(gci, ci)
    scale = np.abs(val).max()
    assert np.abs(gval - val).max() <= 1e-13 * scale


@pytest.mark.parametrize("name", list(DECKS))
def test_spmv_matches_oracle(ctx, name):
    from oofem_amd.hipfem import HipCSRSparseMtrx
    d = DECKS[name]
    rp, ci, val = oracle.csr(d)
    K = HipCSRSparseMtrx(ctx)
    K.buildInternalStructureCSR(rp, ci)
    K.upload_values(val)
    x = np.random.default_rng(2).standard_normal(d.neq)
    y = K.times(x)
    yo = oracle.spmv(rp, ci, val, x)
    assert np.abs(y - yo).max() <= 1e-14 * np.abs(val).max() * np.abs(x).max() * 100
    yt = K.timesT(x)
    assert np.abs(yt - yo).max() <= 1e-14 * np.abs(val).max() * np.abs(x).max() * 100


SHAPES = [("lanes", 8, 6, 0), ("lanes", 16, 4, 1), ("lanes", 32, 2, 0), ("lanes", 16, 1, 0),
          ("stream", 8, 2, None), ("stream", 16, 4, None), ("stream", 32, 6, None), ("stream", 64, 4, None),
          ("nodeblock", 0, 0, None)]


@pytest.mark.parametrize("shape", SHAPES, ids=[f"{s[0]}-{s[1]}-{s[2]}" for s in SHAPES])
@pytest.mark.parametrize("name", ["lspace", "qspace", "tetbar"])
def test_spmv_shapes_match_oracle(ctx, name, shape):
    """Every SpMV kernel shape (plain and inside PCG) against the oracle."""
    from oofem_amd.hipfem import HipCSRSparseMtrx, HipFemError, HipPCGSolver
    d = SHAPE_DECKS[name]
    rp, ci, val = oracle.csr(d)
    K = HipCSRSparseMtrx(ctx)
    K.buildInternalStructureCSR(rp, ci)
    K.upload_values(val)
    kind, a, b, c = shape
    if kind == "lanes":
        K.set_spmv_shape(a, b, c)
    elif kind == "nodeblock":
        # fully free nodes: 3x3 block form; the QSpace beam and the tet bar
        # prescribe single dofs (w only) on their loaded end: row-group form
        assert K.is_nodeblocked() == (1 if name == "lspace" else 2)
        K.set_spmv_nodeblock(True)
    else:
        max_block = max(rp[min(r + a, d.neq)] - rp[r] for r in range(0, d.neq, a))
        if max_block > 4096:
            with pytest.raises(HipFemError):
                K.set_spmv_stream(a, b)
            return
        K.set_spmv_stream(a, b)
    x = np.random.default_rng(2).standard_normal(d.neq)
    y = K.times(x)
    yo = oracle.spmv(rp, ci, val, x)
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(val).max() * np.abs(x).max()
    b_ = np.random.default_rng(3).standard_normal(d.neq)
    for fused in (0, 1):
        s = HipPCGSolver(ctx, lstol=1e-12, lsiter=20000, lsprecond=1)
        s.set_fused(fused)
        u = np.zeros(d.neq)
        reason, it, res = s.solve(K, b_, u)
        uo, ito, reso, rco = oracle.pcg(rp, ci, val, b_, tol=1e-12)
        # summation order differs from the oracle's: near 1e-12 the count may move a few percent
        assert reason == 1 and abs(it - ito) <= max(2, 0.05 * ito)
        assert np.linalg.norm(u - uo) <= 1e-9 * np.linalg.norm(uo)


def test_config0_linear_static_parity(ctx):
    from oofem_amd.drivers import linear_static, reaction_sums
    d = meshgen.config0()
    ref = oracle.linear_static(d, tol=1e-10)
    got = linear_static(ctx, d, lstol=1e-10)
    assert abs(got["iters"] - ref["iters"]) <= 1, (got["iters"], ref["iters"])
    du = np.linalg.norm(got["u"] - ref["u"]) / np.linalg.norm(ref["u"])
    dr = np.linalg.norm(got["reactions"] - ref["reactions"]) / np.linalg.norm(ref["reactions"])
    assert du <= 1e-10, du
    assert dr <= 1e-10, dr
    assert abs(reaction_sums(d, got["reactions"])[1] - 1.0) < 1e-9


@pytest.mark.parametrize("case", list(IML_CASES))
def test_pcg_history_matches_reference_cg(ctx, case):
    """hippcg against the reference's own CG template (iml/cg.h compiled from
    the reference tree; its residual history is the committed fixture):
    iterations within 1, the first 50 residuals ||r_k||/||b|| to 1e-10
    relative, every residual within 10 % (round-off of the different
    dot-product order grows late in the solve)."""
    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx, HipPCGSolver
    d = IML_CASES[case]()
    ref = IML[case]
    K = HipCSRSparseMtrx(ctx)
    ev = BatchedEvaluator(ctx, d)
    ev.build_structure(K)
    ev.assemble_stiffness(K)
    f, _ = oracle.surface_load(d)
    S = HipPCGSolver(ctx, lstol=1e-10, lsiter=100000, lsprecond=1)
    S.set_history(100000)
    u = np.zeros(d.neq)
    reason, it, res = S.solve(K, f, u)
    h = S.history()
    assert reason == 1 and res <= 1e-10
    assert len(h) == it + 1 and h[-1] == res
    href = np.array(ref["hist"])
    assert abs(it - ref["iters"]) <= 1, (it, ref["iters"])
    np.testing.assert_allclose(h[:50], href[:50], rtol=1e-10)
    m = min(len(h), len(href))
    assert np.max(np.abs(h[:m] / href[:m] - 1.0)) < 0.1


def test_config0_tight_solve_matches_direct(ctx):
    """Both sides solved to 1e-13: displacements agree to 1e-12."""
    from oofem_amd.drivers import linear_static
    d = meshgen.config0()
    ref = oracle.linear_static(d, tol=1e-13)
    got = linear_static(ctx, d, lstol=1e-13)
    assert np.linalg.norm(got["u"] - ref["u"]) / np.linalg.norm(ref["u"]) < 1e-12


@pytest.mark.parametrize("name", ["patch301.in", "patch300.in", "cantilever_Qspace.in", "deadweight02.in"])
def test_reference_decks_on_gpu(ctx, name):
    prob = read_deck(os.path.join(GOLDEN, name))
    got = run_deck_gpu(ctx, prob)
    checks = evaluate_checks(prob, got)
    bad = [(c.line, v) for c, v, ok in checks if not ok]
    assert not bad, bad
    ref = run_deck_oracle(prob)
    for k in ref:
        scale = max(np.abs(ref[k]["u"]).max(), 1e-300)
        assert np.abs(got[k]["u"] - ref[k]["u"]).max() <= 1e-10 * scale


def test_tutorialmaterial_deck_on_gpu(ctx):
    """The reference's tutorialmaterial deck (one LSpace brick, every DOF
    prescribed, TutorialMaterial = the MisesMat J2 return) through the
    evaluator's MisesMat kernel: each step one return map per Gauss point
    and a commit.  The 48 #ELEMENT stress checks hold, and at every step the
    stresses, plastic strains and kappa agree with the oracle's to 1e-12 of
    their scale."""
    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx
    from tests.deckrun import _collect
    from tests.test_oracle import run_strain_driven_deck_oracle
    prob = read_deck(os.path.join(GOLDEN, "tutorialmaterial.in"))
    d = prob.deck
    assert d.neq == 0 and d.nelem == 1
    K = HipCSRSparseMtrx(ctx)
    ev = BatchedEvaluator(ctx, d)
    ev.build_structure(K)
    got = {}
    for k in range(1, prob.nsteps + 1):
        up, (_, fp) = prob.step_loading(k)
        _, fpi, _ = ev.internal_forces(np.zeros(0), up, update_state=True)
        ev.commit()
        got[k] = _collect(prob, np.zeros(0), up, fpi - fp, ev.get_state().reshape(-1))
    checks = evaluate_checks(prob, got)
    assert len(checks) == 48
    bad = [(c.line, v) for c, v, ok in checks if not ok]
    assert not bad, bad
    ref = run_strain_driven_deck_oracle(prob)
    for k in ref:
        for sl in (slice(20, 26), slice(0, 7)):       # stress; committed plastic strain and kappa
            a, b = got[k]["state"][:, :, sl], ref[k]["state"][:, :, sl]
            assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-300), (k, sl)


def test_c_abi_refuses_bad_inputs(ctx):
    """The boundary's error behaviour (OOFEM_ERROR in the reference): bad
    indices, unknown materials, unflushed assembly and entries outside the
    pattern are refused with an error before any kernel reads them, and the
    objects stay usable afterwards."""
    import copy

    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx, HipFemError, HipPCGSolver
    d = meshgen.cantilever_hex(2, 2, 3, name="tiny")
    ev = BatchedEvaluator(ctx)
    for field, change, msg in (("conn", lambda a: a.__setitem__((0, 0), d.nnode), "conn"),
                               ("emat", lambda a: a.__setitem__(0, len(d.mat)), "emat"),
                               ("mat", lambda a: a.__setitem__((0, 0), 7), "unknown model")):
        bad = copy.deepcopy(d)
        arr = getattr(bad, field).copy()
        change(arr)
        setattr(bad, field, arr)
        with pytest.raises(HipFemError, match=msg):
            ev.set_mesh(bad)
    ev.set_mesh(d)                                   # still usable
    with pytest.raises(HipFemError, match="build_structure"):
        ev.internal_forces(np.zeros(d.neq), np.zeros(d.npeq))   # no colouring yet
    K0 = HipCSRSparseMtrx(ctx)
    ev.build_structure(K0)
    f, _ = ev.internal_forces(np.zeros(d.neq), np.zeros(d.npeq))[:2]
    assert f.shape == (d.neq,) and not f.any()
    ev.close()
    K = HipCSRSparseMtrx(ctx)
    K.buildInternalStructure(4, np.array([[1, 2, 3, 0], [3, 4, 0, 0]]))
    with pytest.raises(HipFemError, match="not in the pattern"):
        K.assemble(np.array([1, 4]), np.eye(2))      # (1,4): no element couples them
    with pytest.raises(HipFemError, match="row 5"):
        K.assemble(np.array([5]), np.eye(1))
    K.zero()
    K.assembleBegin()
    K.assemble(np.array([1, 2, 3]), 4.0 * np.eye(3))
    K.assemble(np.array([3, 4]), 2.0 * np.eye(2))
    s = HipPCGSolver(ctx, lstol=1e-12, lsiter=50, lsprecond=1)
    x = np.zeros(4)
    with pytest.raises(HipFemError, match="assembleEnd"):
        s.solve(K, np.ones(4), x)                    # contributions not flushed
    K.assembleEnd()
    reason, it, res = s.solve(K, np.ones(4), x)      # diagonal system 4, 4, 6, 2
    assert reason == 1 and res <= 1e-12
    np.testing.assert_allclose(x, [0.25, 0.25, 1.0 / 6.0, 0.5], rtol=1e-14)


def _mises_deck(etype):
    if etype == meshgen.LTRSPACE:
        coords, conn = meshgen.kuhn_tets(3, 3, 4, 1.0, 1.0, 1.3, jitter=0.1, seed=5)
    elif etype == meshgen.QSPACE:
        coords, conn = meshgen._rve_mesh(2, meshgen.QSPACE)
        coords = coords * np.array([1.0, 1.0, 1.3])
        rng = np.random.default_rng(5)
        inner = np.all((coords > 1e-9) & (coords < np.array([1.0, 1.0, 1.3]) - 1e-9), axis=1)
        coords[inner] += rng.uniform(-0.03, 0.03, (int(inner.sum()), 3))
    else:
        d0 = meshgen.cantilever_hex(3, 3, 4, lz=1.3, jitter=0.2, seed=5)
        coords, conn = d0.coords, d0.conn
    bc = np.zeros((coords.shape[0], 3), bool)
    bc[coords[:, 2] <= 1e-12] = True
    mat = np.array([[meshgen.MAT_MISES, 210e3, 0.3, 250.0, 2e3, 0, 0, 0]])
    return meshgen.Deck("mises", etype, coords, conn, np.zeros(len(conn), np.int32), mat, bc, np.zeros_like(coords))


@pytest.mark.parametrize("etype", [meshgen.LSPACE, meshgen.LTRSPACE, meshgen.QSPACE])
def test_mises_internal_forces_state_and_tangent(ctx, etype):
    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx
    d = _mises_deck(etype)
    rng = np.random.default_rng(4)
    up = np.zeros(d.npeq)
    u1 = rng.standard_normal(d.neq) * 4e-3          # large enough to yield most points
    ev = BatchedEvaluator(ctx, d)
    K = HipCSRSparseMtrx(ctx)
    ev.build_structure(K)
    st = oracle.state_new(d)
    for step, u in enumerate((u1, 1.6 * u1)):
        f, fp, ebe = ev.internal_forces(u, up, update_state=True)
        fo, fpo, ebeo = oracle.internal_forces(d, u, up, state=st, update_state=True)
        scale = np.abs(fo).max()
        assert np.abs(f - fo).max() <= 1e-12 * scale
        assert np.abs(fp - fpo).max() <= 1e-12 * scale
        np.testing.assert_allclose(ebe, ebeo, rtol=1e-12)
        gs = ev.get_state()
        os_ = st.reshape(-1, 38)
        assert (os_[:, 13] > os_[:, 6]).mean() > 0.3      # many points plastic in this step
        for lo, hi in ((7, 14), (14, 20), (20, 26)):
            sc = max(np.abs(os_[:, lo:hi]).max(), 1e-300)
            assert np.abs(gs[:, lo:hi] - os_[:, lo:hi]).max() <= 1e-11 * sc, (lo, hi)
        ev.assemble_stiffness(K, tangent=True)
        _, _, gval = K.download()
        _, _, oval = oracle.csr(d, state=st, tangent=True)
        assert np.abs(gval - oval).max() <= 1e-12 * np.abs(oval).max()
        ev.commit()
        oracle.state_commit(d, st)


def test_hipcsr_sparsemtrx_api(ctx):
    """buildInternalStructure from location arrays, assemble(loc, K_e),
    at(), zero(), version counter, timesT, add, scale."""
    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx, HipFemError
    d = DECKS["lspace"]
    loc = np.where(d.eqn[d.conn] > 0, d.eqn[d.conn], 0).reshape(d.nelem, -1)
    A = HipCSRSparseMtrx(ctx)
    A.buildInternalStructure(d.neq, loc)
    B = HipCSRSparseMtrx(ctx)
    ev = BatchedEvaluator(ctx, d)
    ev.build_structure(B)
    ra, ca, _ = A.download()
    rb, cb, _ = B.download()
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(ca, cb)
    # host-side element scatter through assemble(loc, mat)
    v0 = A.giveVersion()
    A.assembleBegin()
    for e in range(d.nelem):
        X = d.coords[d.conn[e]]
        mp = d.mat[d.emat[e]]
        Ke = oracle.element_stiffness(d.etype, X, oracle.isole_D(mp[1], mp[2]))
        A.assemble(loc[e], Ke)
    A.assembleEnd()
    assert A.giveVersion() > v0
    ev.assemble_stiffness(B)
    _, _, va = A.download()
    _, _, vb = B.download()
    assert np.abs(va - vb).max() <= 1e-13 * np.abs(va).max()
    i, j = int(loc[0][loc[0] > 0][0]), int(loc[0][loc[0] > 0][-1])
    assert A.at(i, j) == va[np.searchsorted(ca[ra[i - 1]:ra[i]], j - 1) + ra[i - 1]]
    A.set_at(i, j, 7.5)
    assert A.at(i, j) == 7.5
    assert not A.isAllocatedAt(1, d.neq)           # far-apart nodes share no element
    assert A.at(1, d.neq) == 0.0                     # missing entries read as zero ...
    with pytest.raises(HipFemError):
        A.set_at(1, d.neq, 1.0)                      # ... and cannot be written
    with pytest.raises(HipFemError):
        A.at(0, 1)                                   # 1-based bounds (SparseMtrx::checkBounds)
    x = np.random.default_rng(0).standard_normal(d.neq)
    B.times_scalar(2.0)
    np.testing.assert_allclose(B.times(x), 2.0 * (sp_from(vb, cb, rb) @ x), rtol=1e-12, atol=1e-9)
    A.zero()
    assert np.abs(A.download()[2]).max() == 0.0


def sp_from(v, c, r):
    import scipy.sparse as sp
    return sp.csr_matrix((v, c, r))


def test_pcg_edge_cases(ctx):
    from oofem_amd.hipfem import CR_CONVERGED, CR_DIVERGED_ITS, HipCSRSparseMtrx, HipFemError, HipPCGSolver
    d = DECKS["lspace"]
    rp, ci, val = oracle.csr(d)
    K = HipCSRSparseMtrx(ctx)
    K.buildInternalStructureCSR(rp, ci)
    K.upload_values(val)
    s = HipPCGSolver(ctx, lstol=1e-10, lsiter=1000, lsprecond=1)
    x = np.zeros(d.neq)
    assert s.solve(K, np.zeros(d.neq), x) == (CR_CONVERGED, 0, 0.0)   # b = 0: |b| := 1, r = 0
    assert not x.any()
    b = np.random.default_rng(1).standard_normal(d.neq)
    s2 = HipPCGSolver(ctx, lstol=1e-12, lsiter=3, lsprecond=1)
    reason, it, res = s2.solve(K, b, np.zeros(d.neq))
    assert reason == CR_DIVERGED_ITS and it == 3
    # the oracle's CG after 3 iterations gives the same residual
    xo, ito, reso, rco = oracle.pcg(rp, ci, val, b, tol=1e-12, maxit=3)
    assert rco == 1 and abs(res - reso) <= 1e-12 * reso
    # warm start: a converged x is returned at once
    x = np.zeros(d.neq)
    s.solve(K, b, x)
    assert s.solve(K, b, x)[1] == 0
    # zero diagonal -> DiagPreconditioner error
    val2 = val.copy()
    val2[np.searchsorted(ci[rp[0]:rp[1]], 0) + rp[0]] = 0.0
    K.upload_values(val2)
    with pytest.raises(HipFemError, match="zero diagonal"):
        s.solve(K, b, np.zeros(d.neq))


def test_mixed_precision_refinement(ctx):
    from oofem_amd.drivers import linear_static
    d = meshgen.cantilever_hex(8, 8, 24, jitter=0.2, e_sigma=0.5, seed=3)
    ref = linear_static(ctx, d, lstol=1e-12)
    got = linear_static(ctx, d, lstol=1e-10, precision=1, inner_tol=1e-4)
    assert got["outer"] >= 2
    assert got["resid"] <= 1e-10
    assert np.linalg.norm(got["u"] - ref["u"]) / np.linalg.norm(ref["u"]) < 1e-8


@pytest.mark.parametrize("operator", ["element", "nodeblock", "csr"])
def test_mixed_precision_reliable_updates(ctx, operator):
    """precision 2: FP32 PCG with FP64 reliable updates, with the element-by-
    element FP32 operator or the FP32 copy of the matrix (node-block and
    scalar CSR).  Convergence is decided on the FP64 true residual, so the
    returned residual is the true one, and the solution agrees with the FP64
    solve to the tolerance's level; iterations stay close to the FP64 count
    because the Krylov direction is kept across updates."""
    from oofem_amd.drivers import LinearStaticProblem
    d = meshgen.cantilever_hex(8, 8, 24, jitter=0.2, e_sigma=0.5, seed=3)
    ref = LinearStaticProblem(ctx, d, lstol=1e-12).solve()
    P = LinearStaticProblem(ctx, d, lstol=1e-10, precision=2)
    P.K.set_matrix_free(operator == "element")
    P.K.set_spmv_nodeblock(operator == "nodeblock")
    got = P.solve()
    assert P.solver.last_matrix_free() == (operator == "element")
    assert got["resid"] <= 1e-10
    assert got["outer"] >= 2                       # reliable updates happened
    # true residual of the returned solution, FP64, through the oracle's SpMV
    rp, ci, val = oracle.csr(d)
    fint, _, _ = oracle.internal_forces(d, np.zeros(d.neq), d.prescribed_values())
    rhs = d.external_load()[0] - fint
    r = rhs - oracle.spmv(rp, ci, val, got["u"])
    assert np.linalg.norm(r) <= 1.0001e-10 * np.linalg.norm(rhs)
    assert np.linalg.norm(got["u"] - ref["u"]) / np.linalg.norm(ref["u"]) < 1e-8
    it64 = LinearStaticProblem(ctx, d, lstol=1e-10).solve()["iters"]
    # iteration growth against FP64 is a property of the FP32 operator's
    # accuracy: the element form sums element forces in FP64, while FP32
    # stiffness values lose about (L/h)^2 ulps on smooth directions
    assert got["iters"] <= (2.0 if operator == "element" else 2.5) * it64, (got["iters"], it64)


@pytest.mark.parametrize("form", [0, 1, 2])
@pytest.mark.parametrize("plastic", [False, True])
def test_element_operator_matches_assembled_matrix(ctx, plastic, form):
    """The FP32 element-by-element operator equals the assembled FP64 matrix
    to FP32 accuracy: elastic (IsoLE, log-normal E) and the MisesMat
    consistent tangent of a plastic state (descriptor form c I_dev + K 1x1 -
    m m^T of the tangent)."""
    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx
    from tests import nlcases
    if plastic:
        d = meshgen.cantilever_hex(4, 4, 6, jitter=0.2, seed=5)
        d.mat = nlcases.MISES.copy()
        d.emat[:] = 0
    else:
        d = meshgen.cantilever_hex(6, 6, 10, jitter=0.2, e_sigma=0.5, seed=4)
    ev = BatchedEvaluator(ctx, d)
    K = HipCSRSparseMtrx(ctx)
    ev.build_structure(K)
    if plastic:
        rng = np.random.default_rng(0)
        u = rng.standard_normal(d.neq) * 3e-3       # strains well past yield
        ev.internal_forces(u, np.zeros(d.npeq), update_state=True)
        s = ev.get_state()
        assert (s[:, 13] > s[:, 6]).mean() > 0.3
    ev.assemble_stiffness(K, tangent=plastic)
    ev.set_element_operator(form)
    assert K.is_matrix_free()
    rng = np.random.default_rng(1)
    for _ in range(3):
        x = rng.standard_normal(d.neq)
        y64 = K.times(x)
        y32 = K.times_matrix_free(x)
        assert np.linalg.norm(y32 - y64) <= 2e-5 * np.linalg.norm(y64)
        if form == 2:
            # the slots are summed in color order: the per-color launches' result, bit for bit
            ev.set_element_operator(0)
            y0 = K.times_matrix_free(x)
            ev.set_element_operator(2)
            assert np.array_equal(y0, y32)
    K.zero()                                          # any change of the values drops the view
    assert not K.is_matrix_free()


def test_mixed_precision_operator_forms_same_bits(ctx):
    """Inside the mixed PCG the slot form sums each equation's colour slots
    in one kernel together with the p.q partials (on an eight-colour mesh
    every slot is read at once: the slots of colours with no element at an
    equation hold zero); the per-colour form adds each colour's launch into
    the FP64 product and then forms p.q.  They are the same sums in the same
    order, so the solves take the same iterations to the same bits."""
    from oofem_amd.drivers import LinearStaticProblem
    d = meshgen.cantilever_hex(8, 8, 24, jitter=0.2, e_sigma=0.5, seed=3)
    out = []
    for form in (2, 0):
        P = LinearStaticProblem(ctx, d, lstol=1e-10, precision=2)
        P.K.set_matrix_free(True)
        P.ev.set_element_operator(form)
        assert P.ev.ncolors() == 8
        got = P.solve()
        assert P.solver.last_matrix_free()
        out.append(got)
    assert out[0]["iters"] == out[1]["iters"] and out[0]["outer"] == out[1]["outer"]
    assert np.array_equal(out[0]["u"], out[1]["u"])


@pytest.mark.parametrize("plastic", [False, True])
def test_element_operator_equation_lookup_forms_agree(ctx, plastic):
    """The element operator finds a lane's equations in one stored word when
    every node numbers its free dofs consecutively (node-by-node numbering),
    and through the connectivity and the equation table otherwise.  The same
    mesh and state under both numberings (the second a random permutation of
    the free equations) gives the same element forces, so the same products
    bit for bit once permuted back, in all three operator forms."""
    import copy

    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx
    from tests import nlcases
    d = meshgen.cantilever_hex(5, 4, 7, jitter=0.2, e_sigma=0.5, seed=9)
    if plastic:
        d.mat = nlcases.MISES.copy()
        d.emat[:] = 0
    rng = np.random.default_rng(2)
    perm = rng.permutation(d.neq)          # old equation k -> new perm[k]
    d2 = copy.deepcopy(d)
    free = d2.eqn > 0
    d2.eqn[free] = perm[d2.eqn[free] - 1] + 1
    assert any((np.diff(r[r > 0]) != 1).any() for r in d2.eqn)   # the table lookup is taken
    u = rng.standard_normal(d.neq) * (3e-3 if plastic else 1e-3)
    xs = [rng.standard_normal(d.neq) for _ in range(2)]
    out = []
    for deck, to_new in ((d, np.arange(d.neq)), (d2, perm)):
        ev = BatchedEvaluator(ctx, deck)
        K = HipCSRSparseMtrx(ctx)
        ev.build_structure(K)
        un = np.empty(d.neq)
        un[to_new] = u
        ev.internal_forces(un, np.zeros(deck.npeq), update_state=True)
        ev.assemble_stiffness(K, tangent=plastic)
        ys = []
        for form in (0, 1, 2):
            ev.set_element_operator(form)
            assert K.is_matrix_free()
            for x in xs:
                xn = np.empty(d.neq)
                xn[to_new] = x
                ys.append(K.times_matrix_free(xn)[to_new])
        out.append(ys)
        K.close()
        ev.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    # and the products are the operator's, not zeros
    assert np.linalg.norm(out[0][0]) > 0


@pytest.mark.parametrize("precision", [0, 2])
def test_pcg_multiple_right_hand_sides(ctx, precision):
    """solve(A, B, X) with eight columns, one of them zero, on config 0: every
    column meets the tolerance on its true residual (checked in numpy on the
    oracle's CSR) and the zero column stays zero.  Precision 2 with zero
    initial guesses takes the batched path (six columns per batched solve),
    precision 0 the column-by-column loop of the reference's default."""
    import scipy.sparse as sp

    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx, HipPCGSolver
    d = meshgen.config0()
    rp, ci, val = oracle.csr(d)
    A = sp.csr_matrix((val, ci, rp), shape=(d.neq, d.neq))
    K = HipCSRSparseMtrx(ctx)
    ev = BatchedEvaluator(ctx, d)
    ev.build_structure(K)
    ev.assemble_stiffness(K)
    rng = np.random.default_rng(3)
    B = rng.standard_normal((d.neq, 8))
    B[:, 7] = 0.0
    X = np.zeros_like(B, order="F")
    tol = 1e-9
    s = HipPCGSolver(ctx, lstol=tol, lsiter=100000, lsprecond=1, precision=precision)
    reason, it, ach = s.solve_multi(K, B, X)
    assert reason == 1 and it > 0
    for j in range(7):
        r = np.linalg.norm(B[:, j] - A @ X[:, j]) / np.linalg.norm(B[:, j])
        assert r <= 5 * tol, (j, r)
    assert not X[:, 7].any()
    s.close()
