"""The C++ host layer (class factory, HipCSRSparseMtrx, HipPCGSolver,
LinearStatic and StaticStructural/NRSolver drivers) on the GPU, against the
oracle and closed forms."""
import numpy as np
import pytest

from oofem_amd import meshgen
from oracle import oracle
from tests import nlcases

pytestmark = pytest.mark.gpu


def test_factory_builds_hip_plugins_and_linear_static_parity():
    from oofem_amd.problem import StructuralProblem
    from oofem_amd.drivers import reaction_sums
    d = meshgen.config0()
    P = StructuralProblem(d, lstol=1e-10)
    f, fp = d.external_load()
    P.set_prescribed(d.prescribed_values())
    P.set_load(f, fp)
    it, res = P.linear_static()
    assert P.matrix_class() == "HipCSRSparseMtrx"
    ref = oracle.linear_static(d, tol=1e-10)
    u = P.solution()
    R = P.reactions()
    assert abs(it - ref["iters"]) <= 1
    assert np.linalg.norm(u - ref["u"]) / np.linalg.norm(ref["u"]) <= 1e-10
    assert np.linalg.norm(R - ref["reactions"]) / np.linalg.norm(ref["reactions"]) <= 1e-10
    assert abs(reaction_sums(d, R)[1] - 1.0) < 1e-9
    # SparseMtrx::at through the factory-made matrix
    rp, ci, val = oracle.csr(d)
    assert P.matrix_at(1, 1) == pytest.approx(val[rp[0] + np.searchsorted(ci[rp[0]:rp[1]], 0)], rel=1e-13)


def test_unregistered_matrix_type_is_an_error():
    from oofem_amd.problem import ProblemError, StructuralProblem
    d = meshgen.cantilever_hex(2, 2, 3)
    P = StructuralProblem(d, smtype=4)            # SymCompCol is not part of this build
    P.set_prescribed(d.prescribed_values())
    P.set_load(*d.external_load())
    with pytest.raises(ProblemError, match="not registered"):
        P.linear_static()


@pytest.mark.parametrize("etype", [meshgen.LSPACE, meshgen.LTRSPACE])
def test_newton_uniaxial_closed_form_gpu(etype):
    from oofem_amd.problem import StructuralProblem
    d, top = nlcases.uniaxial_block(etype)
    ey = nlcases.SIG0 / nlcases.E
    P = StructuralProblem(d, lstol=1e-12, rtolf=1e-10, maxiter=30, manrmsteps=1)
    P.set_load(np.zeros(d.neq), np.zeros(d.npeq))
    for k in range(1, 7):
        eps = k * 0.6 * ey
        P.set_prescribed(d.prescribed_values(eps))
        rep = P.nonlinear_step()
        assert rep["status"] == 1 and rep["force_error"] <= 1e-10
        R = P.reactions()
        P.update()
        Rz = nlcases.top_reaction_z(d, top, R)
        expect = nlcases.uniaxial_stress(eps)
        assert abs(Rz - expect) <= 1e-8 * expect, (k, Rz, expect)


def test_newton_tet_bar_matches_oracle_step_by_step():
    """Jittered tet bar with MisesMat, end displacement ramped over 5 steps,
    full Newton: per step the Newton iteration count, the displacements and
    the Gauss-point state match the oracle's."""
    from oofem_amd.problem import StructuralProblem
    d = nlcases.tet_bar()
    P = StructuralProblem(d, lstol=1e-13, rtolf=1e-8, maxiter=40, manrmsteps=1)
    P.set_load(np.zeros(d.neq), np.zeros(d.npeq))
    st = oracle.state_new(d)
    X = np.zeros(d.neq)
    nsteps = 5
    prev = np.zeros(d.npeq)
    for k in range(1, nsteps + 1):
        up = d.prescribed_values(k / nsteps)
        P.set_prescribed(up)
        rep = P.nonlinear_step()
        nit, cg, status, err = oracle.nonlinear_step(d, st, X, up, rtolf=1e-8, maxiter=40, manrmsteps=1, lstol=1e-13,
                                                     u_presc_prev=prev)
        prev = up
        assert rep["status"] == 1 and status == 1
        assert rep["newton"] == nit, (k, rep, nit)
        u = P.solution()
        assert np.linalg.norm(u - X) <= 1e-9 * np.linalg.norm(X)
        gs = P.state()
        os_ = st.reshape(-1, 38)
        assert np.abs(gs[:, 13] - os_[:, 13]).max() <= 1e-9 * max(os_[:, 13].max(), 1e-30)
        P.update()
        oracle.state_commit(d, st)
    assert (st.reshape(-1, 38)[:, 6] > 0).mean() > 0.2      # the bar has yielded


@pytest.mark.parametrize("mixed", [0, 2])
def test_inexact_newton_reaches_the_exact_newton_state(mixed):
    """Plasticity block 8x8x32 (MisesMat), end displacement ramped over 8
    steps, first 3 steps: with forcing terms up to 0.1 every step converges
    to rtolf 1e-8, the displacements and the plastic strain stay within 1e-7
    of the oracle's exact Newton (corrections solved to 1e-12), and the
    steps take fewer PCG iterations in total than exact Newton on the GPU."""
    from oofem_amd.problem import StructuralProblem
    d = meshgen.config3(8, 8, 32)
    st = oracle.state_new(d)
    X = np.zeros(d.neq)
    prev = np.zeros(d.npeq)
    pcg = {}
    sol = {}
    for forcing in (0.0, 0.1):
        P = StructuralProblem(d, lstol=1e-10, rtolf=1e-8, maxiter=60, manrmsteps=1, mixed=mixed, forcing=forcing)
        P.set_load(np.zeros(d.neq), np.zeros(d.npeq))
        total = 0
        for k in range(1, 4):
            P.set_prescribed(d.prescribed_values(k / 8))
            rep = P.nonlinear_step()
            assert rep["status"] == 1 and rep["force_error"] <= 1e-8, (forcing, k, rep)
            total += rep["pcg"]
            P.update()
        pcg[forcing] = total
        sol[forcing] = (P.solution(), P.state()[:, 6].copy())
        P.close()
    for k in range(1, 4):
        up = d.prescribed_values(k / 8)
        nit, cg, status, err = oracle.nonlinear_step(d, st, X, up, rtolf=1e-8, maxiter=60, manrmsteps=1, lstol=1e-12,
                                                     u_presc_prev=prev)
        assert status == 1
        prev = up
        oracle.state_commit(d, st)
    kappa = st.reshape(-1, 38)[:, 6]
    assert (kappa > 0).mean() > 0.05                          # part of the block has yielded
    for forcing, (u, kap) in sol.items():
        assert np.linalg.norm(u - X) <= 1e-7 * np.linalg.norm(X), forcing
        assert np.abs(kap - kappa).max() <= 1e-7 * kappa.max(), forcing
    assert pcg[0.1] < 0.7 * pcg[0.0], pcg


@pytest.mark.parametrize("tangent", [False, True])
def test_tangent_product_matches_oracle(ctx, tangent):
    """hfem_model_tangent_product_dev (sum_g B^T D B x_e dV) against the
    oracle's element-matrix product K_e du_e over prescribed increments,
    for the elastic and the plastic temp state."""
    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx
    d = nlcases.tet_bar()
    st = oracle.state_new(d)
    X = np.zeros(d.neq)
    up = d.prescribed_values(0.6)
    if tangent:
        nit, cg, status, err = oracle.nonlinear_step(d, st, X, up, rtolf=1e-10, maxiter=40, manrmsteps=1)
        assert status == 1
    ev = BatchedEvaluator(ctx, d)
    K = HipCSRSparseMtrx(ctx)
    ev.build_structure(K)
    if tangent:
        ev.internal_forces(X, up, update_state=True)   # temp state on the device from the same point
    du = d.prescribed_values(1.0)
    y, yp = ev.tangent_product(None, du, tangent=True)
    ref = oracle.extrapolated_forces(d, st, du)
    assert np.abs(y - ref).max() <= 1e-10 * np.abs(ref).max()


def test_line_search_matches_oracle():
    """NRSolver with LineSearchNM on the GPU driver (search's point kept): a
    single step to 3/8 of the block's end displacement without the tangent
    predictor converges
    with the oracle's Newton iteration and line-search counts, at its
    displacements (1e-8) and plastic strains."""
    from oofem_amd.problem import StructuralProblem
    d = meshgen.config3(8, 8, 32)
    up = d.prescribed_values(3 / 8)
    st = oracle.state_new(d)
    X = np.zeros(d.neq)
    ref = oracle.nonlinear_step_ex(d, st, X, up, initial_guess=False, maxiter=40, lstol=1e-12, linesearch=True,
                                   ls_keep_point=True)
    assert ref["status"] == 1 and ref["line_searches"] >= 1
    P = StructuralProblem(d, lstol=1e-12, rtolf=1e-8, maxiter=40, manrmsteps=1, initial_guess=0)
    P.set_nr_extras(linesearch=True)
    P.set_nr_criteria(ls_keep_point=True)
    P.set_load(np.zeros(d.neq), np.zeros(d.npeq))
    P.set_prescribed(up)
    rep = P.nonlinear_step()
    assert rep["status"] == 1 and rep["newton"] == ref["newton"], (rep, ref)
    u = P.solution()
    assert np.linalg.norm(u - X) <= 1e-8 * np.linalg.norm(X)
    kap = P.state()[:, 13]                    # temp kappa (the step is not committed)
    kref = st.reshape(-1, 38)[:, 13]
    assert np.abs(kap - kref).max() <= 1e-8 * kref.max()
    P.close()


def test_time_step_reduction_matches_oracle():
    """SimpleReductionStrategy time stepping on the GPU driver against the
    oracle's driver: the same accepted steps (time, Newton iterations,
    restarts) and the same final displacements and plastic strains."""
    from oofem_amd.problem import StructuralProblem
    d = meshgen.config3(8, 8, 32)
    up_ref = d.prescribed_values(1.0)
    red = dict(dtmin=0.01, dtmax=1.0)
    X, st, ref = oracle.static_structural(d, up_ref, 1.0, 1.0, reduction=red, maxiter=8, lstol=1e-12)
    P = StructuralProblem(d, lstol=1e-12, rtolf=1e-8, maxiter=8, manrmsteps=1)
    P.set_load(np.zeros(d.neq), np.zeros(d.npeq))
    got = P.run_steps(up_ref, 1.0, 1.0, reduction=red)
    key = lambda s: (round(s["time"], 12), s["newton"], s["restarts"], s["status"])
    assert [key(s) for s in got] == [key(s) for s in ref], (got, ref)
    assert any(s["restarts"] for s in got)
    u = P.solution()
    assert np.linalg.norm(u - X) <= 1e-8 * np.linalg.norm(X)
    kap = P.state()[:, 6]
    kref = st.reshape(-1, 38)[:, 6]
    assert np.abs(kap - kref).max() <= 1e-8 * kref.max()
    P.close()


@pytest.mark.parametrize("name", ["patch301.in", "patch300.in", "cantilever_Qspace.in", "deadweight02.in",
                                  "tutorialmaterial.in"])
def test_run_reference_deck_end_to_end(tmp_path, name):
    """A reference deck through the deck reader, the C++ StaticStructural
    driver behind the hipcsr / hippcg plug-ins (the deck's own keywords plus
    smtype 11 lstype 9 lstol 1e-12), the .out writer and the errorcheck
    block: every check holds on the computed values, and every #NODE check
    holds on the .out file read back.  tutorialmaterial prescribes every
    DOF (no free equations): each step is one return map per Gauss point
    and the Newton loop converges at its first check, as NRSolver does."""
    import os

    from oofem_amd.deck import read_deck
    from oofem_amd.outfile import check_out, read_out
    from oofem_amd.run import run_deck
    from tests.conftest import GOLDEN
    path = os.path.join(GOLDEN, name)
    out = tmp_path / "deck.out"
    r = run_deck(path, out_path=str(out), keywords=dict(smtype=11, lstype=9, lstol=1e-12))
    bad = [(c.line, v) for c, v, ok in r["checks"] if not ok]
    assert r["checks"] and not bad, bad
    prob = read_deck(path)
    steps = read_out(out.read_text())
    assert len(steps) == prob.nsteps
    oc = check_out(prob, steps)
    assert oc and all(ok for c, v, ok in oc if c.kind == "node"), [(c.line, v) for c, v, ok in oc if not ok]


def test_run_deck_default_solver_is_rejected():
    """The deck's own keywords name the reference's default skyline matrix,
    which is not part of this build: the class factory refuses it."""
    import os

    from oofem_amd.problem import ProblemError
    from oofem_amd.run import run_deck
    from tests.conftest import GOLDEN
    with pytest.raises(ProblemError):
        run_deck(os.path.join(GOLDEN, "patch301.in"))


@pytest.mark.parametrize("mode", [2, 1, 3])
def test_elastic_stiffness_newton_matches_oracle(mode):
    """NonLinearStatic stiffmode 2 on the GPU driver (the elastic matrix is
    assembled once and kept): the oracle's Newton iteration count and root.
    The secant modes 1 and 3 reach the same iterates, because IsoLE and
    MisesMat answer SecantStiffness with the elastic matrix
    (misesmat.C:496-504)."""
    from oofem_amd.problem import StructuralProblem
    d = meshgen.config3(4, 4, 16)
    d.mat = d.mat.copy()
    d.mat[0, 4] = 5e4
    up = d.prescribed_values(3 / 8)
    st = oracle.state_new(d)
    X = np.zeros(d.neq)
    ref = oracle.nonlinear_step_ex(d, st, X, up, initial_guess=False, maxiter=400, lstol=1e-12, stiffmode=2)
    assert ref["status"] == 1
    P = StructuralProblem(d, lstol=1e-12, rtolf=1e-8, maxiter=400, manrmsteps=1, initial_guess=0, stiff_mode=mode)
    P.set_load(np.zeros(d.neq), np.zeros(d.npeq))
    P.set_prescribed(up)
    rep = P.nonlinear_step()
    assert rep["status"] == 1 and abs(rep["newton"] - ref["newton"]) <= 1, (rep, ref)
    assert np.linalg.norm(P.solution() - X) <= 1e-8 * np.linalg.norm(X)
    P.close()


def test_run_nonlinear_static_deck(tmp_path):
    """patch301 rewritten as a NonLinearStatic deck (direct control, elastic
    stiffness, one metastep of one step): every errorcheck record holds."""
    import os

    from oofem_amd.run import run_deck
    from tests.conftest import GOLDEN
    src = open(os.path.join(GOLDEN, "patch301.in")).read().splitlines()
    i = next(k for k, l in enumerate(src) if l.lower().startswith("staticstructural"))
    src[i] = "NonLinearStatic nsteps 1 nmsteps 1 nmodules 1"
    src.insert(i + 1, "nsteps 1 controlmode 1 stiffmode 2 rtolf 1e-10")
    p = tmp_path / "nls.in"
    p.write_text("\n".join(src) + "\n")
    r = run_deck(str(p), keywords=dict(smtype=11, lstype=9, lstol=1e-12))
    bad = [(c.line, v) for c, v, ok in r["checks"] if not ok]
    assert r["checks"] and not bad, bad
    with pytest.raises(ValueError, match="steplength"):
        # indirect control needs its arc length
        src[i + 1] = "nsteps 1 controlmode 0"
        p.write_text("\n".join(src) + "\n")
        run_deck(str(p), keywords=dict(smtype=11, lstype=9))


def test_direct_displacement_control_matches_oracle():
    """NRSolver ddm/ddv on the GPU driver: the uniaxial block past yield with
    its top face's u_z free and driven by the 1e6 diagonal penalty reaches
    the oracle's Newton iteration count, displacements (1e-8) and plastic
    strains."""
    from oofem_amd.problem import StructuralProblem
    from tests import nlcases
    d, top = nlcases.uniaxial_block(meshgen.LSPACE, n=3)
    bc2 = d.bc_mask.copy()
    bc2[top, 2] = False
    d2 = meshgen.Deck("uniaxial_ddm", d.etype, d.coords, d.conn, d.emat, d.mat, bc2, np.zeros_like(d.coords))
    eqs = d2.eqn[top, 2]
    eps = 2.0 * nlcases.SIG0 / nlcases.E
    vals = np.full(len(eqs), eps)
    st = oracle.state_new(d2)
    X = np.zeros(d2.neq)
    ref = oracle.nonlinear_step_ex(d2, st, X, np.zeros(d2.npeq), initial_guess=False, maxiter=40, lstol=1e-13,
                                   ddm=(eqs, vals), ddm_factor=1.0)
    assert ref["status"] == 1
    P = StructuralProblem(d2, lstol=1e-13, rtolf=1e-8, maxiter=40, manrmsteps=1, initial_guess=0)
    P.set_load(np.zeros(d2.neq), np.zeros(d2.npeq))
    P.set_prescribed(np.zeros(d2.npeq))
    P.set_ddm(eqs, vals)
    P.set_ddm_factor(1.0)
    rep = P.nonlinear_step()
    assert rep["status"] == 1 and rep["newton"] == ref["newton"], (rep, ref)
    assert np.linalg.norm(P.solution() - X) <= 1e-8 * np.linalg.norm(X)
    assert np.abs(P.solution()[eqs - 1] - eps).max() <= 1e-5 * eps
    kap = P.state()[:, 13]
    kref = st.reshape(-1, 38)[:, 13]
    assert np.abs(kap - kref).max() <= 1e-8 * kref.max()
    P.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["full_newton", "modified_restart", "psi0", "hpc1", "hpc2"])
def test_arc_length_steps_match_oracle(case):
    """NonLinearStatic indirect control (CylindricalALM) through the C++
    driver on the GPU against the oracle's restatement, step by step, on the
    traction-loaded MisesMat block through yield: the same corrector
    iteration counts and restarts, load levels to 1e-10 and displacements to
    1e-9; the reactions balance the reached load level times the reference
    load."""
    from oracle import oracle
    from oofem_amd.problem import StructuralProblem
    from tests import nlcases
    d, top, R, Rp = nlcases.traction_block()
    opts = dict(full_newton=dict(steplength=20.0, manrmsteps=1, rtolf=1e-10, rtold=1e-10),
                modified_restart=dict(steplength=20.0, rtolf=1e-8, rtold=1e-8),
                psi0=dict(steplength=2e-3, psi=0.0, manrmsteps=1, rtolf=1e-10, rtold=1e-10),
                hpc1=dict(steplength=4e-3, psi=0.0, manrmsteps=1, rtolf=1e-10, rtold=1e-10),
                hpc2=dict(steplength=4e-3, manrmsteps=1, rtolf=1e-10, rtold=1e-10))[case]
    qz = [int(q) for q in d.eqn[top, 2]]
    hpc = dict(hpc1=(1, qz, None), hpc2=(2, qz[:1] + [0], [1.0, 0.5])).get(case)
    okw = dict(hpcmode=hpc[0], hpc_eqs=hpc[1], hpc_weights=hpc[2] or ()) if hpc else {}
    al = oracle.ArcLength(lstol=1e-13, **opts, **okw)
    st = oracle.state_new(d)
    Xo = np.zeros(d.neq)
    P = StructuralProblem(d, lstol=1e-13, initial_guess=0)
    try:
        P.set_arc_length(**opts)
        if hpc:
            P.set_arc_length_hpc(*hpc)
        P.set_reference_load(R, Rp)
        P.set_prescribed(np.zeros(d.npeq))
        restarts = 0
        for k in range(7):
            ro = al.step(d, st, Xo, np.zeros(d.npeq), R)
            oracle.state_commit(d, st)
            rg = P.arc_length_step()
            P.update()
            assert rg["status"] == 1 and ro["converged"]
            assert (rg["iterations"], rg["restarts"]) == (ro["iterations"], ro["restarts"]), (k, rg, ro)
            assert abs(rg["load_level"] - ro["load_level"]) <= 1e-10 * abs(ro["load_level"])
            assert abs(rg["step_length"] - ro["step_length"]) <= 1e-10 * ro["step_length"]
            u = P.solution()
            assert np.abs(u - Xo).max() <= 1e-9 * np.abs(Xo).max(), k
            reac = P.reactions()
            # bottom reactions carry the load: sum of z reactions = -lambda (unit traction on a unit face)
            bz = d.eqn[:, 2] < 0
            assert abs(reac[-d.eqn[bz, 2] - 1].sum() + rg["load_level"]) <= 1e-8 * rg["load_level"]
            restarts += rg["restarts"]
        if case == "modified_restart":
            assert restarts >= 1
    finally:
        P.close()


@pytest.mark.gpu
def test_arc_length_deck_end_to_end(tmp_path):
    """patch301 (linear LTRSpace, surface load) as a NonLinearStatic deck
    with indirect control (controlmode 0, CylindricalALM keywords on the
    metastep): every #LOADLEVEL and #REACTION check holds, their values being
    the closed form of a linear arc-length run (lambda_k = sum of deltaL_j / p,
    deltaL growing by 3^(1/4) per one-iteration step up to steplength, the
    reactions lambda_k times the exact 1.5), and the .out file carries the
    reached load level of every step."""
    import os

    from oofem_amd.deck import read_deck
    from oofem_amd.outfile import read_out
    from oofem_amd.run import run_deck
    from tests.conftest import GOLDEN
    src = open(os.path.join(GOLDEN, "patch301.in")).read().splitlines()
    i = next(k for k, l in enumerate(src) if l.lower().startswith("staticstructural"))
    src[i] = "NonLinearStatic nsteps 1 nmsteps 1 nmodules 1"
    src.insert(i + 1, "nsteps 4 controlmode 0 psi 1.0 steplength 0.5 initialsteplength 0.2 rtolv 1e-10")
    j = next(k for k, l in enumerate(src) if l.startswith("#%BEGIN_CHECK%"))
    base = tmp_path / "base.in"
    base.write_text("\n".join(src[:j]) + "\n")
    prob = read_deck(str(base))
    d = prob.deck
    _, (f, _) = prob.step_loading(1)
    rp, ci, val = oracle.csr(d)
    KR, _, _, _ = oracle.pcg(rp, ci, val, f, tol=1e-14)
    p = np.sqrt(KR @ KR + f @ f)
    lam, dl, checks = 0.0, 0.2, ["#%BEGIN_CHECK% tolerance 1.e-8"]
    levels = []
    for k in range(1, 5):
        lam += dl / p
        dl = min(dl * 3.0 ** 0.25, 0.5)
        levels.append(lam)
        checks.append(f"#LOADLEVEL tStep {k} value {lam:.16e}")
        checks.append(f"#REACTION tStep {k} number 1 dof 3 value {1.5 * lam:.16e}")
    checks.append("#%END_CHECK%")
    path = tmp_path / "calm.in"
    path.write_text("\n".join(src[:j] + checks) + "\n")
    out = tmp_path / "calm.out"
    r = run_deck(str(path), out_path=str(out), keywords=dict(smtype=11, lstype=9, lstol=1e-13))
    assert len(r["checks"]) == 8
    bad = [(c.line, v) for c, v, ok in r["checks"] if not ok]
    assert not bad, bad
    assert all(r["iterations"][k][0] == 1 for k in range(1, 5))
    steps = read_out(out.read_text())
    assert [round(s["load_level"], 6) for s in steps] == [round(v, 6) for v in levels]


@pytest.mark.gpu
def test_arc_length_deck_refuses_what_is_not_built(tmp_path):
    """Dof-group norms, the CALM line search, a missing steplength, hpcmode
    without its map and the linearized control without weights are refused
    before any solve."""
    import os

    from oofem_amd.run import run_deck
    from tests.conftest import GOLDEN
    src = open(os.path.join(GOLDEN, "patch301.in")).read().splitlines()
    i = next(k for k, l in enumerate(src) if l.lower().startswith("staticstructural"))
    for extra, msg in (("steplength 0.5 nccdg 1", "nccdg"), ("steplength 0.5 linesearch 1", "line search"),
                       ("psi 1.0", "steplength"), ("steplength 0.5 hpcmode 2", "HPC Map"),
                       ("steplength 0.5 hpcmode 2 hpc 2 4 3", "weight")):
        s = list(src)
        s[i] = f"NonLinearStatic nsteps 2 controlmode 0 {extra} nmodules 1"
        p = tmp_path / "bad.in"
        p.write_text("\n".join(s) + "\n")
        with pytest.raises(ValueError, match=msg):
            run_deck(str(p), keywords=dict(smtype=11, lstype=9))


@pytest.mark.gpu
def test_nonlinear_static_direct_control_accumulates_reference_loads(tmp_path):
    """NonLinearStatic direct control keeps every step's reference load in
    its initial load vector (updateLoadVectors, nlinearstatic.C:419-430,
    loadLevel 1 under NRSolver): patch301 (linear, constant surface load)
    over three steps carries 1, 2 and 3 times the load with refloadmode 0,
    so the displacements and reactions are 1, 2 and 3 times the exact
    one-step solution; with refloadmode 1 the constant function has no
    increment (Function::evaluate VM_Incremental) and nothing moves."""
    import os

    from oofem_amd.run import run_deck
    from tests.conftest import GOLDEN
    src = open(os.path.join(GOLDEN, "patch301.in")).read().splitlines()
    i = next(k for k, l in enumerate(src) if l.lower().startswith("staticstructural"))
    j = next(k for k, l in enumerate(src) if l.startswith("#%BEGIN_CHECK%"))
    out = {}
    for mode in (0, 1):
        s = list(src[:j])
        s[i] = f"NonLinearStatic nsteps 3 controlmode 1 refloadmode {mode} rtolf 1e-10 nmodules 1"
        p = tmp_path / f"nls{mode}.in"
        p.write_text("\n".join(s) + "\n")
        out[mode] = run_deck(str(p), keywords=dict(smtype=11, lstype=9, lstol=1e-13))
    r0 = out[0]["results"]
    u1 = r0[1]["u_node"]
    assert abs(r0[1]["reactions"][(0, 3)] - 1.5) <= 1e-9
    for k in (2, 3):
        assert np.abs(r0[k]["u_node"] - k * u1).max() <= 1e-9 * np.abs(u1).max()
        assert abs(r0[k]["reactions"][(0, 3)] - 1.5 * k) <= 1e-9 * k
    for k in (1, 2, 3):
        assert np.abs(out[1]["results"][k]["u_node"]).max() == 0.0


@pytest.mark.gpu
def test_arc_length_deck_with_linearized_hyperplane_control(tmp_path):
    """patch301 under indirect control with hpcmode 2 on node 4's z dof
    (hpc 2 4 3, hpcw 1 1.0): every step moves that dof by +steplength (the
    weight's direction), so on this linear problem the load level after k
    steps is k steplength over that dof's displacement u4 under the
    reference load; u4 < 0 here, so the load runs backwards."""
    import os

    from oofem_amd.deck import read_deck
    from oofem_amd.run import run_deck
    from tests.conftest import GOLDEN
    src = open(os.path.join(GOLDEN, "patch301.in")).read().splitlines()
    i = next(k for k, l in enumerate(src) if l.lower().startswith("staticstructural"))
    j = next(k for k, l in enumerate(src) if l.startswith("#%BEGIN_CHECK%"))
    s = list(src[:j])
    s[i] = ("NonLinearStatic nsteps 3 controlmode 0 steplength 0.4 hpcmode 2 hpc 2 4 3 hpcw 1 1.0 "
            "rtolv 1e-10 nmodules 1")
    p = tmp_path / "hpc.in"
    p.write_text("\n".join(s) + "\n")
    prob = read_deck(str(p))
    d = prob.deck
    _, (f, _) = prob.step_loading(1)
    rp, ci, val = oracle.csr(d)
    KR, _, _, _ = oracle.pcg(rp, ci, val, f, tol=1e-14)
    n4 = list(prob.node_numbers).index(4)
    u4 = KR[d.eqn[n4, 2] - 1]
    r = run_deck(str(p), keywords=dict(smtype=11, lstype=9, lstol=1e-13))
    assert u4 < 0.0
    for k in (1, 2, 3):
        assert abs(r["results"][k]["u_node"][n4, 2] - 0.4 * k) <= 1e-9
        assert abs(r["results"][k]["load_level"] - 0.4 * k / u4) <= 1e-9 * 0.4 * k / abs(u4)


@pytest.mark.gpu
def test_constrained_newton_matches_oracle():
    """NRSolver constrainednrminiter 1 on the GPU driver: the oracle's
    Newton iteration count and root on a step whose corrections get
    halved."""
    from oofem_amd.problem import StructuralProblem
    d = meshgen.config3(4, 4, 16)
    up = d.prescribed_values(3 / 8)
    st = oracle.state_new(d)
    X = np.zeros(d.neq)
    ref = oracle.nonlinear_step_ex(d, st, X, up, maxiter=60, lstol=1e-12, initial_guess=False, rtolf=1e-8,
                                   constrainednrminiter=1)
    assert ref["status"] == 1
    P = StructuralProblem(d, lstol=1e-12, rtolf=1e-8, maxiter=60, manrmsteps=1, initial_guess=0)
    try:
        P.set_constrained_nr(1)
        P.set_load(np.zeros(d.neq), np.zeros(d.npeq))
        P.set_prescribed(up)
        rep = P.nonlinear_step()
        assert rep["status"] == 1 and rep["newton"] == ref["newton"], (rep, ref)
        assert np.linalg.norm(P.solution() - X) <= 1e-8 * np.linalg.norm(X)
    finally:
        P.close()


@pytest.mark.gpu
def test_run_deck_time_step_reduction_matches_oracle(tmp_path):
    """A StaticStructural deck with treductiontype simplereduction (the
    8x8x32 MisesMat block pulled to 1/1 of its end displacement in one
    step of dt 1, maxiter 8): run_deck's adaptive steps are the oracle
    driver's accepted steps (time, Newton iterations, restarts), and the
    final displacements agree."""
    from oofem_amd.run import run_deck
    from tests.deckwriter import write_block_deck
    d = meshgen.config3(8, 8, 32)
    path = tmp_path / "red.in"
    write_block_deck(d, path, "StaticStructural nsteps 1 nmsteps 1 nmodules 1",
                     "nsteps 1 deltat 1.0 treductiontype simplereduction dtmin 0.01 dtmax 1.0 maxiter 8 "
                     "rtolf 1e-8 manrmsteps 1")
    X, st, ref = oracle.static_structural(d, d.prescribed_values(1.0), 1.0, 1.0, reduction=dict(dtmin=0.01, dtmax=1.0),
                                          maxiter=8, lstol=1e-12)
    # lsiter: the IMLSolver default of 200 iterations is far too few for this mesh
    r = run_deck(str(path), keywords=dict(smtype=11, lstype=9, lstol=1e-12, lsiter=100000))
    res = r["results"]
    got = [(round(res[k]["time"], 12), r["iterations"][k][0], res[k]["restarts"]) for k in sorted(res)]
    assert got == [(round(s["time"], 12), s["newton"], s["restarts"]) for s in ref], (got, ref)
    assert any(s["restarts"] for s in ref)
    u = res[max(res)]["u"]
    assert np.linalg.norm(u - X) <= 1e-8 * np.linalg.norm(X)
