"""GPU tests of the MisesMat surrogate (hfem_mlp): training data against
the oracle's return mapping, FP32 forward and backward against the FP64
restatement in oracle/mlp_oracle.py, training progress, and BF16 MFMA
inference against the FP32 forward."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from oracle import oracle

pytestmark = pytest.mark.gpu

E, NU, SIG0, H = 210e3, 0.3, 250.0, 2e3
MP = np.array([1, E, NU, SIG0, H, 0, 0, 0], dtype=np.float64)


@pytest.fixture(scope="module")
def net(ctx):
    from oofem_amd.hipfem import MisesSurrogate
    n = MisesSurrogate(ctx, E, NU, SIG0, H, seed=3)
    yield n
    n.close()


def test_forward_matches_oracle(net):
    rng = np.random.default_rng(0)
    X = rng.standard_normal((300, 13)).astype(np.float32)
    p = net.params().astype(np.float64)
    got = net.forward(X)
    ref = M.forward(p, X.astype(np.float64))
    assert np.abs(got - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())


def test_gradients_match_oracle(net):
    """Backward kernels (tiled GEMMs, SiLU derivative, column sums) against
    the FP64 gradient, per parameter block."""
    rng = np.random.default_rng(1)
    X = rng.standard_normal((97, 13)).astype(np.float32)       # not a multiple of the 64-row tile
    Y = rng.standard_normal((97, 27)).astype(np.float32)
    loss, g = net.gradients(X, Y)
    lref, gref = M.loss_and_grad(net.params().astype(np.float64), X.astype(np.float64), Y.astype(np.float64))
    assert abs(loss - lref) <= 1e-5 * lref
    offs = np.cumsum([0] + [int(np.prod(s)) for s in M.SHAPES])
    for a, b in zip(offs[:-1], offs[1:]):
        assert np.linalg.norm(g[a:b] - gref[a:b]) <= 1e-4 * np.linalg.norm(gref[a:b]), (a, b)


def test_training_samples_follow_the_return_mapping(net):
    """Every generated sample is the plain return mapping (oracle) of its own
    inputs: stress and consistent tangent at (eps, epsp_n, kappa_n)."""
    net.generate(50000, seed=5)
    X, Y = net.samples(0, 50000)
    rng = np.random.default_rng(2)
    rows = rng.choice(50000, 300, replace=False)
    s_in = net.in_scale
    sig_all, D_all = net.decode(X[rows], Y[rows])
    nplastic = 0
    iu = np.triu_indices(6)
    for j, r in enumerate(rows):
        x = X[r].astype(np.float64) / s_in
        epsp_n, kap_n = x[6:12], x[12]
        eps = x[:6] + epsp_n
        sig, epsp, kap, sdev = oracle.mises_return(MP, eps, epsp_n, kap_n)
        D = oracle.mises_tangent(MP, kap_n, kap, sdev)
        nplastic += kap > kap_n
        assert np.abs(sig_all[j] - sig).max() <= 2e-4 * SIG0
        assert np.abs(D_all[j] - D[iu]).max() <= 1e-4 * E
        if kap == kap_n:                                   # elastic branch: zero departure
            assert np.abs(Y[r]).max() <= 1e-6
    assert 0.05 * len(rows) < nplastic < 0.95 * len(rows)      # both branches are sampled
    # path starts: virgin state
    assert np.any((X[:, 6:] == 0).all(axis=1))


def test_training_reduces_loss(ctx):
    from oofem_amd.hipfem import MisesSurrogate
    n = MisesSurrogate(ctx, E, NU, SIG0, H, seed=9)
    n.generate(400000, seed=6)
    X, Y = n.samples(0, 4096)
    l0, _ = n.gradients(X, Y)
    l1 = n.train(1500, batch=2048, lr=2e-3, lr_final=1e-4)
    l1b, _ = n.gradients(X, Y)
    assert l1b < 0.1 * l0, (l0, l1, l1b)
    n.close()


def test_bf16_inference_matches_fp32(net):
    """MFMA inference (BF16 operands, FP32 accumulation) against the FP32
    forward, including a ragged last tile."""
    X, _ = net.samples(0, 5000 + 17) if getattr(net, "nsamples", 0) >= 5017 else (None, None)
    if X is None:
        net.generate(6000, seed=8)
        X, _ = net.samples(0, 5017)
    y32 = net.forward(X)
    y16 = net.infer_bf16(X)
    err = np.linalg.norm(y16 - y32, axis=0) / np.maximum(np.linalg.norm(y32, axis=0), 1e-3 * np.sqrt(len(X)))
    assert err.max() <= 3e-2, err
    assert np.abs(y16).max() > 0


@pytest.fixture(scope="module")
def trained(ctx):
    """A network trained briefly on paths through the plain return mapping
    (a reduced training budget; the benchmark trains the full one)."""
    from oofem_amd.hipfem import MisesSurrogate
    n = MisesSurrogate(ctx, E, NU, SIG0, H, seed=21)
    n.generate(1_000_000, seed=22)
    n.loss = n.train(4000, batch=4096, lr=2e-3, lr_final=2e-5, seed=23)
    yield n
    n.close()


def test_surrogate_accuracy_limits(trained):
    """Why the 1e-4 acceptance target is out of this network's reach, each
    cause measured on plastic samples of its own training set (targets from
    the plain return mapping; the stress outputs are in units of sigma_0):
    the FP32 forward's stress error against the targets (what training
    leaves), and the change of the stress outputs by BF16 inference (BF16
    keeps 8 significant bits, so rounding the last hidden layer's 128
    activations alone moves an output by about 2^-9 of the norm of its
    terms).  Each median exceeds 1e-4 sigma_0 on its own."""
    X, Y = trained.samples(0, 200000)
    plastic = np.abs(Y[:, :6]).max(axis=1) > 1e-6
    assert plastic.sum() > 1000
    X, Y = X[plastic], Y[plastic]
    y32 = trained.forward(X)
    y16 = trained.infer_bf16(X)
    e_train = np.abs(y32[:, :6] - Y[:, :6]).max(axis=1)
    e_bf16 = np.abs(y16[:, :6] - y32[:, :6]).max(axis=1)
    rep = {name: dict(median=float(np.median(e)), p90=float(np.percentile(e, 90)))
           for name, e in (("fp32_vs_return_map", e_train), ("bf16_vs_fp32", e_bf16))}
    print(f"surrogate stress-output errors over sigma_0 on {len(X)} plastic samples: {rep}", flush=True)
    assert rep["fp32_vs_return_map"]["median"] > 1e-4
    assert rep["bf16_vs_fp32"]["median"] > 1e-4


def _plastic_block(ctx, amp):
    from oofem_amd import meshgen
    from oofem_amd.hipfem import BatchedEvaluator, HipCSRSparseMtrx
    d = meshgen.cantilever_hex(4, 4, 8, jitter=0.2, seed=5)
    d.mat = np.array([MP])
    d.emat[:] = 0
    ev = BatchedEvaluator(ctx, d)
    K = HipCSRSparseMtrx(ctx)
    ev.build_structure(K)
    rng = np.random.default_rng(4)
    u = amp * SIG0 / E * rng.standard_normal(d.neq)
    return d, ev, K, u


@pytest.mark.parametrize("tol", [1e-4, 3e-2])
def test_surrogate_update_is_checked_and_falls_back(ctx, trained, tol):
    """Internal forces with the surrogate: queued points either pass the
    return-mapping residual check (then their stress is within a few tol *
    sigma_y of the plain return mapping, the bound the check implies) or
    take the plain update; points outside the queue are bitwise the plain
    path's."""
    d, ev, K, u = _plastic_block(ctx, 0.25)
    up = np.zeros(d.npeq)
    f_ref, _, _ = ev.internal_forces(u, up, update_state=True)
    s_ref = ev.get_state()
    ev.reset_state()
    ev.set_surrogate(trained, band=0.05, tol=tol)
    ev.surrogate_stats(reset=True)
    f_sur, _, _ = ev.internal_forces(u, up, update_state=True)
    s_sur = ev.get_state()
    st = ev.surrogate_stats()
    assert st["queued"] > 0.2 * len(s_ref)
    assert st["accepted"] + st["rejected"] == st["queued"]
    if tol >= 1e-2:
        assert st["accepted"] > 0          # the bound below is then exercised on accepted points
    sy = SIG0 + H * s_ref[:, 6]
    err = np.abs(s_sur[:, 20:26] - s_ref[:, 20:26]).max(axis=1) / sy
    assert err.max() <= 4 * tol, err.max()
    same = err == 0
    assert same.sum() >= len(s_ref) - st["accepted"]        # everything not accepted is the plain result
    assert np.linalg.norm(f_sur - f_ref) <= 4 * tol * np.linalg.norm(f_ref)
    ev.set_surrogate(None)


def test_trained_network_gains_over_untrained(ctx, trained):
    """What training buys: at a loose acceptance tolerance (3e-2) the trained
    network serves a share of the plastic points that pass every residual and
    tangent check, the untrained one (same architecture, initial weights)
    none.  At the accuracy target (1e-4) neither is accepted: the fallback
    carries every point, which is the measured state of this surrogate."""
    from oofem_amd.hipfem import MisesSurrogate
    d, ev, K, u = _plastic_block(ctx, 0.25)
    up = np.zeros(d.npeq)
    fresh = MisesSurrogate(ctx, E, NU, SIG0, H, seed=21)
    share = {}
    for name, net in (("trained", trained), ("untrained", fresh)):
        ev.reset_state()
        ev.set_surrogate(net, band=0.05, tol=3e-2)
        ev.surrogate_stats(reset=True)
        ev.internal_forces(u, up, update_state=True)
        st = ev.surrogate_stats()
        share[name] = st["accepted"] / max(st["queued"], 1)
    ev.set_surrogate(None)
    fresh.close()
    assert share["trained"] >= 0.03 and share["untrained"] == 0.0, share


def test_newton_with_surrogate_converges_to_plain_solution(ctx, trained):
    """StaticStructural + NRSolver steps with the surrogate inside the
    material update: every step converges (rtolf 1e-8), and displacements
    stay within the acceptance tolerance's reach of the plain run."""
    from oofem_amd import meshgen
    from oofem_amd.problem import StructuralProblem
    from tests import nlcases
    d = meshgen.cantilever_hex(3, 3, 8, jitter=0.1, seed=2)
    d.mat = np.array([MP])
    d.emat[:] = 0
    top = np.arange(d.nnode - 16, d.nnode)
    d.bc_mask[top, 2] = True
    d.bc_value[top, 2] = 4 * SIG0 / E * d.coords[:, 2].max()
    d2 = meshgen.Deck(d.name, d.etype, d.coords, d.conn, d.emat, d.mat, d.bc_mask, d.bc_value)
    res = {}
    for use in (False, True):
        P = StructuralProblem(d2, lstol=1e-12, rtolf=1e-8, maxiter=40, manrmsteps=1)
        P.set_load(np.zeros(d2.neq), np.zeros(d2.npeq))
        if use:
            from oofem_amd.hipfem import MisesSurrogate
            net = MisesSurrogate(P.context(), E, NU, SIG0, H, seed=21)
            net.set_params(trained.params())
            P.set_surrogate(net, band=0.05, tol=3e-2)
        for k in range(1, 6):
            P.set_prescribed(d2.prescribed_values(k / 5))
            rep = P.nonlinear_step()
            assert rep["status"] == 1, (use, k, rep)
            P.update()
        res[use] = P.solution()
        if use:
            res["stats"] = P.surrogate_stats()
            P.set_surrogate(None)
            net.close()
        P.close()
    assert res["stats"]["queued"] > 0
    assert np.linalg.norm(res[True] - res[False]) <= 5e-2 * np.linalg.norm(res[False])


def test_element_operator_keeps_the_assembled_surrogate_tangent(ctx, trained):
    """The FP32 element operator applies the tangent the FP64 matrix was
    assembled with, also at surrogate points: after a later state update
    (new surrogate tangents at those points) it still equals K to FP32
    accuracy."""
    d, ev, K, u = _plastic_block(ctx, 0.25)
    up = np.zeros(d.npeq)
    ev.set_surrogate(trained, band=0.05, tol=3e-2)
    ev.surrogate_stats(reset=True)
    ev.internal_forces(u, up, update_state=True)
    assert ev.surrogate_stats()["accepted"] > 0
    ev.assemble_stiffness(K, tangent=True)
    ev.set_element_operator(2)
    assert K.is_matrix_free()
    ev.internal_forces(1.3 * u, up, update_state=True)      # the live surrogate tangents change
    assert K.is_matrix_free()
    rng = np.random.default_rng(7)
    for _ in range(2):
        x = rng.standard_normal(d.neq)
        y64 = K.times(x)
        y32 = K.times_matrix_free(x)
        assert np.linalg.norm(y32 - y64) <= 2e-5 * np.linalg.norm(y64)
    ev.set_surrogate(None)


def test_surrogate_against_plain_return_map_on_the_plasticity_block(ctx, trained):
    """What the trained network gains over the plain return mapping on the
    64x64x256 LSpace plasticity block (1,048,576 elements, 8,388,608 Gauss
    points).  Committed state: the uniaxial field of load step 5 of the
    20-step ramp of the block's end displacement (1.25x the yield strain);
    update: the field of step 6, so every plastic point is queued.  Measured, on device vectors: the
    material update (internal forces with state update) with the plain return
    mapping, and with the surrogate at the stated tolerances (3e-2, and 1e-4,
    the accuracy target): queued and accepted points and the update time.
    Asserted: the accepted points meet the bound the residual check implies
    (4 tol sigma_y), every other point is bitwise the plain update, and at
    3e-2 the network serves a non-zero share of the queue.  The surrogate
    does not make the update cheaper: the closed-form radial return costs
    about 100 flops per point against the network's 75.8 kflop."""
    import time

    from oofem_amd import meshgen
    from oofem_amd.hipfem import BatchedEvaluator, DeviceVector, HipCSRSparseMtrx
    d = meshgen.config3()
    ev = BatchedEvaluator(ctx, d)
    K = HipCSRSparseMtrx(ctx)
    ev.build_structure(K)
    zmax = d.coords[:, 2].max()
    fe = d.eqn[:, 2] > 0

    def field(frac):
        up = d.prescribed_values(frac)
        u = np.zeros(d.neq)
        u[d.eqn[fe, 2] - 1] = d.coords[fe, 2] / zmax * frac * d.meta["u_end"]
        return u, up

    u0, up0 = field(5.0 / 20.0)
    ev.internal_forces(u0, up0, update_state=True)
    ev.commit()
    u1, up1 = field(6.0 / 20.0)
    du = DeviceVector(ctx, d.neq, u1)
    dup = DeviceVector(ctx, d.npeq, up1)
    df = DeviceVector(ctx, d.neq)
    dfp = DeviceVector(ctx, d.npeq)

    def update(net, tol):
        ev.set_surrogate(net, band=0.05, tol=tol)
        best = None
        for _ in range(3):
            ev.restart()
            if net is not None:
                ev.surrogate_stats(reset=True)
            ctx.synchronize()
            t = time.perf_counter()
            ev.internal_forces_dev(du.addr, dup.addr, df.addr, dfp.addr, update_state=True)
            ctx.synchronize()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        st = ev.surrogate_stats() if net is not None else None
        s = ev.get_state()
        return best, st, s[:, 20:26].copy(), s[:, 6].copy()

    t_plain, _, s_plain, kappa_n = update(None, 0.0)
    sy = SIG0 + H * kappa_n
    report = {"plain_update_ms": round(t_plain * 1e3, 2)}
    for tol in (3e-2, 1e-4):
        dt, st, s_sur, _ = update(trained, tol)
        err = np.abs(s_sur - s_plain).max(axis=1) / sy
        assert st["accepted"] + st["rejected"] == st["queued"]
        assert st["queued"] > 0.5 * len(s_plain)
        assert err.max() <= 4 * tol, (tol, err.max())
        assert (err == 0).sum() >= len(s_plain) - st["accepted"]
        report[f"tol {tol:g}"] = dict(queued=st["queued"], accepted=st["accepted"],
                                      accepted_share=round(st["accepted"] / st["queued"], 4),
                                      update_ms=round(dt * 1e3, 2))
    ev.set_surrogate(None)
    print(f"surrogate on the plasticity block: {report}", flush=True)
    assert report["tol 0.03"]["accepted"] > 0
    K.close()
    ev.close()
