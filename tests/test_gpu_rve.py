"""GPU tests of the batched RVE solves (hfem_batch): homogenised stiffness
of a mixed LSpace / QSpace batch against the oracle's six FP64 solves per
RVE, and the autotuner's choice."""
import numpy as np
import pytest

from oofem_amd import meshgen
from oracle import oracle

pytestmark = pytest.mark.gpu

SPECS = [(3, meshgen.LSPACE, 0.3, 1), (4, meshgen.QSPACE, 0.2, 2), (6, meshgen.LSPACE, 0.15, 3),
         (3, meshgen.QSPACE, 0.35, 4), (5, meshgen.LSPACE, 0.25, 5), (2, meshgen.QSPACE, 0.1, 6)]


def _oracle_C(spec):
    d = meshgen.rve(*spec)
    C = np.zeros((6, 6))
    for k in range(6):
        e = np.zeros(6)
        e[k] = 1.0
        r = oracle.linear_static(d, tol=1e-13, u_presc=meshgen.affine_values(d, e))
        C[:, k] = meshgen.homogenized_stress(d, r["reactions"])
    return C


# stored entries per block range from about 400 (n = 2, 3) to 60,000 (QSpace
# n = 4): 0 sends every block to the bulk phase, 5000 splits the batch,
# 10**12 keeps every block in the persistent kernel
SPLITS = {"bulk": 0, "bulk_csr": 0, "bulk_wide": 0, "mixed": 5000, "persistent": 10**12}


@pytest.mark.parametrize("mode", sorted(SPLITS))
def test_batch_homogenized_stiffness_matches_oracle(ctx, mode):
    from oofem_amd.rve import RVEBatch
    B = RVEBatch(ctx, SPECS, nnz_budget=2e5)        # small budget: several chunks per element type
    assert len(B.chunks) >= 3
    form = {"bulk_csr": 0, "bulk_wide": 2}.get(mode, 1)
    B.set_options(tol=1e-10, split=SPLITS[mode], bulk_form=form)
    # every RVE pattern is node-blocked (boundary nodes are prescribed whole)
    assert B.bulk_forms() == [form] * len(B.chunks)
    nbulk = sum(nb for nb, _ in B.partition())
    expect = {"bulk": lambda v: v == len(SPECS), "bulk_csr": lambda v: v == len(SPECS),
              "bulk_wide": lambda v: v == len(SPECS),
              "mixed": lambda v: 0 < v < len(SPECS),
              "persistent": lambda v: v == 0}[mode]
    assert expect(nbulk), (mode, nbulk, B.partition())
    out = B.run()
    assert (out["resid"] > 0).all() and (out["resid"] <= 1e-10).all(), out["resid"]
    for i, spec in enumerate(SPECS):
        ref = _oracle_C(spec)
        err = np.abs(out["C"][i] - ref).max() / np.abs(ref).max()
        assert err <= 1e-7, (spec, err)
    B.close()


def test_batch_autotune_picks_a_valid_shape(ctx):
    from oofem_amd.rve import RVEBatch
    B = RVEBatch(ctx, SPECS[:4])
    B.set_options(tol=1e-8, split=5000)
    tuned = B.autotune(sample=2)
    for t in tuned:
        if t["bulk_blocks"] < t["blocks"]:
            assert 0 <= t["best"] < len(t["ms"]) and t["ms"][t["best"]] == min(t["ms"]), t
        if t["bulk_blocks"] > 0:
            assert t["bulk_ms"][t["bulk_lanes_per_row"].bit_length() - 3] == min(t["bulk_ms"]), t
    out = B.run()
    assert (out["resid"] > 0).all()
    B.close()


def test_batch_bulk_maxit_reports_unconverged(ctx):
    """A cap below the iterations needed ends every block of both phases
    with a negative residual and exactly the capped iteration count."""
    from oofem_amd.rve import RVEBatch
    B = RVEBatch(ctx, SPECS[:3])
    B.set_options(tol=1e-12, maxit=5, split=5000)
    out = B.run()
    assert (out["resid"] < 0).all() and (out["iters"] == 5).all(), out
    B.close()
