"""Writes tests/golden/iml_cg_history.json: the residual histories of the
reference's own CG template (iml/cg.h, compiled unchanged into
oracle/_ref/libref_iml_cg.so by oracle/Makefile) with the diagonal
preconditioner, SymCompCol product order (smtype 4) and lstol 1e-10, on two
meshes the mesh generator rebuilds bit for bit:

  config0: the 10x10x40 LSpace cantilever (14,520 equations), right-hand side
           the 1 MPa end traction;
  hex8x8x24: an 8x8x24 LSpace cantilever with node jitter 0.2 h and
           log-normal E (sigma 0.5, seed 11), same traction.

Each case stores the iteration count, cg.h's achieved tolerance and
||r_k||/||b|| for k = 0..iters.  Run in a tree where the reference is
present: python tests/golden/make_iml_cg_golden.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oofem_amd import meshgen  # noqa: E402
from oracle import oracle  # noqa: E402

CASES = {
    "config0": lambda: meshgen.config0(),
    "hex8x8x24": lambda: meshgen.cantilever_hex(8, 8, 24, jitter=0.2, e_sigma=0.5, seed=11, name="hex8x8x24"),
}


def main():
    if not oracle.have_ref_cg():
        oracle.build()
    out = {}
    for name, make in CASES.items():
        d = make()
        rp, ci, val = oracle.csr(d)
        f, _ = oracle.surface_load(d)
        _, flag, it, tol, hist = oracle.ref_cg(rp, ci, val, f, tol=1e-10, storage=0, precond=1)
        assert flag == 0
        out[name] = {"neq": int(d.neq), "nnz": int(len(val)), "lstol": 1e-10, "iters": int(it),
                     "achieved": float(tol), "hist": [float(v) for v in hist]}
        print(name, d.neq, it, tol)
    with open(os.path.join(ROOT, "tests", "golden", "iml_cg_history.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
