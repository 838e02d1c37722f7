"""Meshes of the reference-CG residual-history fixture
(tests/golden/iml_cg_history.json, written by
tests/golden/make_iml_cg_golden.py from the reference's own iml/cg.h)."""
import json
import os

from oofem_amd import meshgen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CASES = {
    "config0": lambda: meshgen.config0(),
    "hex8x8x24": lambda: meshgen.cantilever_hex(8, 8, 24, jitter=0.2, e_sigma=0.5, seed=11, name="hex8x8x24"),
}


def fixture():
    with open(os.path.join(GOLDEN, "iml_cg_history.json")) as fh:
        return json.load(fh)
