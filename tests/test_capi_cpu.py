"""The C-ABI library loads on a host without a GPU and exports every entry
point include/hipfem.h declares (no compute calls here)."""
import ctypes
import os
import re

import numpy as np

from oofem_amd import build, meshgen
from tests.conftest import ROOT


def _declared(header="hipfem.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b((?:hipfem|hipcsr|hippcg|hfem_model|hfem_problem)_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_host_library_exports_every_symbol():
    assert os.path.exists(build.HOST_LIB), "build libhipfem_host.so first (__graft_entry__.build())"
    ctypes.CDLL(build.LIB, mode=ctypes.RTLD_GLOBAL)
    lib = ctypes.CDLL(build.HOST_LIB)
    names = _declared("hipfem_host.h")
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) >= 45
    for must in ("hipcsr_build_structure", "hipcsr_assemble", "hipcsr_times", "hippcg_solve",
                 "hfem_model_assemble_stiffness", "hfem_model_internal_forces"):
        assert must in names


def test_library_exports_every_symbol():
    assert os.path.exists(build.LIB), "build libhipfem.so first (__graft_entry__.build())"
    lib = ctypes.CDLL(build.LIB)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string_without_gpu():
    lib = ctypes.CDLL(build.LIB)
    lib.hipfem_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.hipfem_version()


def test_equation_numbering_follows_node_order():
    bc = np.zeros((4, 3), bool)
    bc[1, 0] = True
    bc[3] = True
    eqn, neq, npeq = meshgen.number_equations(bc)
    assert neq == 8 and npeq == 4
    assert eqn.tolist() == [[1, 2, 3], [-1, 4, 5], [6, 7, 8], [-2, -3, -4]]


def test_tet_mesh_positive_volumes():
    coords, tets = meshgen.kuhn_tets(4, 3, 5, 1.0, 0.8, 1.3, jitter=0.15, seed=3)
    X = coords[tets]
    det = np.einsum("ij,ij->i", X[:, 3] - X[:, 0], np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]))
    assert (det > 0).all()
    assert abs(det.sum() / 6.0 - 1.0 * 0.8 * 1.3) < 1e-12
