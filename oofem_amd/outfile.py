"""The reference's ``.out`` result file for the records this path produces,
and a reader for it (the role of the reference's tools/extractor.py, which
regression tests use to compare a run's output file with expected values).

Layout per solution step (StructuralEngngModel::printOutputAt,
structengngmodel.C:265-274, over EngngModel::printOutputAt engngm.C:664-684):
- the step header ``Output for time ...`` and ``Output for domain``;
- ``DofManager output``: per node ``Node <label> (<number>):`` and one
  ``dof <id> d <value>`` line per dof, total displacement, prescribed or
  free (DofManager::printOutputAt dofmanager.C:507-515,
  Dof::printSingleOutputAt dof.C:76-82);
- ``Element output``: per element ``element <label> (<number>) :`` and per
  Gauss point ``GP <rule>.<gp> :`` with the strains and stresses in full
  Voigt order (StructuralMaterialStatus::printOutputAt structuralms.C), for
  MisesMat also the plastic strain and ``status { damage ..., kappa ...}``
  (MisesMatStatus::printOutputAt misesmat.C:683-703);
- the reaction table, one line per prescribed dof sorted by node and dof
  (StructuralEngngModel::printReactionForces structengngmodel.C:103-137);
- ``User time consumed by solution step`` (engngm.C:651-654).
The printf formats are the reference's.
"""
from __future__ import annotations

import re
import time

import numpy as np

from .meshgen import MAT_MISES, NGP

_RULE = "=============================================================="


def write_header(f, prob):
    """Start of the file (EngngModel::instanciateYourself prints the
    program banner, the start time and the deck's description line)."""
    f.write(f"oofem_amd output for {prob.output_file}\n")
    f.write(f"\nStarting analysis on: {time.ctime()}\n")
    f.write(f"{prob.description}\n")


def _bc_ids(prob):
    ids = {}
    for bc in prob.bcs:
        for dof in bc["dofs"]:
            for n in bc["nodes"]:
                ids[(int(n), int(dof))] = bc.get("number") or 0
    return ids


def write_step(f, prob, step, t, res, seconds=0.0):
    """One solution step.  res: the quantities of a solved step,
    dict(u_node (nnode x 3), reactions {(node index, dof): value},
    state (nelem x ngp x 38 Gauss-point records, the oracle's layout))."""
    d = prob.deck
    f.write(f"\n{_RULE}")
    f.write(f"\nOutput for time {t:.8e} ")
    f.write(f"\n{_RULE}\n")
    f.write(f"Output for domain {1:3d}\n")
    if "load_level" in res:
        # NonLinearStatic::printOutputAt nlinearstatic.C:674-688
        f.write(f"Reached load level : {res['load_level']:20.6f} in {int(res.get('iterations', 0))} iterations\n")
    f.write("\n\nDofManager output:\n------------------\n")
    u = res["u_node"]
    for i in range(d.nnode):
        f.write(f"{'Node':<8s}{int(prob.node_numbers[i]):8d} ({i + 1:8d}):\n")
        for k in range(3):
            f.write(f"  dof {k + 1:<3d} d {u[i, k]: .8e}\n")
    f.write("\n\nElement output:\n---------------\n")
    st = np.asarray(res["state"]).reshape(d.nelem, NGP[d.etype], -1)
    for e in range(d.nelem):
        f.write(f"element {int(prob.elem_numbers[e])} ({e + 1:8d}) :\n")
        mises = int(d.mat[d.emat[e], 0]) == MAT_MISES
        for g in range(st.shape[1]):
            r = st[e, g]
            f.write(f"  GP {1:2d}.{g + 1:<2d} :")
            f.write("  strains " + "".join(f" {v:+.4e}" for v in r[26:32]))
            f.write("\n              stresses" + "".join(f" {v:+.4e}" for v in r[20:26]))
            f.write("\n")
            if mises:
                f.write("              plastic  " + "".join(f"{v:.4e} " for v in r[0:6]) + "\n")
                f.write(f"status {{ damage {0.0:.4e}, kappa  {r[6]:.4e}}}\n")
                f.write("}\n")
    f.write("\n\n\tR E A C T I O N S  O U T P U T:\n\t_______________________________\n\n\n")
    ids = _bc_ids(prob)
    for (n, k), v in sorted(res["reactions"].items()):
        f.write(f"\tNode {int(prob.node_numbers[n]):8d} iDof {k:2d} reaction {v: .4e}    [bc-id: {ids.get((n, k), 0)}]\n")
    f.write(f"\nUser time consumed by solution step {step}: {seconds:.3f} [s]\n\n")


_NODE = re.compile(r"^Node\s+(\d+)\s+\(\s*(\d+)\):")
_DOF = re.compile(r"^\s+dof\s+(\d+)\s+d\s+(\S+)")
_ELEM = re.compile(r"^element\s+(\d+)\s+\(\s*(\d+)\)\s+:")
_GP = re.compile(r"^\s+GP\s+(\d+)\.(\d+)\s+:\s+strains\s+(.*)$")
_STRESS = re.compile(r"^\s+stresses\s+(.*)$")
_REAC = re.compile(r"^\s*Node\s+(\d+)\s+iDof\s+(\d+)\s+reaction\s+(\S+)")


def read_out(text):
    """Steps of a .out file, in order: list of dict(time, u {(node label, dof):
    value}, reactions {(node label, dof): value}, strains / stresses
    {(element label, gp): 6 values})."""
    steps = []
    cur = node = elem = gp = None
    for line in text.splitlines():
        if line.startswith("Output for time"):
            cur = dict(time=float(line.split()[3]), u={}, reactions={}, strains={}, stresses={})
            steps.append(cur)
            continue
        if line.startswith("Reached load level") and cur is not None:
            cur["load_level"] = float(line.split(":")[1].split()[0])
            continue
        if cur is None:
            continue
        m = _NODE.match(line)
        if m:
            node = int(m.group(1))
            continue
        m = _DOF.match(line)
        if m and node is not None:
            cur["u"][(node, int(m.group(1)))] = float(m.group(2))
            continue
        m = _ELEM.match(line)
        if m:
            elem, node = int(m.group(1)), None
            continue
        m = _GP.match(line)
        if m and elem is not None:
            gp = int(m.group(2))
            cur["strains"][(elem, gp)] = [float(v) for v in m.group(3).split()]
            continue
        m = _STRESS.match(line)
        if m and elem is not None and gp is not None:
            cur["stresses"][(elem, gp)] = [float(v) for v in m.group(1).split()]
            continue
        m = _REAC.match(line)
        if m:
            cur["reactions"][(int(m.group(1)), int(m.group(2)))] = float(m.group(3))
    return steps


def check_out(prob, steps):
    """The deck's errorcheck records evaluated on a .out file's values (what
    the reference's extractor-based regression tests do): list of (check,
    value read, within the deck's tolerance).  Values are printed with 9
    (displacements) or 5 (reactions, strains, stresses) significant digits,
    so a window tighter than that precision can fail on the file although
    it holds for the computed values."""
    out = []
    for c in prob.checks:
        if c.tstep < 1 or c.tstep > len(steps):
            continue
        s = steps[c.tstep - 1]
        if c.kind == "node":
            v = s["u"][(c.number, c.dof)]
        elif c.kind == "reaction":
            v = s["reactions"].get((c.number, c.dof), 0.0)
        elif c.kind == "loadlevel":
            v = s.get("load_level", float("nan"))
        else:
            arr = s["stresses"] if c.keyword == 1 else s["strains"]
            v = arr[(c.number, c.gp)][c.component - 1]
        out.append((c, float(v), abs(v - c.value) <= prob.tolerance))
    return out
