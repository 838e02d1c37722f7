"""Text decks of the generated blocks, in the reference's input format, for
the tests that drive run_deck on problems no reference deck holds."""
import numpy as np


def write_block_deck(deck, path, engng, metastep=None, checks=()):
    """An LSpace / MisesMat block with clamped bottom (z = 0) and a top face
    (z = max) whose u_z follows PiecewiseLinFunction 2 (0 at t = 0, the
    deck's prescribed value at t = 1).  engng: the engineering-model record
    (with nmodules 1: the errorcheck module);
    metastep: an optional metastep record; checks: errorcheck lines."""
    c = deck.coords
    z = c[:, 2]
    bot = np.nonzero(np.isclose(z, 0.0))[0]
    top = np.nonzero(np.isclose(z, z.max()))[0]
    uz = float(np.unique(deck.bc_value[top, 2])[0])
    m = deck.mat[0]
    lines = ["block.out", "generated block", engng]
    if metastep:
        lines.append(metastep)
    lines += ["errorcheck", "domain 3d", "OutputManager tstep_all dofman_all element_all",
              f"ndofman {deck.nnode} nelem {deck.nelem} ncrosssect 1 nmat 1 nbc 2 nic 0 nltf 2 nset 3"]
    for i in range(deck.nnode):
        lines.append(f"node {i + 1} coords 3 {float(c[i, 0])!r} {float(c[i, 1])!r} {float(c[i, 2])!r}")
    for e in range(deck.nelem):
        lines.append(f"lspace {e + 1} nodes 8 " + " ".join(str(int(n) + 1) for n in deck.conn[e]))
    lines.append("SimpleCS 1 material 1 set 3")
    lines.append(f"MisesMat 1 d 0.0 E {float(m[1])!r} n {float(m[2])!r} sig0 {float(m[3])!r} H {float(m[4])!r} omega_crit 0 a 0 tAlpha 0.0")
    lines.append("BoundaryCondition 1 loadTimeFunction 1 dofs 3 1 2 3 values 3 0. 0. 0. set 1")
    lines.append(f"BoundaryCondition 2 loadTimeFunction 2 dofs 1 3 values 1 {uz!r} set 2")
    lines.append("ConstantFunction 1 f(t) 1.0")
    lines.append("PiecewiseLinFunction 2 t 2 0. 1. f(t) 2 0. 1.")
    lines.append(f"Set 1 nodes {len(bot)} " + " ".join(str(int(n) + 1) for n in bot))
    lines.append(f"Set 2 nodes {len(top)} " + " ".join(str(int(n) + 1) for n in top))
    lines.append(f"Set 3 elementranges {{(1 {deck.nelem})}}")
    if checks:
        lines += ["#%BEGIN_CHECK% tolerance 1.e-8", *checks, "#%END_CHECK%"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
