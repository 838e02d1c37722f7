"""Nonlinear (MisesMat) test problems shared by the oracle and GPU tests."""
import numpy as np

from oofem_amd import meshgen

E, NU, SIG0, H = 210e3, 0.3, 250.0, 2e3
MISES = np.array([[meshgen.MAT_MISES, E, NU, SIG0, H, 0, 0, 0]])


def uniaxial_block(etype=meshgen.LSPACE, n=2, L=1.0):
    """Cube [0,L]^3: u_z = 0 on z = 0, u_z prescribed on z = L, lateral
    motion constrained only against rigid-body modes, so the exact solution
    is a homogeneous uniaxial stress state."""
    if etype == meshgen.LSPACE:
        d0 = meshgen.cantilever_hex(n, n, n, lx=L, ly=L, lz=L)
        coords, conn = d0.coords, d0.conn
    else:
        coords, conn = meshgen.kuhn_tets(n, n, n, L, L, L)
    nn = coords.shape[0]
    bc = np.zeros((nn, 3), bool)
    bot = np.isclose(coords[:, 2], 0.0)
    top = np.isclose(coords[:, 2], L)
    bc[bot, 2] = True
    bc[top, 2] = True
    o = np.nonzero(bot & np.isclose(coords[:, 0], 0) & np.isclose(coords[:, 1], 0))[0][0]
    xnode = np.nonzero(bot & np.isclose(coords[:, 0], L) & np.isclose(coords[:, 1], 0))[0][0]
    bc[o, 0] = bc[o, 1] = True
    bc[xnode, 1] = True
    val = np.zeros((nn, 3))
    val[top, 2] = 1.0               # scaled by the load factor of the step
    d = meshgen.Deck("uniaxial", etype, coords, conn, np.zeros(len(conn), np.int32), MISES, bc, val)
    return d, top


def uniaxial_stress(eps):
    """Closed-form uniaxial stress of J2 plasticity with linear hardening
    (sigma_y = sig0 + H kappa, monotonic loading)."""
    if E * eps <= SIG0:
        return E * eps
    return E * (SIG0 + H * eps) / (E + H)


def top_reaction_z(deck, top, reactions):
    m = deck.eqn[top, 2]
    return reactions[-m - 1].sum()


def tet_bar(nx=3, ny=3, nz=5, jitter=0.12, seed=9, u_end=None):
    """Jittered Kuhn-tet bar, clamped at z = 0, u_z prescribed at the far end
    (the shape of the tetrahedral plasticity workload, at test size)."""
    return meshgen.config2(nx=nx, ny=ny, nz=nz, seed=seed, jitter=jitter, u_end=u_end)


def traction_block(etype=meshgen.LSPACE, n=2, L=1.0, material=None):
    """The uniaxial block with its top face free and loaded instead by the
    consistent nodal forces of a unit traction in +z (total force L^2):
    the reference load of an arc-length analysis.  Returns (deck, top
    nodes, R free, R prescribed)."""
    d0, top = uniaxial_block(etype, n, L)
    bc = d0.bc_mask.copy()
    bc[top, 2] = False
    coords, conn = d0.coords, d0.conn
    nn = coords.shape[0]
    fz = np.zeros((nn, 3))
    if etype == meshgen.LSPACE:
        # consistent nodal forces of a unit traction on n x n bilinear faces of
        # area h^2: h^2/4 per face and node, so corner, edge and interior
        # nodes carry 1/4, 1/2 and 1 times h^2
        h = L / n
        w = np.zeros(nn)
        for i in np.nonzero(top)[0]:
            cx = np.isclose(coords[i, 0], 0) or np.isclose(coords[i, 0], L)
            cy = np.isclose(coords[i, 1], 0) or np.isclose(coords[i, 1], L)
            w[i] = h * h * (0.25 if (cx and cy) else 0.5 if (cx or cy) else 1.0)
        fz[:, 2] = w
    else:
        raise NotImplementedError("traction_block: LSpace only")
    mat = MISES if material is None else material
    d = meshgen.Deck("traction", etype, coords, conn, np.zeros(len(conn), np.int32), mat, bc, np.zeros((nn, 3)))
    R = np.zeros(d.neq)
    Rp = np.zeros(d.npeq)
    m = d.eqn > 0
    R[d.eqn[m] - 1] = fz[m]
    m = d.eqn < 0
    Rp[-d.eqn[m] - 1] = fz[m]
    return d, np.nonzero(top)[0], R, Rp


def uniaxial_strain(sig):
    """Inverse of uniaxial_stress: eps = sigma/E + (sigma - sig0)/H past yield."""
    if sig <= SIG0:
        return sig / E
    return sig / E + (sig - SIG0) / H
