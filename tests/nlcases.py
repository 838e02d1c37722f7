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
