"""Synthetic structured meshes ("decks") for the assembly-and-solve path.

Each generator returns a :class:`Deck`: node coordinates, element
connectivity in the reference's local node order, a per-element material
table, Dirichlet constraints, equation numbering and the external load
vector.  Equation numbering follows ``EngngModel::forceEquationNumbering``
(engngm.C:469-531) together with ``MasterDof::askNewEquationNumber``
(masterdof.C:105-126): nodes in order, dofs u, v, w; a free dof takes the
next free equation number, a constrained one the next prescribed number.

LSpace local node order (fei3dhexalin.C:46-64): nodes 1-4 on zeta=+1 at
(xi, eta) = (-,-), (-,+), (+,+), (+,-); nodes 5-8 the same on zeta=-1.  With
xi, eta, zeta aligned to x, y, z, local surface 1 (nodes 1 4 3 2,
fei3dhexalin.C:594-613) is the element's +z face.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

LSPACE, LTRSPACE, QSPACE = 1, 2, 3
NEN = {LSPACE: 8, LTRSPACE: 4, QSPACE: 20}
NGP = {LSPACE: 8, LTRSPACE: 1, QSPACE: 27}

MAT_ISOLE, MAT_MISES = 0, 1
MATP = 8  # material row: [model, E, nu, sig0, H, 0, 0, 0]


@dataclass
class Deck:
    name: str
    etype: int
    coords: np.ndarray            # (nnode, 3) float64
    conn: np.ndarray              # (nelem, nen) int32, 0-based
    emat: np.ndarray              # (nelem,) int32
    mat: np.ndarray               # (nmat, MATP) float64
    bc_mask: np.ndarray           # (nnode, 3) bool, True = prescribed
    bc_value: np.ndarray          # (nnode, 3) float64 prescribed values (total, at load factor 1)
    faces: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))   # (nface, 2): element, surface (1-based)
    traction: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))          # (nface, 3)
    eqn: np.ndarray = None        # (nnode, 3) int32: >0 free eq, <0 -prescribed eq
    neq: int = 0
    npeq: int = 0
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        self.coords = np.ascontiguousarray(self.coords, dtype=np.float64)
        self.conn = np.ascontiguousarray(self.conn, dtype=np.int32)
        self.emat = np.ascontiguousarray(self.emat, dtype=np.int32)
        self.mat = np.ascontiguousarray(self.mat, dtype=np.float64)
        if self.eqn is None:
            self.eqn, self.neq, self.npeq = number_equations(self.bc_mask)

    @property
    def nnode(self):
        return self.coords.shape[0]

    @property
    def nelem(self):
        return self.conn.shape[0]

    @property
    def nen(self):
        return NEN[self.etype]

    def prescribed_values(self, factor=1.0):
        """Prescribed displacement per prescribed equation (length npeq)."""
        out = np.zeros(self.npeq)
        m = self.eqn < 0
        out[-self.eqn[m] - 1] = self.bc_value[m] * factor
        return out

    def external_load(self, factor=1.0):
        """(f_free, f_presc): traction loads assembled in both numberings."""
        return surface_load_vector(self, factor)


def number_equations(bc_mask: np.ndarray):
    """Node order, dof order u,v,w; free -> +k, prescribed -> -k (1-based)."""
    flat = np.asarray(bc_mask, dtype=bool).reshape(-1)
    free = ~flat
    eq = np.zeros(flat.shape[0], dtype=np.int64)
    eq[free] = np.arange(1, int(free.sum()) + 1)
    eq[flat] = -np.arange(1, int(flat.sum()) + 1)
    return eq.reshape(-1, 3).astype(np.int32), int(free.sum()), int(flat.sum())


def _grid_nodes(nx, ny, nz, lx, ly, lz):
    x = np.linspace(0.0, lx, nx + 1)
    y = np.linspace(0.0, ly, ny + 1)
    z = np.linspace(0.0, lz, nz + 1)
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    # values as a text deck holds them ("0.3", not 0.30000000000000004)
    return np.round(np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1), 12)


def _hex_conn(nx, ny, nz):
    """LSpace connectivity, element id = ex + nx*(ey + ny*ez)."""
    ez, ey, ex = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    sx, sy = nx + 1, (nx + 1) * (ny + 1)

    def nid(i, j, k):
        return i + sx * j + sy * k

    x0, x1, y0, y1, z0, z1 = ex, ex + 1, ey, ey + 1, ez, ez + 1
    conn = np.stack([
        nid(x0, y0, z1), nid(x0, y1, z1), nid(x1, y1, z1), nid(x1, y0, z1),
        nid(x0, y0, z0), nid(x0, y1, z0), nid(x1, y1, z0), nid(x1, y0, z0),
    ], axis=1)
    return conn.astype(np.int32)


def _jitter(coords, nx, ny, nz, h, amp, seed):
    """Seeded uniform jitter of +-amp*h per component.  Boundary nodes move
    only tangentially, so the clamped and loaded planes stay planar."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(-amp, amp, size=coords.shape) * np.asarray(h)[None, :]
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    d[(i == 0) | (i == nx), 0] = 0.0
    d[(j == 0) | (j == ny), 1] = 0.0
    d[(k == 0) | (k == nz), 2] = 0.0
    return coords + d


def cantilever_hex(nx, ny, nz, lx=1.0, ly=1.0, lz=4.0, E=30e3, nu=0.2, traction=(0.0, -1.0, 0.0),
                   jitter=0.0, e_sigma=0.0, seed=0, name="cantilever_hex"):
    """Block [0,lx]x[0,ly]x[0,lz] of LSpace hexes, face z=0 clamped, uniform
    traction on face z=lz (element surface 1)."""
    coords = _grid_nodes(nx, ny, nz, lx, ly, lz)
    conn = _hex_conn(nx, ny, nz)
    h = (lx / nx, ly / ny, lz / nz)
    if jitter > 0:
        coords = _jitter(coords, nx, ny, nz, h, jitter, seed)
    nelem = conn.shape[0]
    if e_sigma > 0:
        rng = np.random.default_rng(seed + 1)
        Ee = E * np.exp(e_sigma * rng.standard_normal(nelem))
        mat = np.zeros((nelem, MATP))
        mat[:, 0] = MAT_ISOLE
        mat[:, 1] = Ee
        mat[:, 2] = nu
        emat = np.arange(nelem, dtype=np.int32)
    else:
        mat = np.array([[MAT_ISOLE, E, nu, 0, 0, 0, 0, 0]], dtype=np.float64)
        emat = np.zeros(nelem, dtype=np.int32)
    bc = np.zeros((coords.shape[0], 3), dtype=bool)
    bc[: (nx + 1) * (ny + 1)] = True  # z = 0 layer comes first in node order
    top = np.arange(nx * ny * (nz - 1), nx * ny * nz, dtype=np.int32)
    faces = np.stack([top, np.ones_like(top)], axis=1)
    trac = np.tile(np.asarray(traction, dtype=np.float64), (top.shape[0], 1))
    return Deck(name, LSPACE, coords, conn, emat, mat, bc, np.zeros_like(coords, dtype=np.float64),
                faces, trac, meta=dict(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, seed=seed,
                                        jitter=jitter, e_sigma=e_sigma))


def config0():
    """10x10x40 LSpace cantilever 1x1x4 m, IsoLE E=30e3 nu=0.2, clamp z=0,
    traction 1 MPa in -y on z=4."""
    return cantilever_hex(10, 10, 40, name="cantilever_10x10x40")


def config1(nx=64, ny=64, nz=256, seed=20261203):
    """64x64x256 jittered LSpace cantilever (+-0.2 h), log-normal E (median
    30e3, sigma 0.5), nu=0.2; same clamp and traction as config0."""
    return cantilever_hex(nx, ny, nz, jitter=0.2, e_sigma=0.5, seed=seed,
                          name=f"cantilever_{nx}x{ny}x{nz}_jitter_lognormalE")


def config3(nx=64, ny=64, nz=256, seed=20261205, u_end=None):
    """The plasticity problem of config2 on the config1 LSpace mesh: block
    1 x 1 x 4 of nx x ny x nz hexes (jitter +-0.2 h), MisesMat E=210e3
    nu=0.3 sig0=250 H=2e3, face z=0 clamped, u_z prescribed on z=4 (ramped
    by the driver; total 5x the yield strain over the length)."""
    base = cantilever_hex(nx, ny, nz, jitter=0.2, seed=seed)
    lz = base.meta["lz"]
    if u_end is None:
        u_end = 5.0 * 250.0 / 210e3 * lz
    nn = base.coords.shape[0]
    bc = base.bc_mask.copy()
    val = np.zeros((nn, 3))
    top = np.arange(nn - (nx + 1) * (ny + 1), nn)
    bc[top, 2] = True
    val[top, 2] = u_end
    mat = np.array([[MAT_MISES, 210e3, 0.3, 250.0, 2e3, 0, 0, 0]], dtype=np.float64)
    return Deck(f"block_{nx}x{ny}x{nz}_hex_mises", LSPACE, base.coords, base.conn,
                np.zeros(base.conn.shape[0], dtype=np.int32), mat, bc, val,
                meta=dict(nx=nx, ny=ny, nz=nz, seed=seed, u_end=u_end))


# ---------------------------------------------------------------- tetrahedra
# Kuhn (Freudenthal) split of a cube into 6 tets sharing the main diagonal
# (0,0,0)-(1,1,1); every tet is ordered so that the reference's detJ
# (fei3dtetlin.C:138-140) is positive.
_KUHN = [(0, 1, 3, 7), (0, 1, 5, 7), (0, 2, 3, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 6, 7)]


def _tet_det(c):
    a, b, cc, d = c[:, 0], c[:, 1], c[:, 2], c[:, 3]
    return np.einsum("ij,ij->i", d - a, np.cross(b - a, cc - a))


def kuhn_tets(nx, ny, nz, lx, ly, lz, jitter=0.0, seed=0):
    coords = _grid_nodes(nx, ny, nz, lx, ly, lz)
    if jitter > 0:
        coords = _jitter(coords, nx, ny, nz, (lx / nx, ly / ny, lz / nz), jitter, seed)
    ez, ey, ex = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    sx, sy = nx + 1, (nx + 1) * (ny + 1)
    corner = np.stack([(ex + (c & 1)) + sx * (ey + ((c >> 1) & 1)) + sy * (ez + ((c >> 2) & 1))
                       for c in range(8)], axis=1)
    tets = np.concatenate([corner[:, list(t)][:, None, :] for t in _KUHN], axis=1).reshape(-1, 4)
    # reference detJ = (x4-x1).((x2-x1)x(x3-x1)) orientation; flip 2<->3 where negative
    det = _tet_det(coords[tets])
    neg = det < 0
    tets[neg] = tets[neg][:, [0, 2, 1, 3]]
    det = _tet_det(coords[tets])
    if not np.all(det > 0):
        raise ValueError("tet mesh has non-positive volumes; reduce jitter")
    return coords, tets.astype(np.int32)


def config2(nx=70, ny=70, nz=82, seed=20261204, jitter=0.1, u_end=None):
    """Kuhn-split tet block 1 x 1 x nz/nx, MisesMat E=210e3 nu=0.3 sig0=250
    H=2e3, face z=0 clamped, u_z prescribed on z=L (ramped by the driver)."""
    lx, ly, lz = 1.0, 1.0, nz / nx
    coords, conn = kuhn_tets(nx, ny, nz, lx, ly, lz, jitter=jitter, seed=seed)
    mat = np.array([[MAT_MISES, 210e3, 0.3, 250.0, 2e3, 0, 0, 0]], dtype=np.float64)
    emat = np.zeros(conn.shape[0], dtype=np.int32)
    nn = coords.shape[0]
    bc = np.zeros((nn, 3), dtype=bool)
    val = np.zeros((nn, 3))
    bc[: (nx + 1) * (ny + 1)] = True
    top = np.arange(nn - (nx + 1) * (ny + 1), nn)
    if u_end is None:
        u_end = 5.0 * 250.0 / 210e3 * lz  # 5x the yield strain over the length
    bc[top, 2] = True
    val[top, 2] = u_end
    return Deck(f"tet_{nx}x{ny}x{nz}_mises", LTRSPACE, coords, conn, emat, mat, bc, val,
                meta=dict(nx=nx, ny=ny, nz=nz, seed=seed, u_end=u_end))


# ------------------------------------------------------------ surface loads
_HEX_SURF = np.array([[1, 4, 3, 2], [5, 6, 7, 8], [1, 2, 6, 5], [2, 3, 7, 6], [3, 4, 8, 7], [4, 1, 5, 8]]) - 1
_TET_SURF = np.array([[1, 3, 2], [1, 2, 4], [2, 3, 4], [1, 4, 3]]) - 1
_G2 = 0.577350269189626


def surface_load_vector(deck: Deck, factor=1.0):
    """ConstantSurfaceLoad in the global c.s. (structuralelement.C:110-164):
    hexa faces with the 2x2 Gauss rule and |a x b|/16 surface jacobian
    (fei3dhexalin.C:447-458,532-566), tetra faces with the 1-point rule."""
    f_free = np.zeros(deck.neq)
    f_presc = np.zeros(deck.npeq)
    if deck.faces.shape[0] == 0:
        return f_free, f_presc
    el, sf = deck.faces[:, 0], deck.faces[:, 1] - 1
    t = deck.traction * factor
    if deck.etype == LSPACE:
        snodes = deck.conn[el[:, None], _HEX_SURF[sf]]              # (nf, 4)
        X = deck.coords[snodes]                                       # (nf, 4, 3)
        fe = np.zeros((el.shape[0], 4, 3))
        for ksi in (-_G2, _G2):
            for eta in (-_G2, _G2):
                N = np.array([(1 + ksi) * (1 + eta), (1 - ksi) * (1 + eta), (1 - ksi) * (1 - eta), (1 + ksi) * (1 - eta)]) * 0.25
                dk = np.array([1 + eta, -(1 + eta), -(1 - eta), 1 - eta])
                de = np.array([1 + ksi, 1 - ksi, -(1 - ksi), -(1 + ksi)])
                a = np.einsum("k,fkd->fd", dk, X)
                b = np.einsum("k,fkd->fd", de, X)
                dA = np.linalg.norm(np.cross(a, b), axis=1) * 0.0625
                fe += N[None, :, None] * t[:, None, :] * dA[:, None, None]
    elif deck.etype == LTRSPACE:
        snodes = deck.conn[el[:, None], _TET_SURF[sf]]
        X = deck.coords[snodes]
        dA = np.linalg.norm(np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), axis=1) * 0.5
        N = np.array([1.0 / 3.0, 1.0 / 3.0, 1.0 - 2.0 / 3.0])
        fe = N[None, :, None] * t[:, None, :] * dA[:, None, None]
    else:
        raise NotImplementedError("surface loads on QSpace faces")
    q = deck.eqn[snodes]                                               # (nf, ns, 3)
    m = q > 0
    np.add.at(f_free, q[m] - 1, fe[m])
    m = q < 0
    np.add.at(f_presc, -q[m] - 1, fe[m])
    return f_free, f_presc


def _default_rule(etype):
    """Default volume Gauss rule: 2x2x2 (LSpace), 3x3x3 (QSpace), the
    one-point tetrahedron rule (LTRSpace; lspace.C:68, qspace.C:56,
    ltrspace.C:69). Points (ngp, 3) and weights (ngp,)."""
    if etype == LTRSPACE:
        return np.full((1, 3), 0.25), np.array([1.0 / 6.0])
    if etype == LSPACE:
        x, w = np.array([-1.0, 1.0]) / np.sqrt(3.0), np.ones(2)
    else:
        x, w = np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)]), np.array([5.0, 8.0, 5.0]) / 9.0
    I, J, K = np.meshgrid(np.arange(x.size), np.arange(x.size), np.arange(x.size), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    return np.stack([x[I], x[J], x[K]], axis=1), w[I] * w[J] * w[K]


def _shape(etype, xi):
    """N (nen,) and dN/dxi (nen, 3) at one local point, in the node orders of
    fei3dhexalin.C:46-62, fei3dhexaquad.C:46-73 and fei3dtetlin.C:46-54."""
    if etype == LTRSPACE:
        u, v, w = xi
        return (np.array([u, v, w, 1.0 - u - v - w]),
                np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [-1.0, -1.0, -1.0]]))
    nen = NEN[etype]
    N, dN = np.empty(nen), np.empty((nen, 3))
    for a in range(nen):
        p = _QS_POS[a].astype(np.float64)
        g = 1.0 + p * xi                       # per-direction linear factors
        if etype == LSPACE or a < 8:
            lin = 0.125 * g.prod()
            dlin = np.array([0.125 * p[j] * np.prod(np.delete(g, j)) for j in range(3)])
            if etype == LSPACE:
                N[a], dN[a] = lin, dlin
            else:                              # serendipity corner
                s = float(p @ xi) - 2.0
                N[a], dN[a] = lin * s, dlin * s + lin * p
        else:                                  # mid-edge node, p[k] == 0
            k = int(np.nonzero(p == 0)[0][0])
            o = [j for j in range(3) if j != k]
            bub = 1.0 - xi[k] ** 2
            N[a] = 0.25 * bub * g[o[0]] * g[o[1]]
            dN[a, k] = -0.5 * xi[k] * g[o[0]] * g[o[1]]
            dN[a, o[0]] = 0.25 * bub * p[o[0]] * g[o[1]]
            dN[a, o[1]] = 0.25 * bub * p[o[1]] * g[o[0]]
    return N, dN


def body_load_vector(deck: Deck, elems, dens, b, factor=1.0):
    """DeadWeight in the global c.s. (structuralelement.C:262-296): per
    element, sum over the default Gauss rule of N^T b |det J| w rho.
    ``elems`` 0-based element indices, ``dens`` their densities, ``b`` the
    force per unit mass (3,). Returns (f_free, f_presc)."""
    f_free = np.zeros(deck.neq)
    f_presc = np.zeros(deck.npeq)
    elems = np.asarray(elems, dtype=np.int64)
    if elems.size == 0:
        return f_free, f_presc
    bf = np.asarray(b, dtype=np.float64) * factor
    nodes = deck.conn[elems]                                           # (ne, nen)
    X = deck.coords[nodes]                                             # (ne, nen, 3)
    fe = np.zeros(X.shape)
    pts, wts = _default_rule(deck.etype)
    for xi, w in zip(pts, wts):
        N, dN = _shape(deck.etype, xi)
        J = np.einsum("ai,eaj->eij", dN, X)
        dV = np.abs(np.linalg.det(J)) * w * np.asarray(dens, dtype=np.float64)
        fe += N[None, :, None] * bf[None, None, :] * dV[:, None, None]
    q = deck.eqn[nodes]                                                # (ne, nen, 3)
    m = q > 0
    np.add.at(f_free, q[m] - 1, fe[m])
    m = q < 0
    np.add.at(f_presc, -q[m] - 1, fe[m])
    return f_free, f_presc


# ------------------------------------------------------------------- RVEs
# QSpace local node positions in (xi, eta, zeta) (fei3dhexaquad.C node
# order: 8 corners as LSpace, 4 mid-edge nodes of the zeta = +1 face, 4 of
# the zeta = -1 face, 4 mid-height edges); xi, eta, zeta run along x, y, z.
_QS_POS = np.array([(-1, -1, 1), (-1, 1, 1), (1, 1, 1), (1, -1, 1), (-1, -1, -1), (-1, 1, -1), (1, 1, -1), (1, -1, -1),
                    (-1, 0, 1), (0, 1, 1), (1, 0, 1), (0, -1, 1), (-1, 0, -1), (0, 1, -1), (1, 0, -1), (0, -1, -1),
                    (-1, -1, 0), (-1, 1, 0), (1, 1, 0), (1, -1, 0)], dtype=np.int64)


def _rve_mesh(n, etype):
    """Unit cube of n^3 LSpace or QSpace cells: coordinates and connectivity."""
    if etype == LSPACE:
        return _grid_nodes(n, n, n, 1.0, 1.0, 1.0), _hex_conn(n, n, n)
    m = 2 * n + 1   # doubled lattice: corner and mid-edge points only
    I, J, K = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    keep = ((I % 2) + (J % 2) + (K % 2)) <= 1
    lat = np.full(m ** 3, -1, dtype=np.int64)
    order = np.nonzero(keep)[0]
    lat[order] = np.arange(order.size)
    # node order: z-major like the LSpace grid
    key = K[order] * m * m + J[order] * m + I[order]
    perm = np.argsort(key, kind="stable")
    renum = np.empty(order.size, dtype=np.int64)
    renum[perm] = np.arange(order.size)
    coords = np.round(np.stack([I[order][perm], J[order][perm], K[order][perm]], axis=1) / (2.0 * n), 12)
    ez, ey, ex = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    conn = np.empty((ex.size, 20), dtype=np.int64)
    for a, (u, v, w) in enumerate(_QS_POS):
        li = 2 * ex + 1 + u
        lj = 2 * ey + 1 + v
        lk = 2 * ez + 1 + w
        conn[:, a] = renum[lat[li * m * m + lj * m + lk]]
    return coords, conn.astype(np.int32)


def rve(n, etype, vf, seed, E_matrix=10e3, E_incl=100e3, nu=0.25):
    """Two-phase RVE: unit cube of n^3 cells (LSpace or QSpace), seeded
    spherical inclusions (radius 0.08..0.2) placed until the element volume
    fraction reaches vf (an element is inclusion when its centre is inside a
    sphere); every boundary node prescribed (affine displacement of a macro
    strain, set per load case with affine_values)."""
    coords, conn = _rve_mesh(n, etype)
    rng = np.random.default_rng(seed)
    centres = coords[conn[:, :8]].mean(axis=1)
    incl = np.zeros(conn.shape[0], dtype=bool)
    for _ in range(10000):
        if incl.mean() >= vf:
            break
        c = rng.uniform(0.0, 1.0, 3)
        r = rng.uniform(0.08, 0.2)
        incl |= ((centres - c) ** 2).sum(axis=1) <= r * r
    mat = np.array([[MAT_ISOLE, E_matrix, nu, 0, 0, 0, 0, 0], [MAT_ISOLE, E_incl, nu, 0, 0, 0, 0, 0]])
    on_bnd = np.any((coords <= 1e-12) | (coords >= 1.0 - 1e-12), axis=1)
    bc = np.repeat(on_bnd[:, None], 3, axis=1)
    d = Deck(f"rve_{'q' if etype == QSPACE else 'l'}{n}_vf{vf:.2f}_s{seed}", etype, coords, conn,
             incl.astype(np.int32), mat, bc, np.zeros_like(coords),
             meta=dict(n=n, vf_target=vf, vf=float(incl.mean()), seed=seed))
    return d


VOIGT_PAIRS = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


def macro_gradient(eps_voigt):
    """Symmetric displacement gradient of a Voigt macro strain (engineering
    shear split over the two off-diagonal entries)."""
    g = np.zeros((3, 3))
    for k, (i, j) in enumerate(VOIGT_PAIRS):
        if i == j:
            g[i, i] = eps_voigt[k]
        else:
            g[i, j] = g[j, i] = 0.5 * eps_voigt[k]
    return g


def affine_values(deck, eps_voigt, centre=(0.5, 0.5, 0.5)):
    """Prescribed values u = grad . (x - centre) of PrescribedGradient::give
    (prescribedgradient.C:59-93) at the prescribed equations."""
    g = macro_gradient(eps_voigt)
    u = (deck.coords - np.asarray(centre)[None, :]) @ g.T
    m = deck.eqn < 0
    out = np.zeros(deck.npeq)
    out[-deck.eqn[m] - 1] = u[m]
    return out


def homogenized_stress(deck, reactions, centre=(0.5, 0.5, 0.5)):
    """Volume-averaged stress of an RVE under affine boundary displacements
    from its boundary reactions: sigma = (1/V) sum_b R_b (x) (x_b - centre),
    symmetrised (V = 1), in Voigt order."""
    m = deck.eqn < 0
    node, comp = np.nonzero(m)
    R = np.zeros_like(deck.coords)
    R[node, comp] = reactions[-deck.eqn[m] - 1]
    X = deck.coords - np.asarray(centre)[None, :]
    S = R.T @ X
    S = 0.5 * (S + S.T)
    return np.array([S[i, j] for (i, j) in VOIGT_PAIRS])


def config4(count=512, seed=20261206, nmin=8, nmax=32, q_share=0.3):
    """The RVE batch: sizes n ~ U{nmin..nmax}, QSpace in a seeded q_share of
    the RVEs, volume fractions U(0.1, 0.4), and one seeded random macro strain
    of magnitude 1e-3 per RVE.  Returns (specs, macro strains); decks are
    built on demand with rve(*spec)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, count)
    q = rng.random(count) < q_share
    vfs = rng.uniform(0.1, 0.4, count)
    specs = [(int(ns[i]), QSPACE if q[i] else LSPACE, float(vfs[i]), int(seed + 1 + i)) for i in range(count)]
    e = rng.standard_normal((count, 6))
    e *= 1e-3 / np.linalg.norm(e, axis=1, keepdims=True)
    return specs, e
