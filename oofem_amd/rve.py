"""Batched RVE homogenisation on the GPU (the configuration-4 workload).

Each RVE is a unit cube of LSpace or QSpace cells with two isotropic phases
under affine boundary displacements u = grad(eps) . (x - x_c)
(PrescribedGradient::give, prescribedgradient.C:59-93).  Its homogenised
stiffness comes from six unit-strain solves (PrescribedGradient::computeField
style, prescribedgradient.C:161-181, without the PETSc-only computeTangent)
and the boundary-reaction average sigma = (1/V) sum_b R_b (x) (x_b - x_c).

GPU layout: the RVEs of one element type are concatenated into one mesh
(disjoint node ranges), so the batched evaluator assembles one
block-diagonal stiffness; chunks keep every matrix under the int32 index
range.  All blocks and all six right-hand sides of a chunk are solved by
hfem_batch (mixed-precision PCG with FP64 reliable updates): blocks above
the split by a bulk-synchronous phase over the whole GPU, the rest by one
persistent launch, largest blocks first; kernel shapes from the autotuner.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import meshgen
from .hipfem import BatchedEvaluator, DeviceVector, HipCSRSparseMtrx, _check, _p, lib

NNZ_ROW = {meshgen.LSPACE: 81, meshgen.QSPACE: 200}


def _strided(ctx, n, a, x, incx, b, y, incy):
    """y[i incy] = a x[i incx] + b y[i incy] on device addresses."""
    _check(lib().hipfem_dev_axpby_strided(ctx.h, C.c_int64(n), C.c_double(a), C.c_void_p(x), C.c_int64(incx),
                                          C.c_double(b), C.c_void_p(y), C.c_int64(incy)), "dev_axpby_strided")


def _concat(decks):
    """One deck holding the RVEs side by side (node and element offsets; the
    shared two-phase material table)."""
    coords = np.concatenate([d.coords for d in decks])
    off = np.cumsum([0] + [d.nnode for d in decks])
    conn = np.concatenate([d.conn + o for d, o in zip(decks, off[:-1])])
    emat = np.concatenate([d.emat for d in decks])
    bc = np.concatenate([d.bc_mask for d in decks])
    big = meshgen.Deck("rve_chunk", decks[0].etype, coords, conn, emat, decks[0].mat, bc, np.zeros_like(coords))
    return big, off


def block_rows(deck, node_off):
    """First free equation (0-based) of every RVE of a concatenated deck, and
    the total: equations are numbered node by node, so each RVE's free
    equations form one contiguous range."""
    rows = np.zeros(len(node_off), dtype=np.int32)
    for i in range(len(node_off) - 1):
        e = deck.eqn[node_off[i]:node_off[i + 1]]
        free = e[e > 0]
        if free.size == 0 or free.max() - free.min() + 1 != free.size or free.min() - 1 != rows[i]:
            raise ValueError(f"RVE {i}: free equations are not one contiguous range after the previous RVE")
        rows[i + 1] = free.max()
    return rows


class _Chunk:
    def __init__(self, ctx, decks, ids):
        self.ctx = ctx
        self.ids = ids
        self.deck, self.node_off = _concat(decks)
        d = self.deck
        self.ev = BatchedEvaluator(ctx, d)
        self.K = HipCSRSparseMtrx(ctx)
        self.ev.build_structure(self.K)
        self.ev.assemble_stiffness(self.K)
        rows = block_rows(d, self.node_off)
        self.rows = rows
        self.h = C.c_void_p()
        _check(lib().hfem_batch_create(self.K.h, len(decks), _p(rows), C.byref(self.h)), "hfem_batch_create")
        self.n = d.neq
        self.b = DeviceVector(ctx, max(1, 6 * self.n))
        self.x = DeviceVector(ctx, max(1, 6 * self.n))
        self.up = [meshgen.affine_values(d, e) for e in np.eye(6)]
        # device-resident work vectors of the right-hand sides and reactions
        self.d_up = [DeviceVector(ctx, max(1, d.npeq), u) for u in self.up]
        self.d_u = DeviceVector(ctx, max(1, self.n))
        self.d_f = DeviceVector(ctx, max(1, self.n))
        self.d_fp = DeviceVector(ctx, max(1, d.npeq))
        # prescribed dofs: reaction index, RVE, component and the node's
        # position relative to the cell centre
        node, comp = np.nonzero(d.eqn < 0)
        self.p_idx = -d.eqn[node, comp] - 1
        self.p_bin = (np.searchsorted(self.node_off, node, side="right") - 1) * 3 + comp
        self.p_xc = d.coords[node] - 0.5

    def rhs(self):
        """b_k = -f_int(free 0, prescribed affine_k) for the six unit strains,
        interleaved, on the device."""
        self.d_u.fill(0.0)
        for k in range(6):
            self.ev.internal_forces_dev(self.d_u.addr, self.d_up[k].addr, self.d_f.addr, self.d_fp.addr)
            _strided(self.ctx, self.n, -1.0, self.d_f.addr, 1, 0.0, self.b.addr + 8 * k, 6)

    def solve(self):
        nb = len(self.ids)
        it = np.zeros(nb, dtype=np.int32)
        upd = np.zeros(nb, dtype=np.int32)
        res = np.zeros(nb)
        _check(lib().hfem_batch_solve_dev(self.h, C.c_void_p(self.b.addr), C.c_void_p(self.x.addr), _p(it), _p(upd),
                                          _p(res)), "hfem_batch_solve")
        return it, upd, res

    def homogenized(self):
        """6x6 stiffness per RVE from the reactions of the six solves: per RVE
        the sum over its prescribed dofs of R (x) (x - x_c), symmetrised (the
        unit cell has volume 1)."""
        nrve = len(self.ids)
        out = np.zeros((nrve, 6, 6))
        for k in range(6):
            _strided(self.ctx, self.n, 1.0, self.x.addr + 8 * k, 6, 0.0, self.d_u.addr, 1)
            self.ev.internal_forces_dev(self.d_u.addr, self.d_up[k].addr, self.d_f.addr, self.d_fp.addr)
            fp = self.d_fp.download()
            r = fp[self.p_idx]
            S = np.stack([np.bincount(self.p_bin, weights=r * self.p_xc[:, j], minlength=3 * nrve)
                          for j in range(3)], axis=1).reshape(nrve, 3, 3)
            S = 0.5 * (S + np.transpose(S, (0, 2, 1)))
            out[:, :, k] = np.stack([S[:, i, j] for (i, j) in meshgen.VOIGT_PAIRS], axis=1)
        return out

    def close(self):
        if self.h:
            lib().hfem_batch_destroy.argtypes = [C.c_void_p]
            lib().hfem_batch_destroy.restype = None
            lib().hfem_batch_destroy(self.h)
            self.h = C.c_void_p()
        for o in (self.b, self.x, self.d_u, self.d_f, self.d_fp, *self.d_up, self.K, self.ev):
            o.close()


class RVEBatch:
    """All RVEs of a batch on one GPU, grouped by element type into chunks of
    at most nnz_budget stiffness entries."""

    def __init__(self, ctx, specs, nnz_budget=4e8, log=None):
        self.ctx = ctx
        self.specs = specs
        self.chunks = []
        t = time.perf_counter()
        for et in (meshgen.QSPACE, meshgen.LSPACE):
            ids = [i for i, s in enumerate(specs) if s[1] == et]
            cur, cur_nnz = [], 0.0
            for i in ids + [None]:
                if i is not None:
                    n = specs[i][0]
                    est = 3.0 * (n + 1) ** 3 * (4 if et == meshgen.QSPACE else 1) * NNZ_ROW[et]
                if cur and (i is None or cur_nnz + est > nnz_budget):
                    if log:
                        log(f"[rve] building {len(cur)} {'QSpace' if et == meshgen.QSPACE else 'LSpace'} RVEs "
                            f"(about {cur_nnz:.2e} entries)")
                    decks = [meshgen.rve(*specs[j]) for j in cur]
                    self.chunks.append(_Chunk(ctx, decks, cur))
                    if log:
                        log(f"[rve] chunk of {len(cur)} {'QSpace' if et == meshgen.QSPACE else 'LSpace'} RVEs: "
                            f"neq {self.chunks[-1].n}, nnz {self.chunks[-1].K.nnz()}")
                    cur, cur_nnz = [], 0.0
                if i is not None:
                    cur.append(i)
                    cur_nnz += est
        self.t_setup = time.perf_counter() - t

    def set_options(self, tol=1e-8, delta=0.1, maxit=20000, split=-1, nodeblock=True, bulk_form=None):
        """split: blocks with more stored entries than this are solved by the
        bulk-synchronous phase (-1: automatic, 0: every block); bulk_form:
        0 CSR rows, 1 3x3 node-block products from the CSR values, 2 the
        same from a block-major padded FP32 copy (nodeblock=False is 0)."""
        form = bulk_form if bulk_form is not None else int(bool(nodeblock))
        for c in self.chunks:
            _check(lib().hfem_batch_set_options(c.h, C.c_double(tol), C.c_double(delta), C.c_int(maxit)), "options")
            _check(lib().hfem_batch_set_split(c.h, C.c_longlong(split)), "split")
            _check(lib().hfem_batch_set_bulk_form(c.h, int(form)), "bulk form")

    def set_profiling(self, on=True):
        for c in self.chunks:
            _check(lib().hfem_batch_set_profiling(c.h, int(bool(on))), "profiling")

    def profile(self):
        """HIP-event timing of the bulk products summed over the chunks:
        dict(launches, ms, bytes)."""
        n, ms, by = 0, 0.0, 0.0
        for c in self.chunks:
            a, b, d = C.c_longlong(0), C.c_double(0), C.c_double(0)
            _check(lib().hfem_batch_profile(c.h, C.byref(a), C.byref(b), C.byref(d)), "profile")
            n, ms, by = n + a.value, ms + b.value, by + d.value
        return dict(launches=n, ms=ms, bytes=by)

    def bulk_forms(self):
        """1 (node-block) or 0 (CSR rows) per chunk."""
        lib().hfem_batch_bulk_form.argtypes = [C.c_void_p]
        return [lib().hfem_batch_bulk_form(c.h) for c in self.chunks]

    def partition(self):
        """(bulk blocks, split used) per chunk."""
        out = []
        for c in self.chunks:
            nb, thr = C.c_int(0), C.c_longlong(0)
            _check(lib().hfem_batch_partition_info(c.h, C.byref(nb), C.byref(thr)), "partition")
            out.append((nb.value, thr.value))
        return out

    def autotune(self, sample=8, log=None):
        nshape = lib().hfem_batch_nshapes()
        out = []
        nbulk = lib().hfem_batch_nbulk_shapes()
        for c in self.chunks:
            if log:
                log(f"[rve] autotune on {sample} median blocks of a chunk of {len(c.ids)}")
            c.rhs()
            ms = np.zeros(nshape)
            best = C.c_int(0)
            _check(lib().hfem_batch_autotune(c.h, C.c_void_p(c.b.addr), C.c_void_p(c.x.addr), C.c_int(sample),
                                             _p(ms), C.byref(best)), "autotune")
            bs, lpr = C.c_int(0), C.c_int(0)
            lib().hfem_batch_shape_info(best.value, C.byref(bs), C.byref(lpr))
            msb = np.zeros(nbulk)
            bestb = C.c_int(0)
            _check(lib().hfem_batch_autotune_bulk(c.h, C.c_void_p(c.b.addr), C.c_void_p(c.x.addr), _p(msb),
                                                  C.byref(bestb)), "bulk autotune")
            blpr = C.c_int(0)
            lib().hfem_batch_bulk_shape_info(bestb.value, C.byref(blpr))
            nb, thr = C.c_int(0), C.c_longlong(0)
            _check(lib().hfem_batch_partition_info(c.h, C.byref(nb), C.byref(thr)), "partition")
            lib().hfem_batch_bulk_form.argtypes = [C.c_void_p]
            out.append(dict(etype=int(c.deck.etype), blocks=len(c.ids), bulk_blocks=nb.value, split=thr.value,
                            bulk_nodeblock=lib().hfem_batch_bulk_form(c.h),
                            best=best.value, block=bs.value, lanes_per_row=lpr.value,
                            ms=[round(float(v), 3) for v in ms], bulk_lanes_per_row=blpr.value,
                            bulk_ms=[round(float(v), 3) for v in msb]))
        return out

    def run(self, log=None):
        """Six unit-strain solves per RVE; returns per-RVE 6x6 stiffness,
        iterations, reliable updates, residuals and the solve time (the
        persistent launches only; right-hand sides and reactions apart)."""
        nrve = len(self.specs)
        C6 = np.zeros((nrve, 6, 6))
        it = np.zeros(nrve, dtype=np.int32)
        upd = np.zeros(nrve, dtype=np.int32)
        res = np.zeros(nrve)
        t_solve = t_rhs = t_hom = 0.0
        for c in self.chunks:
            t = time.perf_counter()
            c.rhs()
            self.ctx.synchronize()
            t_rhs += time.perf_counter() - t
            t = time.perf_counter()
            i_, u_, r_ = c.solve()
            t_solve += time.perf_counter() - t
            if log:
                log(f"[rve] chunk of {len(c.ids)} solved in {time.perf_counter() - t:.2f} s, iterations max {i_.max()}")
            ids = np.asarray(c.ids)
            it[ids], upd[ids], res[ids] = i_, u_, r_
            t = time.perf_counter()
            C6[ids] = c.homogenized()
            t_hom += time.perf_counter() - t
        return dict(C=C6, iters=it, updates=upd, resid=res, t_solve=t_solve, t_rhs=t_rhs, t_homogenize=t_hom)

    def close(self):
        for c in self.chunks:
            c.close()
        self.chunks = []
