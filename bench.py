#!/usr/bin/env python3
"""Benchmark of the assembly-and-solve path on one MI355X.

Default workload (config 3): the 64x64x256 LSpace block (1,048,576 elements,
3,240,575 free equations, jitter +-0.2 h) with MisesMat (E 210e3, nu 0.3,
sig0 250, H 2e3), clamped at z = 0, u_z prescribed on z = 4 and ramped to
5x the yield strain over 8 load steps.  One step = one StaticStructural load
step: tangent initial guess, full Newton with the consistent tangent to the
1e-8 force criterion, Jacobi-PCG to 1e-6 per correction.  After --c3-warm
untimed steps, --c3-steps are timed for three forms:
  fp64             hipcsr + hippcg FP64 + plain return mapping (value)
  mixed            hippcg mixed 2: FP32 PCG with FP64 reliable updates and
                   the FP32 element-by-element operator
  mixed_surrogate  the same with the trained MisesMat surrogate (BF16 MFMA
                   inference, residual-checked, plain fallback)
value = free DOFs x timed steps / wall time of the FP64 form; the others,
their speed-ups and their displacement / stress errors against it are in
config.variants.

roofline: the FP64 PCG SpMV (row-group form on this pattern: the top face
constrains u_z only), algorithmic bytes per launch over its mean HIP-event
duration.  cpu_baseline: the oracle (one thread) on a 64x64x16 slab of the
same problem, scaled by elements, nnz and the GPU's iteration counts.

Other workloads: --workload cfg1 (LinearStatic cantilever, FP64 vs mixed),
cfg0, cfg4 (512-RVE batch, solves/s); --mode kernels (counter runs).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Newton-step DOFs/s at 1e-8 residual, plain FP64 vs mixed+surrogate; SpMV GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_deck(args):
    from oofem_amd import meshgen
    if args.workload == "cfg0":
        return meshgen.config0()
    if args.workload == "cfg3":
        return meshgen.config3(nx=args.nx, ny=args.ny, nz=args.nz)
    if args.workload == "cfg2":
        return meshgen.config2()
    return meshgen.config1(nx=args.nx, ny=args.ny, nz=args.nz)


def apply_shape(K, spec):
    """--shape lanes:L:U[:nt] or stream:R:U"""
    if not spec:
        return
    p = spec.split(":")
    if p[0] == "lanes":
        K.set_spmv_shape(int(p[1]), int(p[2]), int(p[3]) if len(p) > 3 else 0)
    elif p[0] == "stream":
        K.set_spmv_stream(int(p[1]), int(p[2]))
    else:
        raise SystemExit(f"bad --shape {spec}")


def run_kernels(args, deck):
    """Kernel microbenchmark for counter runs: a copy-like calibration kernel
    of known size (y = x over 2^27 doubles: 1 GiB read + 1 GiB written), then
    --iters PCG iterations (FP64 and FP32-inner) on the assembled K."""
    import numpy as np

    from oofem_amd.hipfem import BatchedEvaluator, Context, DeviceVector, HipCSRSparseMtrx, HipPCGSolver
    ctx = Context(0)
    n_copy = 1 << 27
    a = DeviceVector(ctx, n_copy)
    b = DeviceVector(ctx, n_copy)
    from oofem_amd.hipfem import lib as _lib
    import ctypes as _C
    for _ in range(3):
        b.axpby(1.0, a, 0.0)
    ctx.synchronize()
    copy_gbs = {}
    for name, fn in (("axpby", lambda: b.axpby(1.0, a, 0.0)),
                     ("hipMemcpy", lambda: _lib().hipfem_dev_copy(ctx.h, _C.c_int64(n_copy), a.ptr, b.ptr, 0)),
                     ("wide_kernel", lambda: _lib().hipfem_dev_copy(ctx.h, _C.c_int64(n_copy), a.ptr, b.ptr, 1))):
        fn()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(5):
            fn()
        ctx.synchronize()
        copy_gbs[name] = round(5 * 2 * 8 * n_copy / (time.perf_counter() - t) / 1e9, 1)
    K = HipCSRSparseMtrx(ctx)
    ev = BatchedEvaluator(ctx, deck)
    ev.build_structure(K)
    ev.assemble_stiffness(K)
    apply_shape(K, args.shape)
    f, _ = deck.external_load()
    if not f.any():
        f = np.ones(deck.neq)          # displacement-driven decks carry no load: any non-zero b will do
    db = DeviceVector(ctx, deck.neq, f)
    dx = DeviceVector(ctx, deck.neq)
    out = {}
    for prec, name in ((0, "fp64_nodeblock"), (1, "fp32_nodeblock"), (2, "fp32_element_operator")):
        s = HipPCGSolver(ctx, lstol=1e-30, lsiter=args.iters, lsprecond=1, precision=prec, inner_tol=1e-30)
        s.set_profiling(True)
        dx.fill(0.0)
        s.solve_dev(K, db.addr, dx.addr)
        p = s.profile()
        out[name] = (p["ms64"] / max(p["n64"], 1)) if prec == 0 else (p["ms32"] / max(p["n32"], 1))
        s.close()
    print(json.dumps(dict(mode="kernels", neq=deck.neq, nnz=K.nnz(), shape=K.spmv_shape(), copy_doubles=n_copy,
                          copy_gbs=copy_gbs, spmv_avg_ms=out)), flush=True)
    return 0


def run_gpu(args, deck, precision):
    """LinearStatic steps through the C++ host driver (class factory ->
    hipcsr + hippcg, batched evaluator); data resident on the GPU."""
    from oofem_amd.problem import StructuralProblem
    P = StructuralProblem(deck, lstol=args.tol, lsiter=args.maxit, lsprecond=1, mixed=precision,
                          irtol=args.inner_tol)
    f_ext, f_ext_p = deck.external_load()
    P.set_prescribed(deck.prescribed_values())
    P.set_load(f_ext, f_ext_p)
    t = time.perf_counter()
    for i in range(args.warmup):          # the first step also builds the sparse structure
        it, res = P.linear_static()
        log(f"[bench] {precision=} warmup {i}: iters={it} res={res:.3e} ({time.perf_counter() - t:.1f} s)")
    P.set_profiling(True)
    iters = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        it, res = P.linear_static()
        iters.append(it)
    dt = time.perf_counter() - t0         # linear_static returns after the device work has finished
    p = P.profile()
    u = P.solution()
    R = P.reactions()
    info = P.solver_info()
    out = dict(dt=dt, iters=iters, res=res, nnz=P.matrix_nnz(), n_spmv=p["n64"], ms_spmv=p["ms64"],
               n_spmv32=p["n32"], ms_spmv32=p["ms32"], outer=info["outer"], matrix_free=info["matrix_free"],
               u=u, reactions=R)
    P.close()
    return out


def run_newton(args, deck, precision, warm=1, steps=None, total=None, surrogate=None, keep_state=False, forcing=0.0,
               initial_guess=1):
    """StaticStructural + NRSolver load steps (full Newton, consistent tangent,
    rtolf 1e-8, tangent initial guess) with the prescribed end displacement
    ramped over `total` steps: `warm` untimed steps, then `steps` timed ones.
    surrogate: (params, band, tol) of a trained MisesMat network, or None;
    forcing > 0: inexact Newton with forcing terms up to that value."""
    import numpy as np

    from oofem_amd.problem import StructuralProblem
    steps = args.nl_steps if steps is None else steps
    total = args.nl_total if total is None else total
    P = StructuralProblem(deck, lstol=args.nl_lstol, lsiter=args.maxit, lsprecond=1, mixed=precision,
                          irtol=args.inner_tol, rtolf=1e-8, maxiter=60, manrmsteps=1, forcing=forcing,
                          initial_guess=initial_guess)
    P.set_load(np.zeros(deck.neq), np.zeros(deck.npeq))
    net = None
    if surrogate is not None:
        from oofem_amd.hipfem import MisesSurrogate
        m = deck.mat[0]
        net = MisesSurrogate(P.context(), m[1], m[2], m[3], m[4])
        net.set_params(surrogate[0])
        P.set_surrogate(net, band=surrogate[1], tol=surrogate[2])
    reports = []
    t0 = None
    for k in range(1, warm + steps + 1):
        P.set_prescribed(deck.prescribed_values(k / total))
        if k == warm + 1:
            P.set_profiling(True)
            if net is not None:
                P.surrogate_stats(reset=True)
            t0 = time.perf_counter()
        rep = P.nonlinear_step()
        if rep["status"] != 1:
            raise RuntimeError(f"Newton step {k} failed: {rep}")
        P.update()
        reports.append(rep)
        log(f"[bench] precision {precision}{' +surrogate' if net else ''}{' +forcing' if forcing else ''} "
            f"newton step {k}: {rep}")
    dt = time.perf_counter() - t0
    p = P.profile()
    out = dict(dt=dt, steps=steps, newton=[r["newton"] for r in reports[warm:]], force_error=[r["force_error"] for r in reports],
               reactions=P.reactions(),
               pcg=[r["pcg"] for r in reports[warm:]], n_spmv=p["n64"], ms_spmv=p["ms64"], n_spmv32=p["n32"],
               ms_spmv32=p["ms32"], nnz=P.matrix_nnz(), u=P.solution(), info=P.solver_info(),
               layout=P.matrix_layout())
    if keep_state:
        gs = P.state()
        out["stress"] = gs[:, 20:26].copy()
        out["kappa"] = gs[:, 6].copy()
        del gs
    if net is not None:
        out["surrogate"] = P.surrogate_stats()
        P.set_surrogate(None)
        net.close()
    P.close()
    return out


def train_surrogate(args, deck):
    """Training set and training on the GPU (untimed setup of the surrogate
    variant): --sur-samples strain paths samples through the plain return
    mapping, --sur-steps Adam steps."""
    from oofem_amd.hipfem import Context, MisesSurrogate
    m = deck.mat[0]
    ctx = Context(0)
    net = MisesSurrogate(ctx, m[1], m[2], m[3], m[4], seed=args.sur_seed)
    t = time.perf_counter()
    net.generate(args.sur_samples, seed=args.sur_seed + 1)
    ctx.synchronize()
    t_gen = time.perf_counter() - t
    t = time.perf_counter()
    loss = net.train(args.sur_steps, batch=args.sur_batch, lr=args.sur_lr, lr_final=args.sur_lr * 1e-2,
                     seed=args.sur_seed + 2)
    t_train = time.perf_counter() - t
    params = net.params()
    net.close()
    ctx.close()
    log(f"[bench] surrogate: {args.sur_samples} samples in {t_gen:.1f} s, {args.sur_steps} steps of "
        f"{args.sur_batch} in {t_train:.1f} s, final batch loss {loss:.3e}")
    return params, dict(samples=args.sur_samples, steps=args.sur_steps, batch=args.sur_batch, gen_s=t_gen,
                        train_s=t_train, final_loss=loss)


def measured_traffic(kernel, workload, args):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    counter summary (profiles/pmc_traffic.json: FETCH_SIZE and WRITE_SIZE per
    dispatch, each scaled by its copy-kernel calibration factor), or None."""
    if args.traffic_gb is not None:
        return args.traffic_gb * 1e9
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        tab = json.load(fh)
    for rec in tab.get("kernels", []):
        if rec.get("kernel") == kernel and rec.get("workload") == workload:
            return rec.get("bytes_per_launch")
    return None


def cpu_baseline(args, deck_full, gpu_iters, nnz_full):
    """Oracle on a bounded slab of the same mesh family, scaled to the full step."""
    import numpy as np

    from oofem_amd import meshgen
    from oracle import oracle
    nzs = args.cpu_sample_nz
    s = meshgen.config1(nx=args.nx, ny=args.ny, nz=nzs) if args.workload == "cfg1" else meshgen.config0()
    t = time.perf_counter()
    rp, ci, val = oracle.csr(s)               # structure + assembly of the sample
    t_asm = time.perf_counter() - t
    t = time.perf_counter()
    oracle.internal_forces(s, np.zeros(s.neq), s.prescribed_values())
    t_fint = time.perf_counter() - t
    f, _ = s.external_load()
    nit = args.cpu_sample_iters
    t = time.perf_counter()
    oracle.pcg(rp, ci, val, f, tol=1e-30, maxit=nit)
    t_cg = (time.perf_counter() - t) / nit
    scale_e = deck_full.nelem / s.nelem
    scale_nnz = nnz_full / len(val)
    t_step = (t_asm + 2 * t_fint) * scale_e + gpu_iters * t_cg * scale_nnz
    return dict(value=deck_full.neq / t_step, unit="DOFs/s", cores=1, kind="port",
                sample=(f"oracle (C restatement, 1 thread) on a {args.nx}x{args.ny}x{nzs} slab of the same mesh "
                        f"family ({s.nelem} elements, {len(val)} nnz): K assembly {t_asm:.2f} s, f_int "
                        f"{t_fint:.2f} s, {nit} CG iterations at {t_cg * 1e3:.1f} ms each; scaled by elements "
                        f"x{scale_e:.1f} and nnz x{scale_nnz:.1f} to the full step with the GPU's "
                        f"{gpu_iters} iterations (estimated step {t_step:.1f} s)"))


def cpu_baseline_newton(args, deck, rn):
    """Oracle (one thread) on a slab of the same plasticity problem: tangent
    assembly, f_int with the return mapping and CG iterations timed there,
    scaled to the full step by element count, nnz and the GPU run's Newton
    and PCG iteration counts."""
    import numpy as np

    from oofem_amd import meshgen
    from oracle import oracle
    s = meshgen.config3(nx=args.nx, ny=args.ny, nz=args.cpu_sample_nz)
    st = oracle.state_new(s)
    up = s.prescribed_values(3.0 / args.c3_total)
    t = time.perf_counter()
    _, _, _ = oracle.internal_forces(s, np.zeros(s.neq), up, state=st, update_state=True)
    t_fint = time.perf_counter() - t
    t = time.perf_counter()
    rp, ci, val = oracle.csr(s, st, tangent=True)
    t_asm = time.perf_counter() - t
    f = np.ones(s.neq)
    nit = args.cpu_sample_iters
    t = time.perf_counter()
    oracle.pcg(rp, ci, val, f, tol=1e-30, maxit=nit)
    t_cg = (time.perf_counter() - t) / nit
    scale_e = deck.nelem / s.nelem
    scale_nnz = rn["nnz"] / len(val)
    newton = sum(rn["newton"]) / rn["steps"]
    pcg = sum(rn["pcg"]) / rn["steps"]
    # per Newton iteration: one tangent assembly and one f_int (+1 f_int per step)
    t_step = ((newton + 1) * t_fint + newton * t_asm) * scale_e + pcg * t_cg * scale_nnz
    return dict(value=deck.neq / t_step, unit="DOFs/s", cores=1, kind="port",
                sample=(f"oracle (C restatement, 1 thread) on a {args.nx}x{args.ny}x{args.cpu_sample_nz} slab of the "
                        f"same problem ({s.nelem} elements, {len(val)} nnz): f_int with return mapping {t_fint:.2f} s, "
                        f"tangent assembly {t_asm:.2f} s, {nit} CG iterations at {t_cg * 1e3:.1f} ms; scaled by "
                        f"elements x{scale_e:.1f}, nnz x{scale_nnz:.1f} and the GPU's {newton:.1f} Newton / "
                        f"{pcg:.0f} PCG iterations per step (estimated step {t_step:.1f} s)"))


def spmv_roofline(args, deck, r):
    """Roofline of the FP64 PCG SpMV of a Newton run: algorithmic bytes of the
    matrix's layout over the mean HIP-event launch time."""
    n = deck.neq
    roof = None
    if r["n_spmv"]:
        avg_ms = r["ms_spmv"] / r["n_spmv"]
        lay = r["layout"]
        if lay["form"] == 2:
            # row-group SpMV: values, one int32 column and one int32 offset per
            # chunk, group bounds (2 int32), row pointers, gather p, write q
            bytes_per = 8.0 * r["nnz"] + 8.0 * lay["blocks"] + 8.0 * lay["groups"] + 4.0 * n + 16.0 * n
            kname = "k_pcg_spmv_p_grp<double, double>"
        elif lay["form"] == 1:
            bytes_per = 8.0 * r["nnz"] + 4.0 * r["nnz"] / 9.0 + 4.0 * n / 3.0 + 16.0 * n
            kname = "k_pcg_spmv_p_nb<double, double>"
        else:
            bytes_per = 12.0 * r["nnz"] + 20.0 * n
            kname = "k_pcg_spmv_p<double, double> (lane form)"
        ach = bytes_per / (avg_ms * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4),
                    traffic=measured_traffic(kname, deck.name, args), kernel=kname, avg_launch_ms=round(avg_ms, 4),
                    launches=r["n_spmv"], algorithmic_bytes=int(bytes_per))
    return roof


def main_newton(args):
    """Newton-step throughput on the LSpace plasticity block (config 3):
    plain FP64 (hipcsr + hippcg FP64 + exact MisesMat) against the mixed
    solver and the mixed solver with the trained surrogate."""
    import numpy as np

    from oofem_amd import meshgen
    t = time.perf_counter()
    deck = meshgen.config3(nx=args.nx, ny=args.ny, nz=args.nz)
    log(f"[bench] deck {deck.name}: nelem={deck.nelem} neq={deck.neq} ({time.perf_counter() - t:.1f} s)")
    kw = dict(warm=args.c3_warm, steps=args.c3_steps, total=args.c3_total)
    runs = {"fp64": run_newton(args, deck, 0, keep_state=True, **kw)}
    if args.forcing > 0:
        runs["fp64_inexact"] = run_newton(args, deck, 0, keep_state=True, forcing=args.forcing, **kw)
    if args.precision in ("both", "mixed", "all"):
        runs["mixed"] = run_newton(args, deck, 2, **kw)
        if args.forcing > 0:
            runs["mixed_inexact"] = run_newton(args, deck, 2, keep_state=True, forcing=args.forcing, **kw)
    sur_info = None
    if args.precision in ("both", "all") and args.sur_steps > 0:
        params, sur_info = train_surrogate(args, deck)
        name = "mixed_surrogate_inexact" if args.forcing > 0 else "mixed_surrogate"
        runs[name] = run_newton(args, deck, 2, surrogate=(params, args.sur_band, args.sur_tol), keep_state=True,
                                forcing=args.forcing, **kw)
    base = runs["fp64"]
    n = deck.neq
    variants = {}
    for name, r in runs.items():
        v = dict(dofs_per_s=n * r["steps"] / r["dt"], ms_per_step=r["dt"] / r["steps"] * 1e3,
                 newton_iters=r["newton"], pcg_iters=r["pcg"])
        if name != "fp64":
            v["speedup_vs_fp64"] = v["dofs_per_s"] / (n * base["steps"] / base["dt"])
            v["u_rel_err_vs_fp64"] = float(np.linalg.norm(r["u"] - base["u"]) / np.linalg.norm(base["u"]))
            v["reliable_updates_last_solve"] = r["info"]["outer"]
            v["element_operator"] = r["info"]["matrix_free"]
            if r["n_spmv32"]:
                v["fp32_operator_avg_ms"] = r["ms_spmv32"] / r["n_spmv32"]
        if "surrogate" in r:
            v["surrogate"] = dict(r["surrogate"], band=args.sur_band, tol=args.sur_tol, training=sur_info)
        if "stress" in r and name != "fp64":
            sy = 250.0 + 2e3 * base["kappa"]
            e = np.abs(r["stress"] - base["stress"]).max(axis=1) / sy
            v["stress_err_vs_fp64"] = dict(max_rel=float(e.max()), rms_rel=float(np.sqrt((e * e).mean())))
        variants[name] = v
        log(f"[bench] {name}: {v}")
    r = base
    roof = spmv_roofline(args, deck, r)
    cpu = None if args.no_cpu_baseline else cpu_baseline_newton(args, deck, r)
    out = {
        "metric": METRIC, "value": n * r["steps"] / r["dt"], "unit": "DOFs/s", "n_gpus": 1, "steps": r["steps"],
        "warmup": args.c3_warm, "ms_per_step": r["dt"] / r["steps"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": deck.name, "elements": deck.nelem, "free_dofs": n, "nnz": r["nnz"],
                   "step": (f"StaticStructural load step {args.c3_warm + 1}..{args.c3_warm + args.c3_steps} of "
                            f"{args.c3_total} (prescribed end displacement, 5x yield strain in total): tangent "
                            "initial guess, full Newton with consistent tangent, rtolf 1e-8, Jacobi-PCG to "
                            f"{args.nl_lstol:g}"),
                   "precision": "fp64", "variants": variants},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def main_tet(args):
    """NonLinearStatic J2 plasticity on the 2.4M-element Kuhn-tet block
    (config 2): direct control with the tangent predictor (initialguess 1:
    without it the first step's boundary jump does not converge in 60
    Newton iterations), consistent tangent, rtolf 1e-8, FP64; --c2-warm
    untimed and --c2-steps timed load steps of the 20.  Size-independent
    checks: every step's force error, and the balance of the reactions (no
    external load: they sum to zero)."""
    import numpy as np

    from oofem_amd import meshgen
    t = time.perf_counter()
    deck = meshgen.config2()
    log(f"[bench] deck {deck.name}: nelem={deck.nelem} neq={deck.neq} ({time.perf_counter() - t:.1f} s)")
    r = run_newton(args, deck, 0, warm=args.c2_warm, steps=args.c2_steps, total=20, initial_guess=1)
    n = deck.neq
    R = r["reactions"]
    m = deck.eqn < 0
    node, comp = np.nonzero(m)
    Rz = np.zeros(3)
    top = np.zeros(3)
    for c in range(3):
        sel = comp == c
        Rz[c] = R[-deck.eqn[node[sel], c] - 1].sum()
    ztop = np.isclose(deck.coords[node, 2], deck.coords[:, 2].max()) & (comp == 2)
    top_force = float(R[-deck.eqn[node[ztop], 2] - 1].sum())
    roof = spmv_roofline(args, deck, r)
    out = {
        "metric": METRIC, "value": n * r["steps"] / r["dt"], "unit": "DOFs/s", "n_gpus": 1, "steps": r["steps"],
        "warmup": args.c2_warm, "ms_per_step": r["dt"] / r["steps"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": deck.name, "elements": deck.nelem, "free_dofs": n, "nnz": r["nnz"],
                   "step": (f"NonLinearStatic load steps {args.c2_warm + 1}..{args.c2_warm + args.c2_steps} of 20 "
                            "(prescribed end displacement, 5x yield strain in total): direct control, tangent "
                            "initial guess (initialguess 1), full Newton with consistent tangent, rtolf 1e-8, "
                            f"Jacobi-PCG to {args.nl_lstol:g}"),
                   "newton_iters": r["newton"], "pcg_iters": r["pcg"],
                   "checks": {"max_force_error": float(max(r["force_error"])),
                              "reaction_sum_xyz_over_end_force": [float(v / abs(top_force)) for v in Rz],
                              "end_force": top_force}},
        "roofline": roof, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def main_rve(args):
    """Batched RVE homogenisation (config 4): solves/s of the persistent
    mixed-precision batch, kernel shape from the autotuner, homogenised
    stiffness error against the oracle's FP64 solves on the smallest RVEs."""
    import numpy as np

    from oofem_amd import meshgen
    from oofem_amd.hipfem import Context
    from oofem_amd.rve import RVEBatch
    specs, macro = meshgen.config4(count=args.rve_count)
    ctx = Context(0)
    B = RVEBatch(ctx, specs, log=log)
    log(f"[bench] {len(specs)} RVEs in {len(B.chunks)} chunks, setup {B.t_setup:.1f} s")
    B.set_options(tol=args.rve_tol, split=args.rve_split, bulk_form=args.rve_bulk_form)
    t = time.perf_counter()
    tuned = B.autotune(sample=args.rve_tune_sample, log=log)
    t_tune = time.perf_counter() - t
    log(f"[bench] autotune ({t_tune:.1f} s): {tuned}")
    for i in range(args.warmup):
        B.run(log=log)
    B.set_profiling(True)
    runs = [B.run(log=log) for _ in range(args.steps)]
    prof = B.profile()
    forms = B.bulk_forms()
    t_solve = sum(r["t_solve"] for r in runs) / len(runs)
    t_rhs = sum(r["t_rhs"] for r in runs) / len(runs)
    t_hom = sum(r["t_homogenize"] for r in runs) / len(runs)
    out = runs[-1]
    nsolve = 6 * len(specs)
    # oracle FP64 on the smallest RVEs of the batch
    order = np.argsort([s[0] * (4 if s[1] == meshgen.QSPACE else 1) for s in specs])
    errs = []
    t = time.perf_counter()
    for i in order[:args.rve_check]:
        d = meshgen.rve(*specs[i])
        Cr = np.zeros((6, 6))
        for k in range(6):
            e = np.zeros(6)
            e[k] = 1.0
            from oracle import oracle
            r = oracle.linear_static(d, tol=1e-12, u_presc=meshgen.affine_values(d, e))
            Cr[:, k] = meshgen.homogenized_stress(d, r["reactions"])
        errs.append(float(np.abs(out["C"][i] - Cr).max() / np.abs(Cr).max()))
    t_check = time.perf_counter() - t
    sig = np.einsum("rij,rj->ri", out["C"], macro)
    dofs = sum(c.n for c in B.chunks)
    nq = sum(1 for s in specs if s[1] == meshgen.QSPACE)
    res = {
        "metric": METRIC, "value": nsolve / t_solve, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_solve * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-krylov/f64-residual", "data": "synthetic",
        "config": {"workload": f"rve_batch_{len(specs)}", "rves": len(specs), "qspace_rves": nq,
                   "free_dofs_total": dofs, "solves_per_step": nsolve, "tol": args.rve_tol,
                   "chunks": len(B.chunks), "setup_s": B.t_setup, "autotune": tuned, "autotune_s": t_tune,
                   "split": args.rve_split, "rhs_s": t_rhs, "reactions_s": t_hom,
                   "solves_per_s_with_rhs_and_reactions": nsolve / (t_solve + t_rhs + t_hom),
                   "iters_max": int(out["iters"].max()), "iters_mean": float(out["iters"].mean()),
                   "reliable_updates_mean": float(out["updates"].mean()),
                   "resid_max": float(np.abs(out["resid"]).max()), "all_converged": bool((out["resid"] > 0).all()),
                   "oracle_check": dict(rves=len(errs), C_rel_err_max=max(errs) if errs else None, seconds=t_check),
                   "macro_stress_norm_mean": float(np.linalg.norm(sig, axis=1).mean())},
        "roofline": None, "cpu_baseline": None,
    }
    if prof["launches"]:
        avg_ms = prof["ms"] / prof["launches"]
        ach = prof["bytes"] / (prof["ms"] * 1e-3) / 1e9
        kname = ("k_bulk_spmv_nbw<%d>" if all(f == 2 for f in forms) else
                 "k_bulk_spmv_nb<%d>" if all(forms) else "k_bulk_spmv<%d>")
        lpr = tuned[0]["bulk_lanes_per_row"] if tuned else 0
        res["roofline"] = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                               frac=round(ach / HBM_PEAK_GBS, 4), traffic=None, kernel=kname % lpr,
                               avg_launch_ms=round(avg_ms, 4), launches=prof["launches"],
                               algorithmic_bytes=int(prof["bytes"] / prof["launches"]))
    B.close()
    ctx.close()
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_rve(args, specs, t_solve)
    print(json.dumps(res), flush=True)
    return 0


def cpu_baseline_rve(args, specs, t_gpu):
    """Oracle (C restatement, one thread) on a seeded sample of the batch's
    RVEs: six FP64 Jacobi-PCG solves each to the batch tolerance, timed, and
    scaled to the whole batch by the sum over RVEs of stored entries times
    cube edge (CG iterations grow with the edge length), the sample's
    measured cost per unit of that product."""
    import numpy as np

    from oofem_amd import meshgen
    from oracle import oracle
    from oofem_amd.rve import NNZ_ROW
    rng = np.random.default_rng(7)
    # stored entries (the batch's own estimate) times the cube edge
    w = np.array([3.0 * (s[0] + 1) ** 3 * (4 if s[1] == meshgen.QSPACE else 1) * NNZ_ROW[s[1]] * s[0] for s in specs])
    cut = np.percentile(w, 25)
    order = rng.permutation(len(specs))
    t_cpu, w_done, done = 0.0, 0.0, []
    for i in order:
        if t_cpu > args.rve_cpu_seconds:
            break
        if w[i] > cut:
            continue
        d = meshgen.rve(*specs[i])
        t = time.perf_counter()
        for k in range(6):
            e = np.zeros(6)
            e[k] = 1.0
            oracle.linear_static(d, tol=args.rve_tol, u_presc=meshgen.affine_values(d, e))
        t_cpu += time.perf_counter() - t
        w_done += w[i]
        done.append(f"{'Q' if specs[i][1] == meshgen.QSPACE else 'L'}{specs[i][0]}")
    t_batch = t_cpu * w.sum() / w_done
    return dict(value=float(6 * len(specs) / t_batch), unit="solves/s", cores=1, kind="port",
                sample=(f"oracle FP64 Jacobi-PCG (1 thread), six solves each on {len(done)} RVEs of the batch "
                        f"({', '.join(done)}) in {t_cpu:.1f} s, scaled by stored entries x edge length "
                        f"(sum over the batch / sum over the sample = {w.sum() / w_done:.1f}); estimated "
                        f"batch {t_batch:.0f} s"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["cfg1", "cfg0", "cfg2", "cfg3", "cfg4"], default="cfg3")
    ap.add_argument("--c2-warm", type=int, default=1, help="untimed load steps of config 2")
    ap.add_argument("--c2-steps", type=int, default=2, help="timed load steps of config 2")
    ap.add_argument("--rve-count", type=int, default=512)
    ap.add_argument("--rve-tol", type=float, default=1e-8)
    ap.add_argument("--forcing", type=float, default=0.1,
                    help="inexact-Newton variants with forcing terms up to this value (0: none)")
    ap.add_argument("--rve-tune-sample", type=int, default=16)
    ap.add_argument("--rve-cpu-seconds", type=float, default=15.0,
                    help="CPU time budget of the oracle sample for the RVE cpu_baseline")
    ap.add_argument("--rve-bulk-form", type=int, default=1, choices=[0, 1, 2],
                    help="bulk products: 0 CSR rows, 1 node-block, 2 node-block from a padded block-major copy")
    ap.add_argument("--rve-split", type=int, default=-1,
                    help="stored entries above which an RVE goes to the bulk phase (-1 automatic, 0 all)")
    ap.add_argument("--rve-check", type=int, default=8)
    ap.add_argument("--c3-warm", type=int, default=1, help="untimed load steps before the timed ones (cfg3)")
    ap.add_argument("--c3-steps", type=int, default=1, help="timed load steps (cfg3)")
    ap.add_argument("--c3-total", type=int, default=8, help="load steps of the full ramp (cfg3)")
    ap.add_argument("--sur-samples", type=int, default=20_000_000)
    ap.add_argument("--sur-steps", type=int, default=20000)
    ap.add_argument("--sur-batch", type=int, default=4096)
    ap.add_argument("--sur-lr", type=float, default=2e-3)
    ap.add_argument("--sur-seed", type=int, default=31)
    ap.add_argument("--sur-band", type=float, default=0.05)
    ap.add_argument("--sur-tol", type=float, default=1e-4)
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--maxit", type=int, default=200000)
    ap.add_argument("--precision", choices=["fp64", "mixed", "ir", "both", "all"], default="both",
                    help="fp64: plain PCG; mixed: FP32 PCG with FP64 reliable updates (element-by-element "
                         "operator); ir: FP32 inner PCG in FP64 iterative refinement")
    ap.add_argument("--inner-tol", type=float, default=1e-4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-nz", type=int, default=16)
    ap.add_argument("--cpu-sample-iters", type=int, default=40)
    ap.add_argument("--shape", default="", help="SpMV shape: lanes:L:U[:nt] or stream:R:U")
    ap.add_argument("--mode", choices=["step", "kernels"], default="step")
    ap.add_argument("--iters", type=int, default=50, help="PCG iterations in --mode kernels")
    ap.add_argument("--newton-steps", type=int, default=2, help="timed Newton steps of the tet plasticity variant")
    ap.add_argument("--nl-total", type=int, default=20, help="load steps of the full ramp")
    ap.add_argument("--nl-lstol", type=float, default=1e-6, help="PCG tolerance inside Newton")
    ap.add_argument("--traffic-gb", type=float, default=None,
                    help="HBM bytes per SpMV launch from the calibrated FETCH_SIZE+WRITE_SIZE runs (GB)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU slab decomposition is a design, not code).", file=sys.stderr)
        return 2

    if args.workload == "cfg3" and args.mode == "step":
        return main_newton(args)
    if args.workload == "cfg4":
        return main_rve(args)
    if args.workload == "cfg2" and args.mode == "step":
        return main_tet(args)
    t = time.perf_counter()
    deck = make_deck(args)
    log(f"[bench] deck {deck.name}: nnode={deck.nnode} nelem={deck.nelem} neq={deck.neq} "
        f"({time.perf_counter() - t:.1f} s)")
    if args.mode == "kernels":
        return run_kernels(args, deck)
    modes = {"fp64": [0], "mixed": [2], "ir": [1], "both": [0, 2], "all": [0, 2, 1]}[args.precision]
    names = {0: "fp64", 1: "mixed_ir", 2: "mixed"}
    res = {}
    for prec in modes:
        name = names[prec]
        r = run_gpu(args, deck, prec)
        res[name] = r
        log(f"[bench] {name}: {args.steps} steps in {r['dt']:.3f} s, iters {r['iters']} "
            f"(outer/updates {r['outer']}, element operator {r['matrix_free']})")
    main_name = names[modes[0]]
    r = res[main_name]
    n, nnz = deck.neq, r["nnz"]
    ms_per_step = r["dt"] / args.steps * 1e3
    value = n * args.steps / r["dt"]
    roof = None
    if r["n_spmv"]:
        # node-block PCG SpMV: values 8 B/nnz, one int32 block column per 9
        # values, row pointers 4 B per node row, gather p and write q (16 B/row)
        avg_ms = r["ms_spmv"] / r["n_spmv"]
        bytes_per = 8.0 * nnz + 4.0 * nnz / 9.0 + 4.0 * n / 3.0 + 16.0 * n
        ach = bytes_per / (avg_ms * 1e-3) / 1e9
        kname = "k_pcg_spmv_p_nb<double, double>"
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=measured_traffic(kname, deck.name, args),
                    kernel=kname, avg_launch_ms=round(avg_ms, 4), launches=r["n_spmv"],
                    algorithmic_bytes=int(bytes_per))
    variants = {}
    for name, rv in res.items():
        variants[name] = dict(dofs_per_s=n * args.steps / rv["dt"], ms_per_step=rv["dt"] / args.steps * 1e3,
                              pcg_iters=rv["iters"], final_resid=rv["res"])
        if name != "fp64":
            variants[name]["outer_or_updates"] = rv["outer"]
            variants[name]["element_operator"] = rv["matrix_free"]
    if args.newton_steps > 0:
        from oofem_amd import meshgen
        nd = meshgen.config2()
        log(f"[bench] newton deck {nd.name}: nelem={nd.nelem} neq={nd.neq}")
        args.nl_steps = args.newton_steps
        for prec in ([0, 2] if 2 in modes else [0]):
            rn = run_newton(args, nd, prec)
            variants["newton_tets_" + names[prec]] = dict(
                workload=nd.name, free_dofs=nd.neq, steps=rn["steps"], dofs_per_s=nd.neq * rn["steps"] / rn["dt"],
                ms_per_step=rn["dt"] / rn["steps"] * 1e3, newton_iters=rn["newton"], pcg_iters=rn["pcg"],
                lstol=args.nl_lstol, spmv_avg_ms=rn["ms_spmv"] / max(rn["n_spmv"], 1) if rn["n_spmv"] else None,
                fp32_operator_avg_ms=rn["ms_spmv32"] / max(rn["n_spmv32"], 1) if rn["n_spmv32"] else None)
        if "newton_tets_mixed" in variants:
            variants["newton_tets_mixed"]["speedup_vs_fp64"] = (variants["newton_tets_mixed"]["dofs_per_s"] /
                                                                variants["newton_tets_fp64"]["dofs_per_s"])
    import numpy as np
    for name in ("mixed", "mixed_ir"):
        if "fp64" in res and name in res:
            a, b = res["fp64"], res[name]
            variants[name]["u_rel_err_vs_fp64"] = float(np.linalg.norm(b["u"] - a["u"]) / np.linalg.norm(a["u"]))
            variants[name]["reactions_rel_err_vs_fp64"] = float(
                np.linalg.norm(b["reactions"] - a["reactions"]) / np.linalg.norm(a["reactions"]))
            variants[name]["speedup_vs_fp64"] = variants[name]["dofs_per_s"] / variants["fp64"]["dofs_per_s"]
            if b["n_spmv32"]:
                variants[name]["fp32_operator_avg_ms"] = b["ms_spmv32"] / b["n_spmv32"]
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, deck, int(r["iters"][-1]), nnz)
        log(f"[bench] cpu baseline: {cpu['value']:.3e} DOFs/s")
    out = {
        "metric": METRIC, "value": value, "unit": "DOFs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if main_name == "fp64" else "f32-krylov/f64-residual",
        "data": "synthetic",
        "config": {"workload": deck.name, "elements": deck.nelem, "free_dofs": n, "nnz": nnz,
                   "step": "assemble K + rhs + Jacobi-PCG to 1e-8 + reactions", "tol": args.tol,
                   "precision": main_name, "variants": variants},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
