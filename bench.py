#!/usr/bin/env python3
"""Benchmark of the assembly-and-solve path on one MI355X.

Default workload (config 3): the 64x64x256 LSpace block (1,048,576 elements,
3,240,575 free equations, jitter +-0.2 h) with MisesMat (E 210e3, nu 0.3,
sig0 250, H 2e3), clamped at z = 0, u_z prescribed on z = 4 and ramped to
5x the yield strain over 20 load steps (more when --warmup + --steps
exceed 20).  One step = one StaticStructural load step: tangent initial
guess, full Newton with the consistent tangent to the 1e-8 force criterion,
Jacobi-PCG to 1e-6 per correction.  The plain FP64 form runs --warmup
untimed steps, then --steps timed ones; every other form of --variants runs
the same warm-up and the first --compare-steps of those timed steps:
  fp64                     hipcsr + hippcg FP64 + plain return mapping (value)
  mixed_surrogate          hippcg mixed 2 (FP32 PCG with FP64 reliable updates
                           and the FP32 element-by-element operator) with the
                           trained MisesMat surrogate (BF16 MFMA inference,
                           residual-checked, plain fallback)
  fp64_inexact, mixed_surrogate_inexact
                           the same two with Eisenstat-Walker inexact Newton
  mixed, mixed_inexact     (on request) the mixed form without the surrogate
value = free DOFs x timed steps / wall time of the FP64 form; every other
form's speed-up against the FP64 form with the same Newton forcing over the
same load steps, its errors against plain FP64, the errors of every form
against the FP64 oracle on a slab, and the surrogate's acceptance against
its tolerance are in config.

roofline: the FP64 PCG SpMV (row-group form on this pattern: the top face
constrains u_z only), algorithmic bytes per launch over its mean HIP-event
duration.  cpu_baseline: the oracle on the full mesh with the reference's
OpenMP element loops, timed on the host, with the GPU run's iteration
counts.

Other workloads: --workload cfg1 (LinearStatic cantilever, FP64 vs mixed),
cfg0, cfg2 (tet plasticity), cfg4 (512-RVE batch, solves/s); --mode kernels
(counter runs).
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
    plastic = bool(deck.mat[:, 0].max() >= 1) and args.kernels_plastic
    if plastic:
        # MisesMat tangent of a plastic state: the uniaxial field of 3/8 of
        # the end displacement (about 1.9x the yield strain) at every node
        up = deck.prescribed_values(3.0 / 8.0)
        zmax = deck.coords[:, 2].max()
        uz = deck.coords[:, 2] / zmax * 3.0 / 8.0 * deck.meta["u_end"]
        u = np.zeros(deck.neq)
        fe = deck.eqn[:, 2] > 0
        u[deck.eqn[fe, 2] - 1] = uz[fe]
        ev.internal_forces(u, up, update_state=True)
        s = ev.get_state()
        log(f"[bench] kernels: plastic Gauss points {(s[:, 13] > s[:, 6]).mean():.3f}")
        del s
    ev.assemble_stiffness(K, tangent=plastic)
    apply_shape(K, args.shape)
    f, _ = deck.external_load()
    if not f.any():
        f = np.ones(deck.neq)          # displacement-driven decks carry no load: any non-zero b will do
    db = DeviceVector(ctx, deck.neq, f)
    dx = DeviceVector(ctx, deck.neq)
    out = {}
    iter_ms = {}
    for prec, name in ((0, "fp64_nodeblock"), (1, "fp32_nodeblock"), (2, "fp32_element_operator")):
        s = HipPCGSolver(ctx, lstol=1e-30, lsiter=args.iters, lsprecond=1, precision=prec, inner_tol=1e-30)
        s.set_profiling(True)
        dx.fill(0.0)
        s.solve_dev(K, db.addr, dx.addr)
        p = s.profile()
        out[name] = (p["ms64"] / max(p["n64"], 1)) if prec == 0 else (p["ms32"] / max(p["n32"], 1))
        s.close()
    # whole PCG iterations (wall clock over --iters iterations): the plain
    # PCG with p updated in its own pass (fused 0) and formed inside the
    # SpMV gather (fused 1), and the mixed form with the slot-form element
    # operator (form 2)
    for prec, fl in ((0, (0, 1)), (2, (2,))):
        for flags in fl:
            s = HipPCGSolver(ctx, lstol=1e-30, lsiter=args.iters, lsprecond=1, precision=prec, inner_tol=1e-30)
            if prec == 0:
                s.set_fused(flags)
            else:
                ev.set_element_operator(flags)   # 2: the slot form
            dx.fill(0.0)
            s.solve_dev(K, db.addr, dx.addr)
            ctx.synchronize()
            dx.fill(0.0)
            ctx.synchronize()
            t = time.perf_counter()
            _, nit, _ = s.solve_dev(K, db.addr, dx.addr)
            ctx.synchronize()
            iter_ms[f"precision{prec}_{'fused' if prec == 0 else 'form'}{flags}"] = \
                (time.perf_counter() - t) * 1e3 / max(nit, 1)
            s.close()
    ev.set_element_operator(2)
    log(f"[bench] kernels: ms per PCG iteration {iter_ms}")
    print(json.dumps(dict(mode="kernels", workload=deck.name, plastic_tangent=plastic, neq=deck.neq, nnz=K.nnz(),
                          shape=K.spmv_shape(), copy_doubles=n_copy,
                          copy_gbs=copy_gbs, spmv_avg_ms=out, iteration_ms=iter_ms)), flush=True)
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
               initial_guess=1, keep_last_start=False, snap_step=None):
    """StaticStructural + NRSolver load steps (full Newton, consistent tangent,
    rtolf 1e-8, tangent initial guess) with the prescribed end displacement
    ramped over `total` steps: `warm` untimed steps, then `steps` timed ones.
    surrogate: (params, band, tol) of a trained MisesMat network, or None;
    forcing > 0: inexact Newton with forcing terms up to that value.
    snap_step: also keep the solution and Gauss-point stress at the end of
    that load step (the copies are taken out of the timed region)."""
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
    last_start = None
    snap = None
    t_excluded = 0.0
    for k in range(1, warm + steps + 1):
        if keep_last_start and k == warm + steps:
            # the committed Gauss-point state the last step starts from; the
            # device-to-host copy is taken out of the timed region
            t_copy = time.perf_counter()
            last_start = P.state()
            if t0 is not None:
                t_excluded += time.perf_counter() - t_copy
        P.set_prescribed(deck.prescribed_values(k / total))
        if k == warm + 1:
            P.set_profiling(True)
            if net is not None:
                P.surrogate_stats(reset=True)
            t0 = time.perf_counter()
        ts = time.perf_counter()
        rep = P.nonlinear_step()
        if rep["status"] != 1:
            raise RuntimeError(f"Newton step {k} failed: {rep}")
        P.update()
        rep["wall_s"] = time.perf_counter() - ts
        reports.append(rep)
        if k == snap_step:
            t_copy = time.perf_counter()
            gs = P.state()
            snap = dict(u=P.solution(), stress=gs[:, 20:26].copy(), kappa=gs[:, 6].copy())
            del gs
            t_excluded += time.perf_counter() - t_copy
        log(f"[bench] precision {precision}{' +surrogate' if net else ''}{' +forcing' if forcing else ''} "
            f"newton step {k}: {rep}")
    dt = time.perf_counter() - t0 - t_excluded
    p = P.profile()
    out = dict(dt=dt, steps=steps, newton=[r["newton"] for r in reports[warm:]], force_error=[r["force_error"] for r in reports],
               step_s=[round(r["wall_s"], 3) for r in reports[warm:]],
               reactions=P.reactions(),
               pcg=[r["pcg"] for r in reports[warm:]], n_spmv=p["n64"], ms_spmv=p["ms64"], n_spmv32=p["n32"],
               ms_spmv32=p["ms32"], nnz=P.matrix_nnz(), u=P.solution(), info=P.solver_info(),
               layout=P.matrix_layout())
    if keep_state:
        gs = P.state()
        out["stress"] = gs[:, 20:26].copy()
        out["kappa"] = gs[:, 6].copy()
        del gs
    if snap is not None:
        out["snap"] = snap
    if last_start is not None:
        out["last_start"] = last_start
        out["last_up"] = deck.prescribed_values((warm + steps) / total)
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
    """The CPU path on the full mesh of the plasticity block, timed on this
    host: the oracle (a C restatement of the reference's element loops,
    MisesMat, CompCol storage and IML CG) with its element loops in the
    reference's OpenMP form (parallel elements, critical-section scatter,
    engngm.C:902-923, 1375-1408; CG serial as IMLSolver runs it) on every
    host thread.  Timed: f_int with the return mapping, the consistent
    tangent assembled into the CSR pattern, and --cpu-cg-iters CG
    iterations on that matrix.  Only the iteration counts come from the GPU
    run (its Newton and PCG iterations per step, the same algorithm): step =
    (newton + 1) f_int + (newton + 1) assembly (the initial-guess tangent
    included) + pcg CG iterations.  The single-thread figure repeats the
    element loops serially on a slab of the same problem (scaled by element
    count, stated); skyline LDL^T on configuration 0 is the reference's
    direct solver."""
    import numpy as np

    from oofem_amd import meshgen
    from oracle import oracle
    nthr = oracle.omp_threads()
    t = time.perf_counter()
    rp, ci = oracle.csr_structure(deck)
    t_struct = time.perf_counter() - t
    log(f"[bench] cpu baseline: pattern {t_struct:.1f} s ({nthr} threads)")
    st = oracle.state_new(deck)
    up = deck.prescribed_values(3.0 / 8.0)   # 1.9x the yield strain: a plastic state
    zmax = deck.coords[:, 2].max()
    u = np.zeros(deck.neq)
    fe = deck.eqn[:, 2] > 0
    u[deck.eqn[fe, 2] - 1] = deck.coords[fe, 2] / zmax * up.max()
    t = time.perf_counter()
    oracle.internal_forces(deck, u, up, state=st, update_state=True, omp=True)
    t_fint = time.perf_counter() - t
    t = time.perf_counter()
    val = oracle.assemble(deck, rp, ci, st, tangent=True, omp=True)
    t_asm = time.perf_counter() - t
    log(f"[bench] cpu baseline: f_int {t_fint:.1f} s, tangent assembly {t_asm:.1f} s")
    nit = args.cpu_cg_iters
    t = time.perf_counter()
    oracle.pcg(rp, ci, val, np.ones(deck.neq), tol=1e-300, maxit=nit)
    t_cg = (time.perf_counter() - t) / nit
    log(f"[bench] cpu baseline: CG {t_cg * 1e3:.0f} ms per iteration")
    del val, rp, ci
    plastic = float((st.reshape(-1, 38)[:, 13] > st.reshape(-1, 38)[:, 6]).mean())
    del st
    newton = sum(rn["newton"]) / rn["steps"]
    pcg = sum(rn["pcg"]) / rn["steps"]
    t_step = (newton + 1) * (t_fint + t_asm) + pcg * t_cg
    # one thread: the element loops of a slab, scaled by element count
    s = meshgen.config3(nx=args.nx, ny=args.ny, nz=args.cpu_sample_nz)
    sst = oracle.state_new(s)
    sup = s.prescribed_values(3.0 / 8.0)
    su = np.zeros(s.neq)
    sfe = s.eqn[:, 2] > 0
    su[s.eqn[sfe, 2] - 1] = s.coords[sfe, 2] / s.coords[:, 2].max() * sup.max()
    srp, sci = oracle.csr_structure(s)
    t = time.perf_counter()
    oracle.internal_forces(s, su, sup, state=sst, update_state=True)
    t_fint1 = (time.perf_counter() - t) * deck.nelem / s.nelem
    t = time.perf_counter()
    oracle.assemble(s, srp, sci, sst, tangent=True)
    t_asm1 = (time.perf_counter() - t) * deck.nelem / s.nelem
    t_step1 = (newton + 1) * (t_fint1 + t_asm1) + pcg * t_cg
    log(f"[bench] cpu baseline: one thread f_int {t_fint1:.1f} s, assembly {t_asm1:.1f} s per full mesh")
    # configuration 0: LinearStatic with the reference's direct solver
    c0 = meshgen.config0()
    t = time.perf_counter()
    r0, c0i, v0 = oracle.csr(c0)
    f0, _ = oracle.surface_load(c0)
    oracle.internal_forces(c0, np.zeros(c0.neq), c0.prescribed_values())
    t0a = time.perf_counter() - t
    t = time.perf_counter()
    _, nwk = oracle.skyline_solve(r0, c0i, v0, f0)
    t0s = time.perf_counter() - t
    t = time.perf_counter()
    _, it0, _, _ = oracle.pcg(r0, c0i, v0, f0, tol=1e-10)
    t0c = time.perf_counter() - t
    return dict(value=deck.neq / t_step, unit="DOFs/s", cores=nthr, kind="port",
                sample=(f"oracle on the full {deck.name} mesh ({deck.nelem} elements, {deck.neq} free DOFs, "
                        f"{rn['nnz']} nnz), {nthr} host threads for the element loops (OpenMP, critical-section "
                        f"scatter) and one for CG: f_int with the return mapping {t_fint:.2f} s ({plastic:.2f} of "
                        f"the points plastic), tangent assembly {t_asm:.2f} s, CG {t_cg * 1e3:.0f} ms per "
                        f"iteration ({nit} timed); with the GPU run's {newton:.1f} Newton and {pcg:.0f} PCG "
                        f"iterations per step: {t_step:.1f} s per step (pattern build {t_struct:.1f} s, once)"),
                single_thread=dict(value=deck.neq / t_step1, cores=1,
                                   sample=(f"element loops on one thread on a {args.nx}x{args.ny}x"
                                           f"{args.cpu_sample_nz} slab ({s.nelem} elements) scaled by "
                                           f"x{deck.nelem / s.nelem:.0f} elements: f_int {t_fint1:.1f} s, assembly "
                                           f"{t_asm1:.1f} s per full mesh; CG as above; {t_step1:.1f} s per step")),
                config0=dict(workload=c0.name, free_dofs=c0.neq, assembly_s=t0a,
                             skyline_ldlt_s=t0s, skyline_profile_entries=int(nwk),
                             skyline_dofs_per_s=c0.neq / (t0a + t0s),
                             iml_cg_s=t0c, iml_cg_iters=int(it0), iml_cg_dofs_per_s=c0.neq / (t0a + t0c)))


def oracle_check_newton(args, variants):
    """Errors of every GPU form against the FP64 oracle on the largest slab
    of the plasticity block the oracle solves within the budget: the same
    load steps (ramp of --c3-total), the oracle's Newton with CG to 1e-12
    per correction; displacement and Gauss-point stress errors per form."""
    import numpy as np

    from oofem_amd import meshgen
    from oracle import oracle
    d = meshgen.config3(nx=args.check_nx, ny=args.check_nx, nz=4 * args.check_nx)
    nsteps = args.check_steps
    t = time.perf_counter()
    st = oracle.state_new(d)
    X = np.zeros(d.neq)
    prev = None
    for k in range(1, nsteps + 1):
        up = d.prescribed_values(k / args.c3_total)
        r = oracle.nonlinear_step_ex(d, st, X, up, u_presc_prev=prev, rtolf=1e-8, maxiter=60, lstol=1e-12)
        if r["status"] != 1:
            raise RuntimeError(f"oracle step {k} failed: {r}")
        oracle.state_commit(d, st)
        prev = up
        log(f"[bench] oracle check: step {k} ({time.perf_counter() - t:.1f} s)")
    t_oracle = time.perf_counter() - t
    so = st.reshape(-1, 38)
    sy = 250.0 + 2e3 * so[:, 6]
    out = dict(workload=d.name, free_dofs=d.neq, load_steps=nsteps, oracle_s=round(t_oracle, 1), forms={})
    for name, (precision, forcing, sur) in variants.items():
        sub = argparse.Namespace(**vars(args))
        sub.nl_steps = nsteps
        r = run_newton(sub, d, precision, warm=0, steps=nsteps, total=args.c3_total, surrogate=sur, keep_state=True,
                       forcing=forcing)
        out["forms"][name] = dict(
            u_rel_err=float(np.linalg.norm(r["u"] - X) / np.linalg.norm(X)),
            stress_err_max_over_sigma_y=float((np.abs(r["stress"] - so[:, 20:26]).max(axis=1) / sy).max()))
    return out


def surrogate_curve(args, deck, params, start_state, X, up, tols):
    """Accepted share of the queued Gauss points against the acceptance
    tolerance, at the last timed step's material update of the plain FP64
    run (its committed start state, its converged displacements), with the
    time of the update and the stress error of the accepted points against
    the plain return mapping."""
    import numpy as np

    from oofem_amd.hipfem import BatchedEvaluator, Context, HipCSRSparseMtrx, MisesSurrogate
    ctx = Context(0)
    ev = BatchedEvaluator(ctx, deck)
    K = HipCSRSparseMtrx(ctx)
    ev.build_structure(K)
    m = deck.mat[0]
    net = MisesSurrogate(ctx, m[1], m[2], m[3], m[4])
    net.set_params(params)

    def update(use, tol):
        ev.set_state(start_state)
        ev.restart()
        ev.set_surrogate(net if use else None, band=args.sur_band, tol=tol)
        if use:
            ev.surrogate_stats(reset=True)
        ctx.synchronize()
        t = time.perf_counter()
        ev.internal_forces(X, up, update_state=True)
        ctx.synchronize()
        dt = time.perf_counter() - t
        return dt, ev.get_state()[:, 20:26].copy(), (ev.surrogate_stats() if use else None)

    t_plain, s_plain, _ = update(False, 0.0)
    sy = 250.0 + 2e3 * start_state.reshape(-1, 38)[:, 6]
    curve = []
    for tol in tols:
        dt, s, stt = update(True, tol)
        err = np.abs(s - s_plain).max(axis=1) / sy
        curve.append(dict(tol=tol, queued=int(stt["queued"]), accepted=int(stt["accepted"]),
                          accepted_share=stt["accepted"] / max(stt["queued"], 1),
                          max_stress_err_over_sigma_y=float(err.max()), update_s=round(dt, 4)))
    ev.set_surrogate(None)
    net.close()
    K.close()
    ev.close()
    ctx.close()
    return dict(plain_update_s=round(t_plain, 4), points=int(s_plain.shape[0]), curve=curve)


def spmv_roofline(args, deck, r):
    """Roofline of the FP64 PCG SpMV of a Newton run: algorithmic bytes of the
    matrix's layout over the mean HIP-event launch time."""
    n = deck.neq
    roof = None
    if r["n_spmv"]:
        avg_ms = r["ms_spmv"] / r["n_spmv"]
        lay = r["layout"]
        if lay["form"] == 2:
            # row-group SpMV: values, one int32 column (with the chunk length)
            # per chunk, group bounds (2 int32), row pointers, gather p, write q
            bytes_per = 8.0 * r["nnz"] + 4.0 * lay["blocks"] + 8.0 * lay["groups"] + 4.0 * n + 16.0 * n
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
    """Newton-step throughput on the LSpace plasticity block (config 3).
    The plain FP64 form (hipcsr + hippcg FP64 + MisesMat: the reference's
    algorithm, whose rate is `value`) runs --warmup untimed and --steps timed
    load steps of a ramp of max(--c3-total, warmup + steps) steps.  Every
    other form of --variants runs the same warm-up and the first
    --compare-steps of those timed steps (a plastic load step of this block
    costs about 20 s in FP64, so every form over every requested step would
    hold the GPU for many minutes):
      mixed                    FP32 PCG with FP64 reliable updates, FP32 element operator
      mixed_surrogate          the same with the trained MisesMat surrogate
      fp64_inexact             plain FP64 with Eisenstat-Walker inexact Newton
      mixed_surrogate_inexact  mixed + surrogate with inexact Newton
    A form's speed-up is its time over the compared load steps against the
    time of the FP64 form with the same Newton forcing over the same load
    steps from the same start (speedup_same_newton), and against plain FP64
    (speedup_vs_plain_fp64); its displacement and stress errors are taken
    against plain FP64 at the end of the compared steps."""
    import numpy as np

    from oofem_amd import meshgen
    t = time.perf_counter()
    deck = meshgen.config3(nx=args.nx, ny=args.ny, nz=args.nz)
    log(f"[bench] deck {deck.name}: nelem={deck.nelem} neq={deck.neq} ({time.perf_counter() - t:.1f} s)")
    total = max(args.c3_total, args.warmup + args.steps)
    ncmp = max(1, min(args.compare_steps, args.steps))
    want = [v for v in args.variants.split(",") if v]
    if "fp64" not in want:
        want.insert(0, "fp64")
    params, sur_info = None, None
    if any("surrogate" in v for v in want):
        params, sur_info = train_surrogate(args, deck)
    sur = (params, args.sur_band, args.sur_tol) if params is not None else None
    forms = {"fp64": (0, 0.0, None), "mixed": (2, 0.0, None), "mixed_surrogate": (2, 0.0, sur),
             "fp64_inexact": (0, args.forcing, None), "mixed_inexact": (2, args.forcing, None),
             "mixed_surrogate_inexact": (2, args.forcing, sur)}
    runs = {}
    for name in want:
        precision, forcing, su = forms[name]
        if name == "fp64":
            runs[name] = run_newton(args, deck, precision, warm=args.warmup, steps=args.steps, total=total,
                                    keep_state=True, forcing=forcing, surrogate=su,
                                    keep_last_start=sur is not None, snap_step=args.warmup + ncmp)
        else:
            runs[name] = run_newton(args, deck, precision, warm=args.warmup, steps=ncmp, total=total,
                                    keep_state=True, forcing=forcing, surrogate=su)
    base = runs["fp64"]
    n = deck.neq
    # the plain FP64 reference over the compared steps, and its state there
    ref_cmp = dict(dt=sum(base["step_s"][:ncmp]), u=base["snap"]["u"], stress=base["snap"]["stress"],
                   kappa=base["snap"]["kappa"])
    cmp_dt = {name: (ref_cmp["dt"] if name == "fp64" else sum(r["step_s"])) for name, r in runs.items()}
    variants = {}
    for name, r in runs.items():
        if name == "fp64":
            v = dict(dofs_per_s=n * r["steps"] / r["dt"], ms_per_step=r["dt"] / r["steps"] * 1e3,
                     newton_iters=r["newton"], pcg_iters=r["pcg"], step_s=r["step_s"],
                     compared_steps_s=ref_cmp["dt"])
        else:
            v = dict(dofs_per_s=n * r["steps"] / r["dt"], ms_per_step=r["dt"] / r["steps"] * 1e3,
                     newton_iters=r["newton"], pcg_iters=r["pcg"], step_s=r["step_s"])
            same = "fp64_inexact" if name.endswith("inexact") else "fp64"
            if same in cmp_dt:
                v["speedup_same_newton"] = cmp_dt[same] / cmp_dt[name]
                v["same_newton_baseline"] = same
            v["speedup_vs_plain_fp64"] = cmp_dt["fp64"] / cmp_dt[name]
            v["u_rel_err_vs_fp64_gpu"] = float(np.linalg.norm(r["u"] - ref_cmp["u"]) / np.linalg.norm(ref_cmp["u"]))
            v["reliable_updates_last_solve"] = r["info"]["outer"]
            v["element_operator"] = r["info"]["matrix_free"]
            if r["n_spmv32"]:
                v["fp32_operator_avg_ms"] = r["ms_spmv32"] / r["n_spmv32"]
            sy = 250.0 + 2e3 * ref_cmp["kappa"]
            e = np.abs(r["stress"] - ref_cmp["stress"]).max(axis=1) / sy
            v["stress_err_vs_fp64_gpu"] = dict(max_rel=float(e.max()), rms_rel=float(np.sqrt((e * e).mean())))
        if "surrogate" in r:
            v["surrogate"] = dict(r["surrogate"], band=args.sur_band, tol=args.sur_tol, training=sur_info)
        variants[name] = v
        log(f"[bench] {name}: {v}")
    extra = {"compared_load_steps": [args.warmup + 1, args.warmup + ncmp]}
    if sur is not None and "last_start" in base:
        extra["surrogate_acceptance"] = surrogate_curve(args, deck, params, base["last_start"], base["u"],
                                                        base["last_up"], [1e-4, 1e-3, 1e-2, 3e-2])
        log(f"[bench] surrogate acceptance: {extra['surrogate_acceptance']}")
    if args.check_nx > 0:
        extra["oracle_check"] = oracle_check_newton(args, {name: forms[name] for name in want})
        log(f"[bench] oracle check: {extra['oracle_check']}")
    r = base
    roof = spmv_roofline(args, deck, r)
    for k in ("last_start", "snap"):
        base.pop(k, None)
    cpu = None if args.no_cpu_baseline else cpu_baseline_newton(args, deck, r)
    out = {
        "metric": METRIC, "value": n * r["steps"] / r["dt"], "unit": "DOFs/s", "n_gpus": 1, "steps": r["steps"],
        "warmup": args.warmup, "ms_per_step": r["dt"] / r["steps"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": deck.name, "elements": deck.nelem, "free_dofs": n, "nnz": r["nnz"],
                   "step": (f"StaticStructural load step of a {total}-step ramp of the prescribed end displacement "
                            f"(5x the yield strain over 20 steps); steps {args.warmup + 1}.."
                            f"{args.warmup + args.steps} timed: tangent initial guess, full Newton with consistent "
                            f"tangent, rtolf 1e-8, Jacobi-PCG to {args.nl_lstol:g}"),
                   "precision": "fp64", "variants": variants, **extra},
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
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=4,
                    help="untimed load steps (cfg3: the first four of the 20-step ramp reach the yield strain)")
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
    ap.add_argument("--c3-total", type=int, default=20,
                    help="load steps of the ramp (cfg3; at least warmup + steps)")
    ap.add_argument("--compare-steps", type=int, default=2,
                    help="cfg3: timed load steps of every form other than plain FP64 (the first of its timed steps)")
    ap.add_argument("--variants", default="fp64,mixed_surrogate,fp64_inexact,mixed_surrogate_inexact",
                    help="cfg3 forms to run (comma separated)")
    ap.add_argument("--check-nx", type=int, default=10,
                    help="cfg3 oracle check on an n x n x 4n slab (0: off)")
    ap.add_argument("--check-steps", type=int, default=6, help="load steps of the oracle check")
    ap.add_argument("--cpu-cg-iters", type=int, default=5, help="CG iterations timed on the full mesh")
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
    ap.add_argument("--kernels-plastic", type=int, default=1,
                    help="--mode kernels on cfg3: assemble the MisesMat tangent of a plastic state (1) or K (0)")
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
