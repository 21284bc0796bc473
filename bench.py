#!/usr/bin/env python3
"""Benchmark of the matrix-free p = 2 elasticity hot path on one MI355X.

Workload: the structured SIMP cantilever on a 96x48x48 p = 2 B-spline grid
(221,184 elements, 98x50x50x3 = 735,000 DOFs) over the box [0,2]x[0,1]x[0,1],
E0 = 1, nu = 0.3, densities i.i.d. U[0.01, 1] (seeded), penalty 3,
Emin = 1e-9, x = 0 clamped, unit downward traction on x = 2.

A step is one masked K.u apply (FP64 MFMA contraction) over the whole grid
with x and y resident in HBM; value = applies per second.  The same JSON
line carries the FP32 apply rate, the PCG time-to-solution (true relative
residual 1e-8, FP64 operator) of plain FP64 PCG and of the mixed solver
(FP32 PCG with FP64 reliable updates) with the errors of the latter against
the former, the
roofline of the FP64 apply kernel, and the CPU baseline (the oracle's
restatement of the reference's rayon CSR SpMV, on this host's cores).

Profiling modes: --apply-only times only the apply in --apply-prec;
--copy-calibration runs a plain device copy of 735,000 doubles 20 times, the
known-size kernel against which the FETCH_SIZE / WRITE_SIZE counters are
scaled.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Ku applies/s + PCG time-to-1e-8 at 735k DOF: mixed-prec vs FP64, 1 GPU"
NEL = (96, 48, 48)
BOUNDS = ((0.0, 2.0), (0.0, 1.0), (0.0, 1.0))
SEED = 20261203
PEAK_F64 = 78.6e12   # MI355X dense FP64 MFMA, FLOP/s
PEAK_F32 = 157.3e12  # MI355X dense FP32, FLOP/s
PEAK_BF16 = 2.5e15   # MI355X dense BF16 MFMA, FLOP/s
PEAK_HBM = 8.0e12    # B/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def problem(np):
    rng = np.random.default_rng(SEED)
    ncp = [n + 2 for n in NEL]
    nel = NEL[0] * NEL[1] * NEL[2]
    rho = rng.uniform(0.01, 1.0, nel)
    scale = 1e-9 + (1.0 - 1e-9) * rho ** 3
    nsc = ncp[0] * ncp[1] * ncp[2]
    i = np.arange(nsc) % ncp[0]
    free = np.ones(3 * nsc, dtype=np.uint8)
    for c in range(3):
        free[c * nsc:(c + 1) * nsc][i == 0] = 0
    x = rng.standard_normal(3 * nsc)
    x /= np.linalg.norm(x)
    return rho, scale, free, x


def traffic_from_profile():
    """HBM bytes per FP64 apply launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated against a copy kernel)."""
    path = os.path.join(ROOT, "profiles", "apply_f64_96x48x48_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def time_applies(torch, icc, ctx, op, x, y, prec, steps, warmup, accumulate=False):
    stream = torch.cuda.ExternalStream(ctx.stream)
    fn = op.apply_accumulate if accumulate else (lambda xx, yy, prec: op.apply(xx, prec=prec, out=yy))
    for _ in range(warmup):
        fn(x, y, prec=prec)
    ctx.sync()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn(x, y, prec=prec)
    e1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps  # ms


def op_apply_atomic(op, x):
    """The operator's apply with the atomic tile scatter (deterministic mode off)."""
    op.set_deterministic(False)
    y = op.apply(x)
    op.set_deterministic(True)
    return y


def pcg_c0(np, torch, icc, ctx):
    """FP64 Jacobi PCG to 1e-8 on the 16x8x8 grid (rho = 0.5, penal 3,
    Emin 1e-9, x = 0 clamped, unit downward traction on x = 2)."""
    nel0 = (16, 8, 8)
    g = icc.Grid(ctx, nel0, BOUNDS)
    ncp = [n + 2 for n in nel0]
    nsc = ncp[0] * ncp[1] * ncp[2]
    fr = np.ones(3 * nsc, dtype=np.uint8)
    for c in range(3):
        fr[c * nsc:(c + 1) * nsc][np.arange(nsc) % ncp[0] == 0] = 0
    op = icc.Operator(g, scale=np.full(g.nel_tot, 1e-9 + (1 - 1e-9) * 0.5 ** 3), free=fr)
    t = icc.CantileverTopOpt(ctx, icc.CantileverTopOptConfig(nel=nel0, size=(2.0, 1.0, 1.0), load=1, emin=1e-9))
    f = t.force * fr
    t.close()
    fd = torch.from_numpy(f).cuda()
    u = torch.zeros(g.ndof, dtype=torch.float64, device="cuda")
    stats = np.zeros(8)
    icc.lib().icc_solve_ir_dev(op.h, icc.ICC_F64, fd.data_ptr(), 1e-8, 1e-8, 200, 20 * g.ndof, u.data_ptr(),
                               stats.ctypes.data)  # warm-up (graph capture)
    ctx.sync()
    u.zero_()
    t0 = time.time()
    rc = icc.lib().icc_solve_ir_dev(op.h, icc.ICC_F64, fd.data_ptr(), 1e-8, 1e-8, 200, 20 * g.ndof, u.data_ptr(),
                                    stats.ctypes.data)
    ctx.sync()
    return dict(time_s=time.time() - t0, status=int(rc), pcg_iterations=int(stats[1]),
                true_rel_residual=float(stats[2]))


def cpu_baseline(np, scale, x, budget_s):
    """The oracle's restatement of the reference CPU K.u (row-parallel CSR
    SpMV, sparse.rs:60-68, OpenMP for rayon) at the full 735k-DOF size."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    sp = O.Space([list(b) for b in BOUNDS], list(NEL), 2)
    lam, mu = O.lame(1.0, 0.3)
    sp.elem_mats(lam, mu)
    t0 = time.time()
    nnz = sp.assemble_structured(scale)
    t_asm = time.time() - t0
    y = sp.matvec(x)  # warm-up
    reps = 0
    t0 = time.time()
    while True:
        y = sp.matvec(x)
        reps += 1
        if time.time() - t0 > budget_s:
            break
    dt = (time.time() - t0) / reps
    cores = int(O.lib().orc_num_threads())
    out = dict(value=1.0 / dt, unit="applies/s", cores=cores, kind="port",
               sample=f"{reps} CSR SpMVs (nnz {nnz}) of the 96x48x48 operator, row-parallel over {cores} "
                      f"OpenMP threads; CSR built in {t_asm:.1f}s (not timed)", csr_build_s=t_asm)
    # the same SpMV on one thread (a few products)
    O.lib().orc_set_threads(1)
    t0 = time.time()
    n1 = 0
    while n1 < 2 or time.time() - t0 < 3.0:
        sp.matvec(x)
        n1 += 1
    out["one_thread_applies_per_s"] = n1 / (time.time() - t0)
    O.lib().orc_set_threads(cores)
    # Jacobi PCG (solver.rs) on the same operator: a bounded sample of 100
    # iterations, free DOFs only, traction load
    return out, y, sp, O


def cpu_pcg(np, O, sp, free, force, gpu_iters):
    """The oracle's restatement of PcgSolver::solve (serial vector loops,
    the OpenMP CSR SpMV) on the 96x48x48 system: 100 iterations timed, the
    time to 1e-8 extrapolated with the FP64 GPU iteration count; and the
    whole solve to 1e-8 on the 16x8x8 grid (rho = 0.5)."""
    b = force * free
    d = np.maximum(sp.diag(), 1e-12)
    t0 = time.time()
    _, its, _ = sp.pcg(b, d, free, 1e-30, 100)
    per_it = (time.time() - t0) / max(its, 1)
    out = dict(c1_ms_per_iteration=per_it * 1e3, c1_sample_iterations=its,
               c1_time_to_1e8_s_extrapolated=per_it * gpu_iters, c1_gpu_fp64_iterations=gpu_iters)
    nel0 = [16, 8, 8]
    sp0 = O.Space([list(bb) for bb in BOUNDS], nel0, 2)
    sp0.elem_mats(*O.lame(1.0, 0.3))
    sp0.assemble_structured(np.full(int(np.prod(nel0)), 1e-9 + (1 - 1e-9) * 0.5 ** 3))
    ncp = [n + 2 for n in nel0]
    nsc = ncp[0] * ncp[1] * ncp[2]
    fr0 = np.ones(3 * nsc, dtype=np.uint8)
    for c in range(3):
        fr0[c * nsc:(c + 1) * nsc][np.arange(nsc) % ncp[0] == 0] = 0
    f0 = sp0.traction_xmax([0.0, -1.0, 0.0]) * fr0
    d0 = np.maximum(sp0.diag(), 1e-12)
    t0 = time.time()
    _, its0, res0 = sp0.pcg(f0, d0, fr0, 1e-8, 20 * 3 * nsc)
    out.update(c0_time_to_1e8_s=time.time() - t0, c0_iterations=its0, c0_residual=res0)
    return out


def cpu_imm_topopt(np, icc):
    """The oracle's restatement of ImmersedTopOpt::run (assembled CSR, Jacobi
    PCG, OpenMP) for the first design iteration of the default 28x20x16
    tilted-box problem."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    mesh = tilted_box(np, icc)
    t0 = time.time()
    comp, _, _, _, its = O.immersed_topopt_run(mesh.vertices, mesh.triangles, (28, 20, 16), 1, 0.35, 3.0, 1.8, 1e-2,
                                               emin=1e-4, pcg_tol=1e-8, stab_kind=2, full=True)
    dt = time.time() - t0
    return dict(s_per_iteration=dt, pcg_iterations=int(its[0]), compliance=float(comp[0]),
                cores=int(O.lib().orc_num_threads()), kind="port",
                sample="first design iteration of the 28x20x16 tilted-box immersed SIMP run (set-up included)")


def immersed_leg(np, icc, ctx, seed=20261204):
    """Immersed solve on a 128x64x64 background grid over [0,2]x[0,1]^2 minus
    three seeded z-cylinders (radius U[0.08, 0.15]); cells with volume fraction
    < 0.05 dropped, cut cells scaled by w; penalty u = 0 on x = 0 (beta =
    1e3 E / h_min); unit downward traction on x = 2; FP64 PCG and the mixed
    solver (FP32 and split-FP16 inner operators), all to a relative residual
    of 1e-8."""
    rng = np.random.default_rng(seed)
    cyl = np.column_stack([rng.uniform(0.4, 1.6, 3), rng.uniform(0.3, 0.7, 3), rng.uniform(0.08, 0.15, 3)])
    res, sol = {}, {}
    for name, solver in (("fp64", 0), ("mixed_fp32", 1), ("mixed_f16x", 2)):
        cfg = icc.ImmersedConfig(res=(128, 64, 64), mode=1, solver=solver, pcg_tol=1e-8)
        t0 = time.time()
        h = icc.ImmersedSdf(ctx, cfg, cyl)
        t_build = time.time() - t0
        t0 = time.time()
        u, st, rc = h.solve()
        dt = time.time() - t0
        res[name] = dict(build_s=t_build, solve_s=dt, status=int(rc), iterations=st["iterations"],
                         true_rel_residual=st["true_res"], compliance=st["compliance"])
        sol[name] = (u, st["compliance"])
        res["problem"] = dict(nel=h.nel, ndof=h.ndof, active_cells=h.n_active, free_dofs=h.n_free,
                              penalty_nnz=h.nnz_penalty, beta=h.beta, cylinders=cyl.round(6).tolist())
        h.close()
    uf, cf = sol["fp64"]
    for name in ("mixed_fp32", "mixed_f16x"):
        um, cm = sol[name]
        res[name].update(u_rel_err_vs_fp64=float(np.linalg.norm(um - uf) / np.linalg.norm(uf)),
                         compliance_rel_err_vs_fp64=abs(cm - cf) / abs(cf))
        res["speedup_" + name + "_vs_fp64"] = res["fp64"]["solve_s"] / res[name]["solve_s"]
    res["speedup_mixed_vs_fp64"] = max(res["speedup_mixed_fp32_vs_fp64"], res["speedup_mixed_f16x_vs_fp64"])
    return res


def batch_leg(np, icc, ctx, seed=20261205):
    """24 SIMP cantilevers of uneven sizes, nx_k = 24 + round(136 k / 23) made
    even, ny = nz = nx / 2, densities U[0.01, 1], penalty 3, Emin 1e-9, own
    traction magnitude U[0.5, 2]; all instances solved in one graph-batched
    mixed solve to a true relative residual of 1e-8, against the same
    instances solved one after another."""
    nx = [2 * int(round((24 + round(136 * k / 23)) / 2)) for k in range(24)]
    nel = [(n, n // 2, n // 2) for n in nx]
    rng = np.random.default_rng(seed)
    setups = []
    for k in range(24):
        ne = nel[k][0] * nel[k][1] * nel[k][2]
        rho = rng.uniform(0.01, 1.0, ne)
        setups.append((1e-9 + (1 - 1e-9) * rho ** 3, rng.uniform(0.5, 2.0)))
    b = icc.Batch(ctx, nel)
    for k, (s, mag) in enumerate(setups):
        b.set(k, scale=s)
        b.set_traction(k, (0.0, -mag, 0.0))
    t0 = time.time()
    b.autotune(icc.ICC_F32, reps=3)  # tile shapes per instance size
    t_tune = time.time() - t0
    ctx.sync()
    t0 = time.time()
    us, st, rc = b.solve(prec=icc.ICC_F32, tol=1e-8, delta=0.1)
    dt = time.time() - t0
    ndof = [int(v) for v in b.ndof]
    b.close()
    seq = 0.0
    for k, (s, mag) in enumerate(setups):
        b1 = icc.Batch(ctx, [nel[k]])
        b1.set(0, scale=s)
        b1.set_traction(0, (0.0, -mag, 0.0))
        b1.autotune(icc.ICC_F32, reps=3)
        ctx.sync()
        t0 = time.time()
        b1.solve(prec=icc.ICC_F32, tol=1e-8, delta=0.1)
        seq += time.time() - t0
        b1.close()
    return dict(instances=len(nel), nx=nx, ndof=ndof, total_ndof=int(sum(ndof)), time_s=dt, status=int(rc),
                iterations=[int(v) for v in st[:, 0]], max_true_rel_residual=float(st[:, 2].max()),
                sequential_time_s=seq, speedup_vs_sequential=seq / dt, autotune_s=t_tune)


def adjoint_leg(np, icc, ctx, oc_iters=20, nfd=32, seed=20261206):
    """Adjoint SIMP at 128x64x64 (1,698,840 DOF): density filter radius 1.5
    elements, unit downward traction on x = 2, volume fraction 0.3, 20
    optimality-criteria iterations (FP64 PCG to 1e-8); then dc/dx at seeded
    densities U[0.2, 1] against central differences of the compliance on 32
    seeded elements, h = 1e-6, every solve to 1e-12.

    c(x') - c(x) is evaluated without cancelling against c: with u the state
    of x, u' = u + d that of x' = x +- h e and K' the operator of x',
        c(x') - c(x) = -sum_f (s'_f - s_f) ee_f(u) + 2 d.(F - K' u) - d.K' d
    (exact for exact states; the error of the variational form is quadratic in
    the solve residual), so the difference keeps its own relative precision."""
    import math
    base = dict(nel=(128, 64, 64), size=(2.0, 1.0, 1.0), volfrac=0.3, penal=3.0, rmin=1.5, emin=1e-9, load=1,
                filter=1)
    t = icc.CantileverTopOpt(ctx, icc.CantileverTopOptConfig(max_iter=oc_iters, pcg_tol=1e-8, **base))
    t0 = time.time()
    h = t.run()
    dt = time.time() - t0
    out = dict(oc_iterations=len(h.compliance), time_s=dt, s_per_oc_iteration=dt / max(1, len(h.compliance)),
               compliance_first=float(h.compliance[0]), compliance_last=float(h.compliance[-1]),
               pcg_iterations_total=int(sum(h.pcg_iterations)))
    t.close()
    t0 = time.time()
    tf = icc.CantileverTopOpt(ctx, icc.CantileverTopOptConfig(max_iter=1, pcg_tol=1e-12, **base))
    grid = icc.Grid(ctx, base["nel"], [[0, 2.0], [0, 1.0], [0, 1.0]])
    flt = icc.Filter(grid, base["rmin"])
    op = icc.Operator(grid, free=tf.free)
    emin, pen = base["emin"], base["penal"]
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.2, 1.0, tf.nel)
    c, g, u = tf.compliance(x)
    s0 = emin + (1 - emin) * flt.fwd(x) ** pen
    ee0 = tf.element_energies(u)

    def dcomp(xd):
        sd = emin + (1 - emin) * flt.fwd(xd) ** pen
        _, _, ud = tf.compliance(xd, u0=u, grad=False)
        d = ud - u
        op.set_scale(sd)
        ku, kd = op.apply(u), op.apply(d)
        return (-math.fsum((sd - s0) * ee0) + 2.0 * math.fsum(d * (tf.force - ku)) - math.fsum(d * kd))

    hstep = 1e-6
    errs = []
    for e in rng.choice(tf.nel, nfd, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[e] += hstep
        xm[e] -= hstep
        fd = (dcomp(xp) - dcomp(xm)) / (2 * hstep)
        errs.append(abs(g[e] - fd) / abs(fd))
    op.close()
    flt.close()
    grid.close()
    tf.close()
    out.update(fd_elements=nfd, fd_step=hstep, fd_max_rel_err=float(max(errs)),
               fd_median_rel_err=float(np.median(errs)), fd_time_s=time.time() - t0)
    return out


def wq_leg(np, torch, icc, ctx, grid, rho, reps=5):
    """Weighted-quadrature stiffness formation (wq_stiffness, interior layout;
    also timed with the Calabro layout) on the 96x48x48 grid with the benchmark's element densities (penal 3,
    emin 1e-9), densities already on the device: formation time (SIMP factor
    at the WQ points, sum-factorised rows, symmetrisation) and the time of
    one product with the formed 5x5x5-stencil matrix (HIP events)."""
    m = icc.WqMatrix(grid)
    xd = torch.from_numpy(rho).cuda()
    v = torch.randn(grid.ndof, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(v)
    stream = torch.cuda.ExternalStream(ctx.stream)
    m.form(icc.ICC_DENSITY_ELEMENT, xd, 3.0, 1e-9)
    m.apply(v, out=y)
    ctx.sync()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record(stream)
    for _ in range(reps):
        m.form(icc.ICC_DENSITY_ELEMENT, xd, 3.0, 1e-9)
    e1.record(stream)
    for _ in range(reps):
        m.apply(v, out=y)
    e2.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    form_ms = e0.elapsed_time(e1) / reps
    apply_ms = e1.elapsed_time(e2) / reps
    nsc = grid.ndof_sc
    out = dict(rows=nsc, stored_doubles=9 * 125 * nsc, form_ms=form_ms, apply_ms=apply_ms,
               apply_gb_per_s=9 * 125 * nsc * 8 / (apply_ms * 1e-3) / 1e9)
    m.close()
    # the same formation with the Calabro point layout (wq_stiffness_layout)
    mc = icc.WqMatrix(grid, "calabro")
    mc.form(icc.ICC_DENSITY_ELEMENT, xd, 3.0, 1e-9)
    ctx.sync()
    e0.record(stream)
    for _ in range(reps):
        mc.form(icc.ICC_DENSITY_ELEMENT, xd, 3.0, 1e-9)
    e1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    out["form_ms_calabro"] = e0.elapsed_time(e1) / reps
    mc.close()
    return out


def cpu_wq(np):
    """The oracle's restatement of wq_stiffness (element densities, assembled
    CSR) on a 32x16x16 sample of the same box."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    nel = [32, 16, 16]
    sp = O.Space([list(b) for b in BOUNDS], nel, 2)
    x = np.random.default_rng(SEED).uniform(0.01, 1.0, int(np.prod(nel)))
    t0 = time.time()
    sp.assemble_wq(1.0, 0.3, 1, x, 3.0, 1e-9, layout=0)
    dt = time.time() - t0
    rows = (nel[0] + 2) * (nel[1] + 2) * (nel[2] + 2)
    return dict(sample="WQ formation of the 32x16x16 grid (11,016 test functions), element densities",
                seconds=dt, us_per_row=dt / rows * 1e6, cores=int(O.lib().orc_num_threads()), kind="port")


def tilted_box(np, icc):
    """The reference's tilted-box B-rep (8 vertices, 12 triangles) from its
    immersed topology-optimisation fixture."""
    d = json.load(open(os.path.join(ROOT, "tests", "fixtures", "reference", "imm_topopt_tilted_box.json")))
    return icc.TriangleMesh3D(np.asarray(d["nodes"], dtype=np.float64).reshape(-1, 3),
                              np.asarray(d["elements"], dtype=np.int64).reshape(-1, 3))


def imm_topopt_leg(np, icc, ctx):
    """Immersed SIMP topology optimisation in the tilted-box B-rep with
    ImmersedTopOptConfig::default (ghost penalty gamma 1e-2, volfrac 0.35,
    penal 3, rmin 1.8, emin 1e-4, pcg_tol 1e-8): the default 28x20x16 grid
    for its 30 iterations, and a 96x72x42 grid (957,264 DOF) for 3
    iterations with FP64 PCG and with the mixed solver (FP32 and split-FP16
    inner operators)."""
    mesh = tilted_box(np, icc)
    out = {}
    for name, res, iters, solver in (("default_28x20x16", (28, 20, 16), 30, 0),
                                     ("fp64_96x72x42", (96, 72, 42), 3, 0),
                                     ("mixed_96x72x42", (96, 72, 42), 3, 1),
                                     ("mixed_f16x_96x72x42", (96, 72, 42), 3, 2)):
        cfg = icc.ImmersedTopOptConfig(grid_res=res, max_iter=iters, solver=solver)
        t0 = time.time()
        opt = icc.ImmersedTopOpt(ctx, mesh, cfg)
        t_build = time.time() - t0
        stamps = []
        t0 = time.time()
        h = opt.run(lambda it, c, v, ch, n: stamps.append(time.time()))
        dt = time.time() - t0
        out[name] = dict(grid=list(res), ndof=opt.ndof, design_cells=opt.n_active, faces=opt.n_faces,
                         iterations=iters, build_s=t_build, run_s=dt, s_per_iteration=dt / iters,
                         first_iteration_s=stamps[0] - t0, pcg_iterations=h.pcg_iterations,
                         compliance_first=h.compliance[0], compliance_last=h.compliance[-1],
                         volume_last=h.volume[-1], change_last=h.change[-1])
        opt.close()
    a = out["fp64_96x72x42"]
    for name, key in (("mixed_96x72x42", "speedup_mixed_fp32_vs_fp64"),
                      ("mixed_f16x_96x72x42", "speedup_mixed_f16x_vs_fp64")):
        b = out[name]
        b["compliance_rel_diff_vs_fp64"] = abs(b["compliance_last"] - a["compliance_last"]) / abs(a["compliance_last"])
        out[key] = a["run_s"] / b["run_s"]
    out["speedup_mixed_vs_fp64"] = out["speedup_mixed_f16x_vs_fp64"]
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-pcg", action="store_true", help="skip the PCG time-to-solution legs")
    ap.add_argument("--delta", type=float, default=0.1, help="FP32 residual drop between reliable updates")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU SpMV timing")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the immersed, batched and adjoint configuration legs")
    ap.add_argument("--apply-only", action="store_true", help="time only the apply (for profilers)")
    ap.add_argument("--apply-prec", choices=("f64", "f32", "bf16", "f16x"), default="f64",
                    help="precision of --apply-only")
    ap.add_argument("--tiling", default=None, metavar="V,W",
                    help="fixed tiling variant and workgroups per CU for the apply legs (skips the autotuner; "
                         "for profiling the configuration a default run picked)")
    ap.add_argument("--split-groups", type=int, default=0,
                    help="element groups per split-FP16 tile (1..16; 0 keeps the default)")
    ap.add_argument("--wq-only", action="store_true",
                    help="run only the WQ formation leg on the 96x48x48 grid and exit")
    ap.add_argument("--imm-only", action="store_true",
                    help="run one design iteration of the 96x72x42 immersed topology optimisation and exit")
    ap.add_argument("--immersed-only", action="store_true",
                    help="run only the immersed 128x64x64 leg (FP64 and both mixed solvers) and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain device copy of 735,000 doubles 20 times and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this implementation runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = 735000
        src = torch.randn(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(20):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 20}), flush=True)
        return

    import numpy as np
    import torch
    import immersed_catmull_clark_amd as icc

    if args.imm_only:
        ctx = icc.Context(0)
        cfg = icc.ImmersedTopOptConfig(grid_res=(96, 72, 42), max_iter=1)
        opt = icc.ImmersedTopOpt(ctx, tilted_box(np, icc), cfg)
        t0 = time.time()
        h = opt.run()
        print(json.dumps({"workload": "immersed topopt, tilted box, 96x72x42, 1 iteration, ghost penalty 1e-2",
                          "ndof": opt.ndof, "faces": opt.n_faces, "run_s": time.time() - t0,
                          "pcg_iterations": h.pcg_iterations, "compliance": h.compliance}), flush=True)
        return

    if args.immersed_only:
        ctx = icc.Context(0)
        print(json.dumps({"immersed_128x64x64": immersed_leg(np, icc, ctx)}), flush=True)
        return

    log("bench: setting up the 96x48x48 grid")
    ctx = icc.Context(0)
    grid = icc.Grid(ctx, NEL, BOUNDS)
    rho, scale, free, xh = problem(np)
    op = icc.Operator(grid, scale=scale, free=free)
    nel = grid.nel_tot
    ndof = grid.ndof
    x = torch.from_numpy(xh).cuda()
    y = torch.zeros(ndof, dtype=torch.float64, device="cuda")
    x32 = x.float()
    y32 = torch.zeros(ndof, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    if args.wq_only:
        print(json.dumps({"wq_form_96x48x48": wq_leg(np, torch, icc, ctx, grid, rho)}), flush=True)
        return

    flops = 2.0 * 81 * 81 * nel
    # tile shape and workgroups per CU picked by timing (FP64 apply first)
    fixed = tuple(int(a) for a in args.tiling.split(",")) if args.tiling else None
    if fixed:
        grid.set_tiling(*fixed)
        tuned = {"f64": (fixed[0], fixed[1], None)}
    else:
        tuned = {"f64": op.autotune(icc.ICC_F64, reps=5)}
    if args.split_groups:
        grid.set_split_tiling(args.split_groups)
    if args.apply_only and args.apply_prec in ("f32", "bf16", "f16x"):
        pr = {"f32": icc.ICC_F32, "bf16": icc.ICC_BF16, "f16x": icc.ICC_F16X}[args.apply_prec]
        if not fixed and pr != icc.ICC_F16X:
            op.autotune(pr, reps=5)
        ms32 = time_applies(torch, icc, ctx, op, x32, y32, pr, args.steps, args.warmup, accumulate=True)
        name = {"f32": "apply_tiles<float, false, 8>", "bf16": "apply_tiles<float, true, 8>",
                "f16x": "apply_split_kernel"}[args.apply_prec]
        print(json.dumps({"kernel": name, "kernel_ms": ms32, "tflops": flops / (ms32 * 1e-3) / 1e12}), flush=True)
        return
    log("bench: FP64 applies")
    t0 = time.time()
    ms_f64 = time_applies(torch, icc, ctx, op, x, y, icc.ICC_F64, args.steps, args.warmup)
    wall = time.time() - t0
    ms_kernel = time_applies(torch, icc, ctx, op, x, y, icc.ICC_F64, args.steps, args.warmup, accumulate=True)
    out = {
        "metric": METRIC, "value": 1e3 / ms_f64, "unit": "applies/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_f64, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded densities U[0.01,1], seeded unit-norm x)",
        "config": {"workload": "SIMP cantilever K.u apply, 96x48x48 p=2 (735,000 DOF, 221,184 elements)",
                   "grid": list(NEL), "ndof": ndof, "emin": 1e-9, "penal": 3, "parallelism": "single GPU"},
    }
    alg_bytes = 8 * (2 * ndof + nel) + ndof / 8
    ach = flops / (ms_kernel * 1e-3)
    out["roofline"] = {"bound": "mfma", "achieved": ach / 1e12, "peak": PEAK_F64 / 1e12, "unit": "TFLOP/s",
                       "frac": ach / PEAK_F64, "traffic": traffic_from_profile(),
                       "kernel": "apply_tiles<double>", "kernel_ms": ms_kernel,
                       "algorithmic_flops_per_launch": flops, "algorithmic_bytes_per_launch": alg_bytes}
    if args.apply_only:
        print(json.dumps(out), flush=True)
        return

    log("bench: deterministic FP64 applies")
    op.set_deterministic(True)
    ms_det = time_applies(torch, icc, ctx, op, x, y, icc.ICC_F64, args.steps, args.warmup)
    yd = op.apply(xh)
    out["apply_f64_deterministic"] = {
        "applies_per_s": 1e3 / ms_det, "ms": ms_det,
        "rel_diff_vs_atomic": float(np.linalg.norm(yd - op_apply_atomic(op, xh)) / np.linalg.norm(yd)),
        "bitwise_repeatable": bool(np.array_equal(yd, op.apply(xh))),
        "element_output_bytes": int(8 * 81 * nel)}
    op.set_deterministic(False)

    log("bench: FP32 applies")
    tuned["f32"] = op.autotune(icc.ICC_F32, reps=5)
    ms_f32 = time_applies(torch, icc, ctx, op, x32, y32, icc.ICC_F32, args.steps, args.warmup)
    ms_k32 = time_applies(torch, icc, ctx, op, x32, y32, icc.ICC_F32, args.steps, args.warmup, accumulate=True)
    yh64 = op.apply(xh)
    yh32 = op.apply(xh, prec=icc.ICC_F32)
    out["apply_f32"] = {"applies_per_s": 1e3 / ms_f32, "ms": ms_f32, "kernel_ms": ms_k32,
                        "tflops": flops / (ms_k32 * 1e-3) / 1e12, "frac_of_fp32_peak": flops / (ms_k32 * 1e-3) / PEAK_F32,
                        "rel_err_vs_f64": float(np.linalg.norm(yh32 - yh64) / np.linalg.norm(yh64))}
    log("bench: BF16 applies")
    tuned["bf16"] = op.autotune(icc.ICC_BF16, reps=5)
    ms_b = time_applies(torch, icc, ctx, op, x32, y32, icc.ICC_BF16, args.steps, args.warmup)
    ms_kb = time_applies(torch, icc, ctx, op, x32, y32, icc.ICC_BF16, args.steps, args.warmup, accumulate=True)
    yhb = op.apply(xh, prec=icc.ICC_BF16)
    out["apply_bf16"] = {"applies_per_s": 1e3 / ms_b, "ms": ms_b, "kernel_ms": ms_kb,
                         "tflops": flops / (ms_kb * 1e-3) / 1e12, "frac_of_bf16_peak": flops / (ms_kb * 1e-3) / PEAK_BF16,
                         "rel_err_vs_f64": float(np.linalg.norm(yhb - yh64) / np.linalg.norm(yh64))}

    log("bench: split-FP16 applies")
    ms_s = time_applies(torch, icc, ctx, op, x32, y32, icc.ICC_F16X, args.steps, args.warmup)
    ms_ks = time_applies(torch, icc, ctx, op, x32, y32, icc.ICC_F16X, args.steps, args.warmup, accumulate=True)
    yhs = op.apply(xh, prec=icc.ICC_F16X)
    out["apply_f16x"] = {"applies_per_s": 1e3 / ms_s, "ms": ms_s, "kernel_ms": ms_ks,
                         "tflops": flops / (ms_ks * 1e-3) / 1e12,
                         "frac_of_fp16_peak": flops / (ms_ks * 1e-3) / PEAK_BF16,
                         "executed_mfma_flops_per_element": 2 * 96 * 112 * 3,
                         "rel_err_vs_f64": float(np.linalg.norm(yhs - yh64) / np.linalg.norm(yh64))}

    if not args.no_pcg:
        # traction load on x = 2 (consistent: F = t int N dA), via the driver's builder
        cfg = icc.CantileverTopOptConfig(nel=NEL, size=(2.0, 1.0, 1.0), load=1, emin=1e-9)
        t = icc.CantileverTopOpt(ctx, cfg)
        f = t.force * free
        t.close()
        force_c1 = f
        fd = torch.from_numpy(f).cuda()
        res = {}
        legs = (("fp64", None), ("mixed_fp32", icc.ICC_F32), ("mixed_f16x", icc.ICC_F16X))
        for name, inner in legs:
            log(f"bench: PCG {name}")
            if inner != icc.ICC_F16X:
                tuned[name] = op.autotune(icc.ICC_F64 if name == "fp64" else icc.ICC_F32, reps=5)
            u = torch.zeros(ndof, dtype=torch.float64, device="cuda")
            stats = np.zeros(8)
            torch.cuda.synchronize()
            ctx.sync()
            t0 = time.time()
            if name == "fp64":
                # FP64 Jacobi PCG to 1e-8 plus the FP64 true-residual check
                rc = icc.lib().icc_solve_ir_dev(op.h, icc.ICC_F64, fd.data_ptr(), 1e-8, 1e-8, 200, 20 * ndof,
                                                u.data_ptr(), stats.ctypes.data)
            else:
                # FP32 recurrence with FP64 reliable updates every `delta` drop
                rc = icc.lib().icc_solve_mixed_prec_dev(op.h, inner, fd.data_ptr(), 1e-8, args.delta, 20 * ndof,
                                                        u.data_ptr(), stats.ctypes.data)
            ctx.sync()
            dt = time.time() - t0
            res[name] = dict(time_s=dt, status=int(rc), updates=int(stats[0]), pcg_iterations=int(stats[1]),
                             true_rel_residual=float(stats[2]), fp64_applies=int(stats[3]),
                             ms_per_iteration=1e3 * dt / max(1, int(stats[1])), u=u.cpu().numpy())
        uf = res["fp64"]["u"]
        cf = float(f @ uf)
        dcf = op.sens(uf, rho, 3.0, 1e-9)
        for name, _ in legs[1:]:
            um = res[name]["u"]
            cm = float(f @ um)
            dcm = op.sens(um, rho, 3.0, 1e-9)
            res[name].update(
                u_rel_err_vs_fp64=float(np.linalg.norm(um - uf) / np.linalg.norm(uf)),
                compliance_rel_err_vs_fp64=abs(cm - cf) / abs(cf),
                dc_rel_err_vs_fp64=float(np.linalg.norm(dcm - dcf) / np.linalg.norm(dcf)))
            res["speedup_" + name + "_vs_fp64"] = res["fp64"]["time_s"] / res[name]["time_s"]
        for k in ("fp64",) + tuple(n for n, _ in legs[1:]):
            res[k].pop("u")
        res["speedup_mixed_vs_fp64"] = res["speedup_mixed_f16x_vs_fp64"]
        res["delta"] = args.delta
        res["fp64_c0_16x8x8"] = pcg_c0(np, torch, icc, ctx)
        out["pcg_time_to_1e-8"] = res

    if not args.no_configs:
        log("bench: immersed 128x64x64")
        out["immersed_128x64x64"] = immersed_leg(np, icc, ctx)
        log("bench: batch of 24 instances")
        out["batch_24_instances"] = batch_leg(np, icc, ctx)
        log("bench: adjoint SIMP 128x64x64")
        out["adjoint_simp_128x64x64"] = adjoint_leg(np, icc, ctx)
        log("bench: immersed topology optimisation (tilted box)")
        out["imm_topopt_tilted_box"] = imm_topopt_leg(np, icc, ctx)
        log("bench: WQ formation 96x48x48")
        out["wq_form_96x48x48"] = wq_leg(np, torch, icc, ctx, grid, rho)

    if not args.no_cpu:
        log("bench: CPU baseline")
        cb, ycpu, csp, O = cpu_baseline(np, scale, (free * xh), args.cpu_budget)
        if "pcg_time_to_1e-8" in out:
            log("bench: CPU PCG (100-iteration sample at 96x48x48, full solve at 16x8x8)")
            cb["pcg"] = cpu_pcg(np, O, csp, free, force_c1, out["pcg_time_to_1e-8"]["fp64"]["pcg_iterations"])
            g0 = out["pcg_time_to_1e-8"].get("fp64_c0_16x8x8")
            if g0:
                cb["pcg"]["c0_speedup_gpu_vs_cpu"] = cb["pcg"]["c0_time_to_1e8_s"] / g0["time_s"]
            cb["pcg"]["c1_speedup_gpu_fp64_vs_cpu"] = (cb["pcg"]["c1_time_to_1e8_s_extrapolated"]
                                                      / out["pcg_time_to_1e-8"]["fp64"]["time_s"])
        yg = op.apply(xh)
        cb["rel_diff_vs_gpu_f64"] = float(np.linalg.norm(free * ycpu - yg) / np.linalg.norm(yg))
        if not args.no_configs:
            log("bench: CPU immersed topology optimisation (1 iteration, 28x20x16)")
            cb["imm_topopt_28x20x16"] = cpu_imm_topopt(np, icc)
            leg = out["imm_topopt_tilted_box"]["default_28x20x16"]
            gpu_it = leg["build_s"] + leg["first_iteration_s"]
            cb["imm_topopt_28x20x16"]["gpu_setup_and_first_iteration_s"] = gpu_it
            cb["imm_topopt_28x20x16"]["speedup_gpu_vs_cpu"] = cb["imm_topopt_28x20x16"]["s_per_iteration"] / gpu_it
            log("bench: CPU WQ formation (32x16x16 sample)")
            cb["wq_form"] = cpu_wq(np)
            wq = out["wq_form_96x48x48"]
            cb["wq_form"]["gpu_us_per_row_96x48x48"] = wq["form_ms"] * 1e3 / wq["rows"]
            cb["wq_form"]["speedup_gpu_vs_cpu_per_row"] = cb["wq_form"]["us_per_row"] / (wq["form_ms"] * 1e3 / wq["rows"])
        out["cpu_baseline"] = cb
    out["wall_s_f64_apply_loop"] = wall
    out["tiling"] = {k: {"variant": v[0], "workgroups_per_cu": v[1], "us_per_apply": v[2]} for k, v in tuned.items()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
