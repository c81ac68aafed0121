#!/usr/bin/env python
"""Assembly throughput of the MI355X drop-in for GridapDistributed's
assemble_matrix_and_vector on 3D Poisson Q2, 64^3 cells on (0,1)^3, mesh
partition (2,2,2) (8 subdomains on one GPU), kappa = exp(g) with g a seeded sum
of 16 Fourier modes, quadrature degree 4, uD = x+y+z, f = -div(kappa grad uD).

One step = one numeric assembly (coefficients at the quadrature points, RHS,
matrix values of every part, owner/ghost exchange folded in) into the CSR
structure built once by the symbolic phase.  Inputs are resident in HBM.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu-baseline]

Prints one JSON line.  metric: Gnnz/s = owned free x free nnz / step time.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gridapdistributed.jl_amd"))

HBM_PEAK_GBS = 8000.0


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def _oracle_legs(p, nnz_of, label):
    """Times the oracle's in-place numeric re-assembly (LocalView + assemble!)
    of workload p twice: one thread (the parts one after another, as in the
    reference's debug mode) and one OpenMP thread per part on the host's cores.
    Returns the cpu_baseline record; its value is the all-cores leg."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    nparts = int(np.prod(p["parts"]))
    ncpu = os.cpu_count() or 1
    threads = max(1, min(ncpu, 16, nparts))
    t0 = time.perf_counter()
    r1 = oracle.run(p, inplace_reps=1, threads=1)
    nnz = nnz_of(r1)
    t1 = r1.times[2]
    t_fresh = r1.times[1]
    del r1
    rT = oracle.run(p, inplace_reps=1, threads=threads)
    tT = rT.times[2]
    del rT
    wall = time.perf_counter() - t0
    return {
        "value": nnz / tT / 1e9,
        "unit": "Gnnz/s",
        "cores": threads,
        "kind": "port",
        "one_thread": nnz / t1 / 1e9,
        "sample": f"{label}; numeric re-assembly into the allocated PSparseMatrix (per-entry CSR search + "
                  f"ghost-row assemble!), nnz={nnz}: {threads} threads (one part each) t={tT:.3f}s, "
                  f"1 thread t={t1:.3f}s; fresh assemble_matrix_and_vector (COO+exchange+compress, 1 thread) "
                  f"{nnz / t_fresh / 1e9:.5f} Gnnz/s; host {cpu_model()}, nproc={ncpu}, "
                  f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')} (the legs set their thread "
                  f"counts explicitly); wall {wall:.1f}s",
    }


def _owned_nnz(r):
    return sum(int(P["rowptr"][P["n_own"]]) for P in r.parts)


def cpu_baseline(n=32):
    """The oracle's in-place re-assembly of the 3D Q2 Fourier-kappa workload on
    a bounded sample (32^3 cells, same parts, kappa and f)."""
    from gridap_amd import workloads as W

    p = W.poisson3d_q2_fourier(n, (2, 2, 2))
    return _oracle_legs(p, lambda r: W.nnz_free(p), f"3D Q2 {n}^3 cells, parts (2,2,2), same kappa/f")


def cpu_baseline_flux(n=48):
    """The oracle's in-place residual_and_jacobian! of the p-Laplacian at a
    smooth state on a bounded sample."""
    from gridap_amd import workloads as W

    p = W.plaplacian3d_q1(n, (4, 2, 2))
    x = W.lattice_coords(p)
    p["state"] = np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1]) * np.sin(np.pi * x[:, 2])
    return _oracle_legs(p, lambda r: W.nnz_free(p),
                        f"p-Laplacian Jacobian+residual, 3D Q1 {n}^3, parts (4,2,2), smooth state")


def cpu_baseline_elasticity(cells=(12, 10, 9)):
    """The oracle's in-place re-assembly of the elasticity problem (81 x 81 Q2
    cell matrices) on a bounded sample."""
    from gridap_amd import workloads as W

    p = W.elasticity3d_q2(cells=cells, parts=(3, 2, 2))
    return _oracle_legs(p, _owned_nnz, f"linear elasticity Q2 {cells[0]}x{cells[1]}x{cells[2]} cells, "
                                       "parts (3,2,2), E=1, nu=0.3, clamped at x=0")


def cpu_baseline_q1(n=32):
    """The oracle's in-place re-assembly of one Q1 cube of the batch's kind
    (8-mode Fourier kappa, f = 1) on a bounded sample."""
    from gridap_amd import workloads as W

    p = W.poisson3d_q1(n, (2, 2, 2), modes=W.fourier_modes(7, 8, 4, 0.5), rhs_const=1.0)
    return _oracle_legs(p, _owned_nnz, f"3D Poisson Q1 {n}^3, parts (2,2,2), 8-mode Fourier kappa")


def cpu_baseline_solve(n=28):
    """The reference's A\\b (README.md:76: the system gathered and solved by a
    sparse LU) on the host, one thread: scipy's SuperLU on the oracle's matrix
    of a bounded Q1 sample (Fourier kappa; the factorisation time does not
    depend on the coefficient's values)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import scipy.sparse.linalg as spla
    import oracle
    from _helpers import owned_global
    from gridap_amd import workloads as W

    p = W.poisson3d_q1(n, (2, 2, 2), modes=W.fourier_modes(7, 8, 4, 0.5), rhs_const=1.0)
    r = oracle.run(p)
    A, b = owned_global(r.parts, r.ngdofs)
    A = A.tocsc()
    t0 = time.perf_counter()
    x = spla.spsolve(A, b)
    t = time.perf_counter() - t0
    rr = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
    return {"value": round(t, 3), "unit": "s", "cores": 1, "kind": "port",
            "sample": f"sparse LU solve (scipy SuperLU) of the oracle's 3D Q1 {n}^3 matrix, {r.ngdofs} free dofs, "
                      f"relres {rr:.1e}; host {cpu_model()}"}


def bench_mixed(prob, args, nnz, ms_fp64, variant, rows_z):
    """The mixed form (FP32 sum-factorised contractions, FP64 CSR values, RHS
    and lifting) on the same problem and the same kernel family as the FP64
    form: the direct-store kernel, and the z-marching kernel at its best column
    length beside it; Gnnz/s over the FP64 form's, and the error against the
    FP64 values (max and Frobenius, relative)."""
    import numpy as np

    def values():
        prob.sync()
        parts = [prob.export_part(k) for k in range(prob.nparts)]
        return np.concatenate([P["val"] for P in parts]), np.concatenate([P["b"] for P in parts])

    def timed():
        for _ in range(args.warmup):
            prob.numeric(precision=1)
        prob.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            prob.numeric(precision=1)
        prob.sync()
        return (time.perf_counter() - t0) * 1e3 / args.steps

    prob.numeric()
    a64, b64 = values()
    forms = {}
    prob.set_kernel(2)
    forms["direct-store"] = (timed(), 2, 0)
    best = None
    for rz in (16, 32, 64):
        prob.set_kernel(0, rz)
        ms = timed()
        if best is None or ms < best[0]:
            best = (ms, 0, rz)
    forms["z-marching"] = best
    out = {}
    for name, (ms, var, rz) in forms.items():
        prob.set_kernel(var, rz if rz else 32)
        prob.numeric(precision=1)
        a32, b32 = values()
        d = a32 - a64
        out[name] = {"value_gnnz_s": round(nnz / (ms * 1e-3) / 1e9, 3), "ms_per_step": round(ms, 4),
                     "speedup_vs_fp64": round(ms_fp64 / ms, 3),
                     "max_rel_err": float(np.abs(d).max() / np.abs(a64).max()),
                     "fro_rel_err": float(np.linalg.norm(d) / np.linalg.norm(a64)),
                     "rhs_max_rel_err": float(np.abs(b32 - b64).max() / np.abs(b64).max())}
        if var == 0:
            out[name]["z_rows_per_workgroup"] = rz
    prob.set_kernel(variant, rows_z)  # back to the FP64 form's tuned kernel
    prob.numeric()
    return out


def bench_plaplacian(args):
    """Newton on the p = 3 p-Laplacian, 3D Q1 96^3 cells, parts (4,2,2), f = 1,
    u = 0 on the boundary, sigma = (eps^2 + |grad u|^2)^(1/2) grad u with
    eps = 1e-2; Poisson initial guess; stops at ||r|| <= 1e-10 ||r(0)||.  One
    step = one full Newton solve (every Jacobian re-assembled in place), its
    corrections solved by FPCG with the multilevel preconditioner rebuilt from
    each Jacobian (coarse network off: it does not gain, see the learned
    workload); the same solve with Jacobi-PCG corrections is timed beside it,
    and so are the forms the configuration names: FP32-contracted Jacobians
    (mixed-precision re-assembly), and those with the coarse network on,
    trained on this problem's own Jacobian."""
    import numpy as np

    import gridap_amd as G
    from gridap_amd import workloads as W

    n = args.cells if args.cells != 64 else 96
    p = W.plaplacian3d_q1(n, (4, 2, 2), p=3.0, eps=1e-2)
    ctx = G.DeviceContext(0)
    op = G.flux_operator_from_workload(ctx, p)
    t0 = time.perf_counter()
    prob = op._problem()
    t_symbolic = time.perf_counter() - t0
    nnz = prob.owned_nnz()
    net = G.LearnedCoarseSolver(ctx, n // 2, 0)
    net.attach(prob)
    net.set_options(8, use_network=False)
    kw = dict(rtol=1e-10, maxiter=40, lin_rtol=1e-8, lin_maxiter=20000, init=G._lib.NEWTON_INIT_POISSON)

    def timed(precond, jac_precision=G._lib.PREC_FP64):
        for _ in range(args.warmup):
            prob.newton(None, precond=precond, jac_precision=jac_precision, **kw)
        prob.sync()
        reps = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            x, rep = prob.newton(None, precond=precond, jac_precision=jac_precision, **kw)
            reps.append(rep)
        prob.sync()
        wall = (time.perf_counter() - t0) / args.steps
        # the solve with x resident on the device (the copy of x to the host is outside)
        return x, reps[-1], float(np.mean([r["ms_total"] for r in reps])) / 1e3, wall

    x_j, rep_j, s_jacobi, wall_jacobi = timed(None)
    x, rep, s_per_solve, wall_per_solve = timed(net)
    sol_rel_diff = float(np.linalg.norm(x - x_j) / np.linalg.norm(x_j))
    # mixed-precision re-assembly: FP32 cell Jacobians (residual and solves FP64)
    x32, rep32, s32, _ = timed(net, G._lib.PREC_FP32)
    prob.set_state(x)

    def asm_time(precision):
        acc = 0.0
        for _ in range(10):
            prob.numeric(precision)
            acc += prob.phase_times()[3]
        return acc / 10

    asm_ms = asm_time(G._lib.PREC_FP64)
    asm32_ms = asm_time(G._lib.PREC_FP32)
    # the coarse network on, trained on the Jacobian at the solution (smooth
    # coarse errors of several widths), with FP32 Jacobians
    prob.numeric()
    net.attach(prob)
    t0 = time.perf_counter()
    for k in range(C5_TRAIN_STEPS):
        net.train_step(31 * k + 7, 2e-3 if k < C5_TRAIN_STEPS // 2 else 1e-3, smooth_passes=(8, 16, 32, 64)[k % 4])
    prob.sync()
    t_train = time.perf_counter() - t0
    net.set_options(8, use_network=True)
    x_nn, rep_nn, s_nn, _ = timed(net, G._lib.PREC_FP32)
    net.set_options(8, use_network=False)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_k_flux_fused_plaplacian_{n}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("traffic_bytes")
    out = {
        "metric": "p-Laplacian Newton solve time to 1e-10 (s)",
        "value": round(s_per_solve, 4),
        "unit": "s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(s_per_solve * 1e3, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"p-Laplacian p=3, eps=1e-2, 3D Q1 {n}^3 on (0,1)^3, parts (4,2,2), f=1, u=0 on the "
                               "boundary; Newton (Poisson initial guess) with multilevel FPCG corrections to 1e-8",
                   "free_dofs": prob.ngdofs, "nnz_owned": nnz, "symbolic_s": round(t_symbolic, 3),
                   "newton_iters": rep["iters"], "fpcg_iters": rep["lin_iters"], "converged": rep["converged"],
                   "fpcg_fallbacks": rep["lin_fallbacks"],
                   "res_ref": rep["res_ref"], "res_final": rep["res_final"],
                   "ms_assembly_total": round(rep["ms_assembly"], 2), "ms_solve_total": round(rep["ms_solve"], 2),
                   "s_wall_incl_host_copy_of_x": round(wall_per_solve, 4),
                   "jacobi_pcg_newton": {"s": round(s_jacobi, 4), "s_wall_incl_host_copy_of_x": round(wall_jacobi, 4),
                                         "newton_iters": rep_j["iters"],
                                         "pcg_iters": rep_j["lin_iters"], "converged": rep_j["converged"],
                                         "res_final": rep_j["res_final"], "ms_solve_total": round(rep_j["ms_solve"], 2)},
                   "speedup_vs_jacobi_pcg_newton": round(s_jacobi / s_per_solve, 3),
                   "solution_rel_diff_vs_jacobi": sol_rel_diff,
                   "jacobian_assembly_ms": round(asm_ms, 4),
                   "jacobian_assembly_gnnz_s": round(nnz / (asm_ms * 1e-3) / 1e9, 2),
                   "mixed_jacobians": {"s": round(s32, 4), "newton_iters": rep32["iters"],
                                       "fpcg_iters": rep32["lin_iters"], "converged": rep32["converged"],
                                       "res_final": rep32["res_final"],
                                       "solution_rel_diff_vs_fp64": float(np.linalg.norm(x32 - x) / np.linalg.norm(x)),
                                       "jacobian_assembly_ms": round(asm32_ms, 4),
                                       "speedup_assembly_vs_fp64": round(asm_ms / asm32_ms, 3)},
                   "network_on_mixed_jacobians": {"s": round(s_nn, 4), "newton_iters": rep_nn["iters"],
                                                  "fpcg_iters": rep_nn["lin_iters"], "converged": rep_nn["converged"],
                                                  "fpcg_fallbacks": rep_nn["lin_fallbacks"],
                                                  "res_final": rep_nn["res_final"],
                                                  "solution_rel_diff_vs_fp64": float(np.linalg.norm(x_nn - x) /
                                                                                     np.linalg.norm(x)),
                                                  "training_steps": C5_TRAIN_STEPS, "training_s": round(t_train, 2)}},
        "roofline": {"bound": "hbm", "achieved": round(8.0 * nnz / (asm_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(8.0 * nnz / (asm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": "k_flux_fused<2,2,4,4,256,double>", "alg_bytes_per_launch": int(8 * nnz)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_flux()
    net.close()
    prob.close()
    ctx.close()
    print(json.dumps(out), flush=True)


# seed of the mapped-grid bench's vertex displacements
MAPPED_SEED = 11

# seed of the elasticity bench's Young's modulus field
ELAST_CELLWISE_SEED = 31

# Adam steps of the p-Laplacian workload's coarse network (on the Jacobian at
# the solution)
C5_TRAIN_STEPS = 600

# damped-Jacobi sweeps before and after each coarse-grid correction of the
# learned workload's V-cycle (at 128^3: 15 FPCG iterations with 1, 13 with 2,
# for 7% less time)
NU_C4 = 2


def bench_learned(args):
    """3D Poisson Q1, 128^3 cells on (0,1)^3, parts (2,2,2), kappa = exp(2 g) with g
    a unit-variance Gaussian random field of squared-exponential correlation length
    0.1, homogeneous Dirichlet, random N(0,1) right-hand side.  The BF16 network
    (the coarse-grid guess of the multilevel preconditioner) is trained on
    --fields seeded fields (training time reported, not part of the metric), then
    the solvers run to the FP64 true residual 1e-8 on held-out fields.  One step =
    one solve with the learned preconditioner; Jacobi-PCG and the forms without
    the network or without the V-cycle are timed beside it."""
    import numpy as np

    import gridap_amd as G
    from gridap_amd import workloads as W

    cells = args.cells if args.cells != 64 else 128
    n = cells // 2
    ctx = G.DeviceContext(0)
    p = W.poisson3d_q1(cells, (2, 2, 2), rhs_const=0.0)
    assem, data = G.problem_from_workload(ctx, p)
    prob = assem._problem_for(data)
    net = G.LearnedCoarseSolver(ctx, n, 1)
    # training: one field per step block, smooth coarse errors of several widths
    t0 = time.perf_counter()
    losses = []
    nsteps = 0
    for f in range(args.fields):
        prob.cell_coefficient_grf(100000 + f, 0.1, 2.0)
        prob.numeric()
        net.attach(prob)
        for k in range(2):
            lr = 2e-3 if f < args.fields // 2 else (1e-3 if f < 3 * args.fields // 4 else 4e-4)
            losses.append(net.train_step(7919 * f + k, lr, smooth_passes=(8, 16, 32, 64)[(2 * f + k) % 4]))
            nsteps += 1
        if f % 200 == 0:
            print(f"train field {f}: loss {np.mean(losses[-20:]):.4f}", file=sys.stderr, flush=True)
    prob.sync()
    t_train = time.perf_counter() - t0
    # held-out fields: Jacobi-PCG and FPCG with the preconditioner in four forms
    forms = {"vcycle_net": (True, True), "vcycle": (True, False), "two_grid_net": (False, True),
             "two_grid": (False, False)}
    res = {k: [] for k in ["jacobi"] + list(forms)}
    for k in range(args.warmup + args.steps):
        prob.cell_coefficient_grf(900000 + k, 0.1, 2.0)
        prob.numeric()
        net.attach(prob)
        rhs = np.random.default_rng(5000 + k).normal(size=prob.ngdofs)
        def solve(nn):
            # (x, iterations, true relres, device ms, fell back, wall s with the host copies of rhs and x)
            t0 = time.perf_counter()
            r = G.fpcg(prob, nn, rhs, 1e-8, 50000)
            return (*r, time.perf_counter() - t0)

        rows = {"jacobi": solve(None)}
        for name, (mg, use_net) in forms.items():
            net.set_multigrid(mg, NU_C4)
            net.set_options(8, use_network=use_net)
            rows[name] = solve(net)
        if k >= args.warmup:
            for name, r in rows.items():
                res[name].append(r[1:])  # iterations, true relres, ms, fell back, wall s
    levels = net.set_multigrid(True, NU_C4)
    # roofline of the solves' dominant kernel, the CSR SpMV, timed live with hipEvents
    nnz, nd = prob.owned_nnz(), prob.ngdofs
    # bytes: value (8 B) + column id (4 B, 2 B where a block holds 16-bit offsets) per entry;
    # row start, length, x and y per row
    ms_spmv, spmv_bytes = prob.spmv_device_timing(50)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_k_spmv_blk_learned_q1_{cells}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("traffic_bytes")
    summ = {}
    for name, rs in res.items():
        it, rr, ms, fb, wall = (np.array([r[i] for r in rs], dtype=float) for i in range(5))
        summ[name] = {"s": round(float(ms.mean()) / 1e3, 4), "s_wall_incl_host_copies": round(float(wall.mean()), 4),
                      "iters": int(round(it.mean())),
                      "relres_max": float(rr.max()), "fallbacks": int(fb.sum())}
    s_j = summ["jacobi"]["s"]
    for name in forms:
        summ[name]["time_ratio_over_jacobi"] = round(summ[name]["s"] / s_j, 3)
    s_l = summ["vcycle_net"]["s"]
    out = {
        "metric": "learned-preconditioned FPCG time to 1e-8 (s)",
        "value": s_l, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(s_l * 1e3, 2), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 (network bf16/fp32)", "data": "synthetic",
        "config": {"workload": f"3D Poisson Q1 {cells}^3 on (0,1)^3, parts (2,2,2), kappa=exp(2 g), g GRF corr 0.1, "
                               "random N(0,1) RHS, homogeneous Dirichlet; FPCG with the learned multilevel "
                               "preconditioner (network guess on the first coarse grid, then a Galerkin V-cycle) vs "
                               "Jacobi-PCG, true FP64 residual 1e-8",
                   "free_dofs": prob.ngdofs, "network_params": net.num_params, "training_fields": args.fields,
                   "training_steps": nsteps, "training_s": round(t_train, 1), "vcycle_grids": levels,
                   "vcycle_smoothing_sweeps": NU_C4,
                   "loss_first": round(float(np.mean(losses[:20])), 4), "loss_last": round(float(np.mean(losses[-20:])), 4),
                   "time_ratio_learned_over_jacobi": round(s_l / s_j, 3),
                   "solvers": summ, "spmv_ms": round(ms_spmv, 4), "nnz_owned": nnz},
        "roofline": {"bound": "hbm", "achieved": round(spmv_bytes / (ms_spmv * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(spmv_bytes / (ms_spmv * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": "k_spmv_blk<128,512>", "alg_bytes_per_launch": int(spmv_bytes)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_solve()
    net.close()
    prob.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def elasticity_problem(ctx, G, W):
    """The elasticity problem of the batched workload: Q2 VectorValue{3}, 96x80x72
    cells on (0,1)^3, parts (3,2,2), E = 1, nu = 0.3, clamped at x = 0, body
    force (0,0,-1)."""
    p = W.elasticity3d_q2(cells=(96, 80, 72), parts=(3, 2, 2))
    t0 = time.perf_counter()
    assem, data = G.problem_from_workload(ctx, p)
    prob = assem._problem_for(data)
    return assem, prob, time.perf_counter() - t0


def elasticity_kernel_ms(prob, reps=5):
    """Average k_elast duration from the library's hipEvents around its launch."""
    acc = 0.0
    for _ in range(reps):
        prob.numeric()
        acc += prob.phase_times()[3]
    return acc / reps


def elasticity_roofline(prob, ms):
    nnz = prob.owned_nnz()
    alg = 8 * nnz  # FP64 value written per stored owned entry
    traffic = None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_k_elast_elasticity_q2_96x80x72.json")
    if os.path.exists(path):
        with open(path) as fh:
            traffic = json.load(fh).get("traffic_bytes")
    ach = alg / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "k_elast<2>",
            "alg_bytes_per_launch": int(alg)}


def bench_elasticity(args):
    """Linear elasticity Q2 (81 x 81 cell matrices) on 96x80x72 cells, parts
    (3,2,2), clamped at x = 0: the numeric assembly alone (one step = one
    assembly of the 2.5 G-entry matrix and the body-force vector)."""
    import gridap_amd as G
    from gridap_amd import workloads as W

    ctx = G.DeviceContext(0)
    assem, prob, t_sym = elasticity_problem(ctx, G, W)
    nnz = prob.owned_nnz()
    for _ in range(args.warmup):
        prob.numeric()
    prob.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prob.numeric()
    prob.sync()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    kms = elasticity_kernel_ms(prob)
    # the same assembly with a piecewise-constant material: a lognormal Young's
    # modulus field (kappa_c scaling lambda and mu), drawn on the device
    prob.cell_coefficient_grf(ELAST_CELLWISE_SEED, 0.1, 1.0)
    for _ in range(args.warmup):
        prob.numeric()
    prob.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prob.numeric()
    prob.sync()
    ms_cw = (time.perf_counter() - t0) * 1e3 / args.steps
    kms_cw = elasticity_kernel_ms(prob)
    cw = elasticity_roofline(prob, kms_cw)
    cellwise = {"ms": round(ms_cw, 4), "gnnz_s": round(nnz / (ms_cw * 1e-3) / 1e9, 2), "kernel_ms": round(kms_cw, 4),
                "kernel": "k_elast_cw<2,2,1>", "achieved_gbs": cw["achieved"], "frac": cw["frac"],
                "field": f"kappa = exp(g), g a unit-variance GRF, correlation length 0.1, seed {ELAST_CELLWISE_SEED}"}
    prob.set_cell_coefficient(None)
    out = {
        "metric": "3D Q2 linear elasticity assembly Gnnz/s (FP64)",
        "value": round(nnz / (ms * 1e-3) / 1e9, 2), "unit": "Gnnz/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "linear elasticity Q2 VectorValue{3}, 96x80x72 cells on (0,1)^3, parts (3,2,2), E=1, "
                               "nu=0.3, clamped at x=0, body force (0,0,-1), SubAssembledRows",
                   "free_dofs": prob.ngdofs, "nnz_owned": nnz, "symbolic_s": round(t_sym, 2),
                   "kernel_ms": round(kms, 4), "cellwise_young_modulus": cellwise},
        "roofline": elasticity_roofline(prob, kms),
    }
    prob.close()
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_elasticity()
    print(json.dumps(out), flush=True)


def bench_mapped(args):
    """The C2 problem (3D Poisson Q2 64^3, parts (2,2,2), Fourier kappa, quad
    degree 4) on a mapped grid: interior vertices moved at random by up to a
    quarter cell plus a smooth bulge (MAPPED_SEED), every cell with its own
    multilinear map.  One step = one numeric assembly: the cell records (the
    Jacobians, kappa at the mapped Gauss points, the 27 x 27 cell matrices and
    cell vectors), then the rows."""
    import gridap_amd as G
    from gridap_amd import workloads as W

    p = W.poisson3d_q2_fourier(args.cells, (2, 2, 2))
    p["vertex_coords"] = W.mapped_vertices(p, amp=0.5, seed=MAPPED_SEED, smooth=0.05)
    ctx = G.DeviceContext(0)
    assem, data = G.problem_from_workload(ctx, p)
    prob = assem._problem_for(data)
    nnz = prob.owned_nnz()
    for _ in range(args.warmup):
        prob.numeric()
    prob.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prob.numeric()
    prob.sync()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    acc = [0.0, 0.0, 0.0, 0.0]
    for _ in range(args.steps):
        prob.numeric()
        acc = [a + b for a, b in zip(acc, prob.phase_times())]
    cells_ms, _, rows_ms, _ = [a / args.steps for a in acc]
    ncell = int(np.prod(p["cells"]))
    rec = 8 * ncell * (27 * 27 + 27)  # cell records written by the first kernel
    # second kernel: 8 B per stored value written, every record entry read once
    alg_rows = 8.0 * nnz + rec
    ach = alg_rows / (rows_ms * 1e-3) / 1e9
    traffic = None  # HBM bytes of the row kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE passes at 64^3
    tpath = os.path.join(ROOT, "profiles", "traffic_k_mapped_rows_poisson3d_q2_64_mapped.json")
    if args.cells == 64 and os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh)["traffic_bytes"]
    out = {
        "metric": "3D Q2 assembly Gnnz/s (FP64), mapped grid",
        "value": round(nnz / (ms * 1e-3) / 1e9, 2), "unit": "Gnnz/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D Poisson Q2 {args.cells}^3 on (0,1)^3 mapped (interior vertices moved up to a "
                               f"quarter cell at random, seed {MAPPED_SEED}, plus a smooth bulge 0.05), parts (2,2,2), "
                               "kappa=exp(16 seeded Fourier modes) at the mapped Gauss points, quad degree 4",
                   "nnz_owned": nnz, "free_dofs": prob.ngdofs,
                   "kernel_ms": {"k_mapped_cells": round(cells_ms, 4), "k_mapped_rows": round(rows_ms, 4)},
                   "cell_record_bytes": rec,
                   "cells_kernel_gflop_s": round(ncell * 27 * 729 * 8.0 / (cells_ms * 1e-3) / 1e9, 1)},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "k_mapped_rows<3,2,1>",
                     "alg_bytes_per_launch": int(alg_rows)},
    }
    prob.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_batched(args):
    """The batched workload: 32 independent 3D Poisson Q1 problems (cubes of
    16^3..48^3 cells drawn from a seed, 1..8 parts each, 8-mode Fourier kappa,
    f = 1) plus the linear-elasticity Q2 problem (96x80x72 cells, parts (3,2,2)),
    every numeric assembly in one hipGraph launch after per-problem autotuning,
    against the same launches issued one by one.  One step = one batch (all 33
    problems).  --no-elasticity leaves the elasticity problem out."""
    import gridap_amd as G
    from gridap_amd import workloads as W

    ctx = G.DeviceContext(0)
    ws = W.batched_q1(seed=7, count=32)
    t0 = time.perf_counter()
    probs = []
    for q in ws:
        assem, data = G.problem_from_workload(ctx, q)
        probs.append((assem, assem._problem_for(data)))
    t_sym = time.perf_counter() - t0
    nnz_q1 = sum(p.owned_nnz() for _, p in probs)
    el = None
    if not args.no_elasticity:
        ea, ep, t_sym_el = elasticity_problem(ctx, G, W)
        el = (ea, ep)
        probs.append(el)
    nnz = sum(p.owned_nnz() for _, p in probs)
    # eager: one problem after the other
    for _ in range(args.warmup):
        for _, p in probs:
            p.numeric()
    probs[0][1].sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for _, p in probs:
            p.numeric()
    probs[0][1].sync()
    ms_eager = (time.perf_counter() - t0) * 1e3 / args.steps
    b = G.AssemblyBatch(ctx)
    for _, p in probs:
        b.add(p)
    chosen = b.autotune(reps=3)
    b.capture()
    for _ in range(args.warmup):
        b.launch()
    b.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.launch()
    b.sync()
    ms_graph = (time.perf_counter() - t0) * 1e3 / args.steps
    chosen = chosen[:len(ws)]
    cfg = {"workload": "32 independent 3D Poisson Q1 problems, cubes 16^3..48^3 cells (seed 7), 1..8 parts each, "
                       "8-mode Fourier kappa, f=1" +
                       ("" if el is None else " + linear elasticity Q2 96x80x72, parts (3,2,2), E=1, nu=0.3, clamped "
                                              "at x=0, body force (0,0,-1)") +
                       "; numeric assembly of all in one hipGraph launch",
           "nnz_owned_total": nnz, "nnz_owned_q1": nnz_q1, "symbolic_s_q1_total": round(t_sym, 3),
           "ms_eager_one_by_one": round(ms_eager, 4), "speedup_graph_vs_eager": round(ms_eager / ms_graph, 3),
           "kernel_chosen_q1": {"z-marching 16": int((chosen == 16).sum()), "z-marching 32": int((chosen == 32).sum()),
                                "z-marching 64": int((chosen == 64).sum()), "per-tile": int((chosen == -1).sum()),
                                "per-tile direct-store": int((chosen == -2).sum()),
                                "per-tile wave-slab": int((chosen == -3).sum())}}
    out = {
        "metric": "batched assembly Gnnz/s (FP64), 32 Q1 problems" + ("" if el is None else " + elasticity Q2") +
                  " per hipGraph launch",
        "value": round(nnz / (ms_graph * 1e-3) / 1e9, 3), "unit": "Gnnz/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_graph, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic", "config": cfg,
    }
    if el is not None:
        kms = elasticity_kernel_ms(el[1])
        cfg["elasticity"] = {"nnz_owned": el[1].owned_nnz(), "free_dofs": el[1].ngdofs,
                             "symbolic_s": round(t_sym_el, 2), "kernel_ms": round(kms, 4),
                             "gnnz_s": round(el[1].owned_nnz() / (kms * 1e-3) / 1e9, 2)}
        out["roofline"] = elasticity_roofline(el[1], kms)
    b.close()
    for _, p in probs:
        p.close()
    ctx.close()
    if not args.no_cpu_baseline:
        if el is not None:  # the elasticity problem holds 98.6% of the batch's entries
            out["cpu_baseline"] = cpu_baseline_elasticity()
        else:
            out["cpu_baseline"] = cpu_baseline_q1()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["poisson-q2", "plaplacian", "learned", "batched", "elasticity", "mapped"],
                    default="poisson-q2",
                    help="poisson-q2: 3D Q2 assembly throughput (default); plaplacian: p=3 Newton on 3D Q1; "
                         "learned: FPCG with the learned preconditioner vs Jacobi-PCG on 3D Q1 128^3; batched: 32 Q1 "
                         "problems + elasticity Q2 in one hipGraph; elasticity: the elasticity Q2 assembly alone")
    ap.add_argument("--no-elasticity", action="store_true", help="batched workload without the elasticity problem")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="default workload only: skip the other configurations' runs summarised in other_configs")
    ap.add_argument("--fields", type=int, default=2000, help="training fields (learned workload)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 1 GiB (scale reference for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        n = 1 << 27  # 2^27 doubles = 1 GiB
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: all parts of the partition live on a single MI355X; --gpus {args.gpus} "
              "is not supported in this build (see DESIGN.md for how the path would shard).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "plaplacian":
        return bench_plaplacian(args)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "elasticity":
        return bench_elasticity(args)
    if args.workload == "mapped":
        return bench_mapped(args)

    import gridap_amd as G
    from gridap_amd import workloads as W

    p = W.poisson3d_q2_fourier(args.cells, (2, 2, 2))
    ctx = G.DeviceContext(0)
    assem, data = G.problem_from_workload(ctx, p)
    t0 = time.perf_counter()
    prob = assem._problem_for(data)  # symbolic phase, once
    t_symbolic = time.perf_counter() - t0
    t0 = time.perf_counter()
    prob.symbolic()  # again on the same problem: the phase alone, past first-call set-up
    t_symbolic_again = time.perf_counter() - t0
    nnz = prob.owned_nnz()
    # autotuner: z-marching columns of 16/32/64 rows, the per-tile kernel or its direct-store / wave-slab forms
    tuner = G.AssemblyBatch(ctx)
    tuner.add(prob)
    chosen = int(tuner.autotune(reps=3)[0])
    tuner.close()
    variant, rows_z = prob.get_kernel()
    kernel = {0: f"k_march3d<2,3,4,8> ({rows_z} z rows per workgroup)", 1: "k_assemble3d<2,3,4,8,8>",
              2: "k_assemble3d_direct<2,3,4,8,8>", 3: "k_assemble3d_wslab<3,8>"}[variant]

    for _ in range(args.warmup):
        prob.numeric()
    prob.sync()
    # timed region: K numeric assemblies back to back
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prob.numeric()
    prob.sync()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    # per-kernel durations from hipEvents recorded on the library's stream
    acc = [0.0, 0.0, 0.0, 0.0]
    for _ in range(args.steps):
        prob.numeric()
        t = prob.phase_times()
        acc = [a + b for a, b in zip(acc, t)]
    coef_ms, rhs_ms, mat_ms, tot_ms = [a / args.steps for a in acc]

    value = nnz / (ms_per_step * 1e-3) / 1e9
    mixed = bench_mixed(prob, args, nnz, ms_per_step, variant, rows_z)
    alg_bytes = 8.0 * nnz  # FP64 CSR values written per launch (8 B per stored owned entry)
    achieved = alg_bytes / (mat_ms * 1e-3) / 1e9
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    # (copy-calibrated) on this workload; only valid at the default size
    traffic = None
    tname = {0: "k_march3d", 1: "k_assemble3d", 2: "k_assemble3d_direct", 3: "k_assemble3d_wslab"}[variant]
    tpath = os.path.join(ROOT, "profiles", f"traffic_{tname}_poisson3d_q2_64.json")
    if args.cells == 64 and os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh)["traffic_bytes"]
    out = {
        "metric": "3D Q2 assembly Gnnz/s (FP64)",
        "value": round(value, 3),
        "unit": "Gnnz/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D Poisson Q2 {args.cells}^3 cells on (0,1)^3, parts (2,2,2) on one GPU, "
                        "kappa=exp(16 seeded Fourier modes), quad degree 4, SubAssembledRows",
            "nnz_owned": nnz,
            "free_dofs": prob.ngdofs,
            "parts": 8,
            "symbolic_s": round(t_symbolic, 3),
            "symbolic_s_again": round(t_symbolic_again, 4),
            "autotuned": {"variant": {0: "z-marching", 1: "per-tile", 2: "per-tile direct-store",
                                      3: "per-tile wave-slab"}[variant],
                          "z_rows_per_workgroup": rows_z},
            "mixed_fp32": mixed,
            "kernel_ms": {"coefficients": round(coef_ms, 4), "rhs": round(rhs_ms, 4), "matrix": round(mat_ms, 4),
                          "step_events": round(tot_ms, 4)},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": kernel,
            "alg_bytes_per_launch": int(alg_bytes),
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    prob.close()
    ctx.close()
    if not args.no_other_configs:
        out["other_configs"] = other_configs()
    print(json.dumps(out), flush=True)


# the key fields of each other configuration's bench line
OTHER = {
    "plaplacian": ("C5", ["newton_iters", "fpcg_iters", "converged", "speedup_vs_jacobi_pcg_newton", "jacobian_assembly_ms",
                          "symbolic_s", "mixed_jacobians", "network_on_mixed_jacobians"]),
    "learned": ("C4", ["training_fields", "training_s", "time_ratio_learned_over_jacobi"]),
    "elasticity": ("C3 elasticity alone", ["nnz_owned", "symbolic_s", "kernel_ms", "cellwise_young_modulus"]),
    "mapped": ("C2 on a mapped grid", ["kernel_ms", "cells_kernel_gflop_s"]),
    "batched": ("C3", ["nnz_owned_total", "ms_eager_one_by_one", "speedup_graph_vs_eager"]),
}


def other_configs():
    """The other configurations, each run as its own bench process (default
    steps, no CPU baseline) after the headline measurement, summarised."""
    import subprocess

    res = {}
    for w, (name, keys) in OTHER.items():
        cmd = [sys.executable, os.path.abspath(__file__), "--workload", w, "--no-cpu-baseline"]
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=400)
            line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1] if r.returncode == 0 else None
        except (subprocess.TimeoutExpired, IndexError):
            line, r = None, None
        if line is None:
            res[w] = {"config": name, "error": (r.stderr[-400:] if r is not None else "timed out or no result")}
            continue
        d = json.loads(line)
        item = {k: d[k] for k in ("metric", "value", "unit", "ms_per_step", "higher_is_better") if k in d}
        item["config"] = name
        item["workload"] = d.get("config", {}).get("workload")
        cfg = d.get("config", {})
        for k in keys:
            if k in cfg:
                item[k] = cfg[k]
        if w == "learned":
            item["solvers"] = {k: {kk: v.get(kk) for kk in ("s", "iters", "time_ratio_over_jacobi") if kk in v}
                               for k, v in cfg.get("solvers", {}).items()} if isinstance(cfg.get("solvers"), dict) else None
        if "roofline" in d:
            item["roofline"] = {k: d["roofline"].get(k) for k in ("bound", "achieved", "peak", "unit", "frac", "kernel")}
        res[w] = item
    return res


if __name__ == "__main__":
    main()
