#!/usr/bin/env python3
"""Benchmark of the MI355X matrix-free DG operator on the P4 tetrahedral elasticity workload.

Workload: 3D linear elasticity, polynomial degree 4 on tetrahedra, [0,1]^3 split into 24^3 cubes of
6 Kuhn tetrahedra (82,944 elements, 35 x 3 basis functions each: 8,709,120 DoF), degree-4 geometry
warped by 0.02 sin(2 pi x) sin(2 pi y) sin(2 pi z), seeded log-normal lambda/mu (means 2 and 1,
sigma 0.3, correlation length 0.2), Dirichlet on all faces.  A step is one matrix-free apply
y = A x of a seeded uniform[-1,1] vector in FP64, inputs resident in HBM; each precision is timed
after 0.3 s of untimed applies (the GPU's clocks ramp up under sustained load) and W warm-up steps.

Prints one JSON line: GDoF/s of the FP64 apply (`value`), the roofline of the apply (its two
kernels, facet then element, back to back on one stream) measured with HIP events on that stream,
the CPU baseline (the CPU FP64 oracle timed on the host cores on the same mesh and vector), the FP32
apply, the launch shapes the autotuner chose, and the time to a 1e-10 relative residual of FP64 CG
versus FP32 CG with FP64 residual replacement on the same mesh and body force, without and with the
element-block Jacobi preconditioner.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-solve] [--no-cpu-baseline] [--no-autotune]
    python bench.py --workload precision|greens|adjoint|multimesh|poisson|gradient   (the other studies; one JSON line each)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GDoF/s of P4 elasticity apply; time-to-1e-10 solve FP32+IR vs FP64, 1 GPU"
SEED = 20261203
PEAK_FP64 = 78.6e12   # MI355X dense FP64 (matrix and vector), FLOP/s
PEAK_FP32 = 157.3e12  # MI355X dense FP32 matrix, FLOP/s
PEAK_HBM = 8.0e12     # bytes/s
# the FP64 apply: the facet kernel then the element kernel of the two-kernel matrix-core apply
KERNELS = ["elasticity_facet_kernel<3, 4, double, ...>", "elasticity_element_kernel<3, 4, double, ...>"]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def on_stream(op):
    """Run the operator on torch's current stream: a non-default stream (set by main for the solver
    workloads), so that the solver iterations are replayed from captured HIP graphs and torch's
    own kernels are ordered with the library's."""
    import torch

    s = torch.cuda.current_stream()
    op.set_stream(None if s.cuda_stream == 0 else s)
    return op


def build_workload(cells, degree=4, seed=SEED):
    import tandem_amd as T
    from tandem_amd import scenario as S

    mesh = T.Mesh.box(3, cells, scheme=1, bc=T.BC_DIRICHLET)
    nodes = S.warp(mesh.affine_nodes(degree))
    op = T.DGOperator(mesh, T.ELASTICITY, degree, nodes)
    X = op.volume_points()
    lam, mu = S.lame_fields(3, seed)
    op.material_Q = (lam(X), mu(X))
    op.set_material(*op.material_Q)
    return op, mesh, nodes, (lam, mu)


def prewarm(op, x, y, seconds=0.3):
    """Untimed applies for `seconds` of wall time: from idle the GPU needs ~0.1 s of sustained work
    before the apply runs at its steady rate (measured: the first ~100 applies of 24^3-cube P4 run
    up to 10% slower), so every timing below starts from the steady state."""
    import torch

    op.apply(x, y)  # the first apply also uploads the operator
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(10):
            op.apply(x, y)
        torch.cuda.synchronize()


def time_applies(op, x, y, steps, warmup, stream):
    import torch

    prewarm(op, x, y)
    for _ in range(warmup):
        op.apply(x, y)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for _ in range(steps):
        op.apply(x, y)
    ev[1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return ev[0].elapsed_time(ev[1]) * 1e-3 / steps, wall / steps


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh, nodes, lam_Q, mu_Q, budget_s=30.0, max_reps=100):
    """The CPU FP64 oracle (element-centric restatement of tandem's apply) on the same mesh,
    geometry, coefficients and input vector as the GPU apply, timed as app/static.cpp:94-121 does:
    one warm-up apply, then up to max_reps applies (as many as fit in budget_s of wall time) on the
    host's cores, mean and min; plus one apply on a single thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    threads = min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    om = O.Mesh.from_arrays(mesh.vertices, mesh.elements, mesh.boundary_facets, mesh.boundary_bc)
    op = O.Operator(0, om, 4, lam_Q, mu_Q, nodes=nodes)
    log("cpu baseline: oracle set-up %.1fs" % (time.perf_counter() - t0))
    x = np.random.default_rng(SEED).uniform(-1, 1, op.n)
    op.apply(x, threads)
    times = []
    t_start = time.perf_counter()
    while len(times) < max_reps and (len(times) < 3 or time.perf_counter() - t_start < budget_s):
        t0 = time.perf_counter()
        op.apply(x, threads)
        times.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    op.apply(x, 1)
    t1 = time.perf_counter() - t0
    tm = float(np.mean(times))
    return {
        "value": op.n / tm * 1e-9,
        "unit": "GDoF/s",
        "cores": threads,
        "kind": "port",
        "sample": "CPU FP64 oracle apply on the full workload (%d elements, %d DoF, same x): 1 warm-up apply, then %d "
                  "applies (as many as fit in %.0f s, at most %d) on %d OpenMP threads of %s: mean %.3f s "
                  "(%.4f GDoF/s), min %.3f s (%.4f GDoF/s); 1 thread, one apply: %.3f s (%.5f GDoF/s)"
                  % (op.E, op.n, len(times), budget_s, max_reps, threads, cpu_model(), tm, op.n / tm * 1e-9,
                     min(times), op.n / min(times) * 1e-9, t1, op.n / t1 * 1e-9),
        "mean_s": tm,
        "min_s": float(min(times)),
        "applies": len(times),
        "one_thread_s": t1,
        "cpu_model": cpu_model(),
    }


def traffic_from_profile():
    """HBM bytes per FP64 apply (both kernels) from the committed counter summary, if it covers them."""
    p = os.path.join(ROOT, "profiles", "elasticity_split_p4_box24_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_apply") if d.get("kernels") == KERNELS else None


def solve_comparison(op, mesh, nodes, fields, rtol=1e-10):
    """Time to a 1e-10 FP64 true relative residual: FP64 CG vs FP32 CG with FP64 residual
    replacement, each without and with the element-block Jacobi preconditioner, and FP64 flexible CG
    with a p-multigrid V-cycle (AMG on its coarsest level), on the manufactured cosine solution
    (right-hand side built on the GPU)."""
    import torch
    from tandem_amd import scenario as S

    lam, mu = fields
    X = op.volume_points()
    f = S.elasticity_body_force(X, lam, mu)               # (E, Nq, 3)
    Xf = op.facet_points()
    g, _ = S.cosine_solution(Xf)                            # (F, nq, 3)
    bd = torch.empty(op.n, dtype=torch.float64, device="cuda")
    fd = torch.from_numpy(np.ascontiguousarray(f.transpose(0, 2, 1)).ravel()).cuda()
    gd = torch.from_numpy(np.ascontiguousarray(g.transpose(0, 2, 1)).ravel()).cuda()
    op.rhs_device(bd, fd, gd)
    torch.cuda.synchronize()  # fd, gd stay alive until the right-hand side is built
    del fd, gd
    out = {}
    for pc in (None, "block-jacobi"):
        tag = "none" if pc is None else pc
        if pc is not None:
            t0 = time.perf_counter()
            op.block_jacobi_build()
            torch.cuda.synchronize()
            out["block_jacobi_setup_s"] = time.perf_counter() - t0
        xs = {}
        # with block Jacobi the FP64 solve runs twice: with the FP64 inverse blocks, and with the
        # FP32-stored blocks that the FP32 solve applies (the same preconditioner data for both; the
        # speedup is taken against this one)
        runs = [("fp64", pc), ("fp32-ru", pc)] if pc is None else [("fp64", pc), ("fp64", "block-jacobi-f32"),
                                                                   ("fp32-ru", pc)]
        for mode, p in runs:
            key = "%s/%s" % (mode, tag if p == pc else p)
            x = torch.zeros_like(bd)
            r = op.solve(bd, x, rtol=rtol, mode=mode, max_iter=40000, precond=p)
            log("solve", key, r)
            out[key] = {"seconds": r["seconds"], "iterations": r["iterations"], "fp64_replacements": r["outer_iterations"],
                        "rel_residual": r["rel_residual"], "converged": bool(r["converged"])}
            xs[key] = x
        ref = "fp64/" + tag if pc is None else "fp64/block-jacobi-f32"
        out["speedup_fp32ru_vs_fp64/" + tag] = out[ref]["seconds"] / out["fp32-ru/" + tag]["seconds"]
        if pc is not None:
            out["speedup_fp32ru_vs_fp64_fp64blocks/" + tag] = (out["fp64/" + tag]["seconds"] /
                                                                out["fp32-ru/" + tag]["seconds"])
        out["solution_rel_diff_fp32ru_vs_fp64/" + tag] = float(torch.linalg.norm(xs["fp32-ru/" + tag] - xs[ref]) /
                                                               torch.linalg.norm(xs[ref]))
        if pc is None:
            x_plain = xs["fp64/none"]
    rec, xp = pmg_solve(op, mesh, fields, bd, rtol)
    out["pmg-fp32vcycle/amg"] = rec
    out["solution_rel_diff_pmg_vs_fp64/none"] = float(torch.linalg.norm(xp - x_plain) / torch.linalg.norm(x_plain))
    out["speedup_pmg_vs_fp64/none"] = out["fp64/none"]["seconds"] / rec["seconds"]
    # the same hierarchy and preconditioners with the V-cycle (and its coarse AMG-CG) in FP64
    rec64, _ = pmg_solve(op, mesh, fields, bd, rtol, fp32=False)
    out["pmg-fp64vcycle/amg"] = rec64
    out["speedup_pmg_fp32vcycle_vs_fp64vcycle"] = rec64["seconds"] / rec["seconds"]
    return out


def pmg_solve(op, mesh, fields, b, rtol, fp32=True):
    """p-multigrid: FP64 flexible CG with an FP32 V-cycle over P4 -> P2 -> P1 (coarse levels on the
    same mesh, warped geometry and material, their own facet penalties), the P1 level's CG
    preconditioned by smoothed-aggregation AMG on its assembled block rows (with fp32=False the
    V-cycle and that coarse solve run in FP64).  Returns the record and the solution."""
    import torch
    import tandem_amd as T
    from tandem_amd import scenario as S

    lam, mu = fields
    t0 = time.perf_counter()
    levels = [op]
    for p in (2, 1):
        c = T.DGOperator(mesh, T.ELASTICITY, p, S.warp(mesh.affine_nodes(p)))
        Xc = c.volume_points()
        c.set_material(lam(Xc), mu(Xc))
        levels.append(c)
    M = T.PMultigrid(levels, galerkin_penalty=0, coarse_fp32=int(fp32), vcycle_fp32=int(fp32))
    M.setup()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    B = b.view(1, -1).clone()
    Xp = torch.zeros_like(B)
    r = M.solve_multi(B, Xp, rtol=rtol, max_iter=1000)
    log("solve pmg/amg %s V-cycle" % ("FP32" if fp32 else "FP64"),
        {k: r[k] for k in ("iterations", "outer_iterations", "rel_residual", "seconds", "converged")})
    rec = {"seconds": r["seconds"], "setup_seconds": t_setup, "iterations": int(r["iterations"]),
           "coarse_cg_iterations": int(r["outer_iterations"]), "levels": [4, 2, 1],
           "rel_residual": float(r["rel_residual"]), "converged": bool(r["converged"])}
    return rec, Xp[0]


def _boundary_facets(op):
    f = op.facets()
    return np.nonzero(f[:, 0] == f[:, 1])[0]


def run_precision(args):
    """FP32 MFMA + FP64 residual replacement vs all-FP64 solve on the P4 tet mesh at cells^3 (the
    precision study: 40^3 = 384,000 elements by default), manufactured cosine solution with the
    seeded variable Lame fields, CG to 1e-10 (block Jacobi, or none with --precond none).  Errors of the FP32-MFMA
    solution against the FP64 solution (whose apply matches the CPU FP64 oracle to 1e-12 in the
    parity tests) for the solution and for the traction on the boundary faces, and both solutions'
    errors against the manufactured solution."""
    import torch
    from tandem_amd import scenario as S

    t0 = time.perf_counter()
    op, mesh, nodes, (lam, mu) = build_workload(args.cells)
    on_stream(op)
    op.upload()
    X = op.volume_points()
    f = S.elasticity_body_force(X, lam, mu)                       # (E, Nq, 3)
    Xf = op.facet_points()
    g, _ = S.cosine_solution(Xf)                                  # (F, nq, 3)
    force = torch.from_numpy(np.ascontiguousarray(f.transpose(0, 2, 1)).ravel()).cuda()
    gdat = torch.from_numpy(np.ascontiguousarray(g.transpose(0, 2, 1)).ravel()).cuda()
    b = torch.empty(op.n, dtype=torch.float64, device="cuda")
    op.rhs_device(b, force, gdat)
    del force
    log("precision: setup %.1fs, %d elements, %d DoF" % (time.perf_counter() - t0, op.E, op.n))
    pc = None if args.precond == "none" else "block-jacobi"
    out = {"workload": "P4 tet elasticity, %d^3 Kuhn cubes x 6 = %d elements, %d DoF" % (args.cells, op.E, op.n),
           "preconditioner": args.precond}
    if pc:
        t1 = time.perf_counter()
        out["colours"] = op.block_jacobi_build()
        torch.cuda.synchronize()
        out["block_jacobi_setup_s"] = time.perf_counter() - t1
    xs = {}
    # with block Jacobi the FP64 solve also runs with the FP32-stored blocks the FP32 solve applies
    runs = [("fp64", pc), ("fp32-ru", pc)] + ([("fp64", "block-jacobi-f32")] if pc else [])
    for mode, p in runs:
        key = mode if p == pc else mode + "/fp32-blocks"
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        r = op.solve(b, x, rtol=1e-10, mode=mode, precond=p, max_iter=100000)
        log("precision solve", key, r)
        out[key] = {k: r[k] for k in ("seconds", "iterations", "outer_iterations", "rel_residual", "converged")}
        xs[key] = x
    out["speedup_fp32ru_vs_fp64"] = out["fp64"]["seconds"] / out["fp32-ru"]["seconds"]
    if pc:
        out["speedup_fp32ru_vs_fp64_fp32blocks"] = out["fp64/fp32-blocks"]["seconds"] / out["fp32-ru"]["seconds"]
    # the same problem by p-multigrid (FP64 flexible CG, FP32 V-cycle, AMG on the P1 level)
    out["pmg"], xs["pmg"] = pmg_solve(op, mesh, (lam, mu), b, 1e-10)
    x64, x32 = xs["fp64"], xs["fp32-ru"]
    out["solution_rel_l2_fp32ru_vs_fp64"] = float(torch.linalg.norm(x32 - x64) / torch.linalg.norm(x64))
    # traction on the boundary faces
    bf = torch.from_numpy(_boundary_facets(op)).cuda()
    t64 = op.traction(x64, gdat)[bf]
    t32 = op.traction(x32, gdat)[bf]
    out["boundary_traction_rel_l2_fp32ru_vs_fp64"] = float(torch.linalg.norm(t32 - t64) / torch.linalg.norm(t64))
    # errors against the manufactured solution at the volume points (L2 with W |det J|)
    EQ = op.host_array("E_Q").reshape(op.Nq, -1).T               # (Nbf, Nq)
    wj = op.host_array("absDetJ").reshape(op.E, op.Nq) * op.host_array("volWeights")[None, :]
    ue, _ = S.cosine_solution(X)                                  # (E, Nq, 3)
    for mode, x in xs.items():
        U = x.view(op.E, 3, -1).cpu().numpy()                      # (E, 3, Nbf)
        uh = np.einsum("epk,kq->eqp", U, EQ)
        out[mode]["l2_error_vs_manufactured"] = float(np.sqrt(np.sum(wj[..., None] * (uh - ue) ** 2) /
                                                              np.sum(wj[..., None] * ue ** 2)))
    return out


def run_greens(args):
    """Discrete Green's function of a 2D plane-strain fault problem (P6 triangles, ~200k elements,
    512 fault facets x 7 slip basis functions = 3,584 right-hand sides): a block of 64 unit fault
    states turned into facet slip (fault adapter) and right-hand sides on the GPU, one batched
    launch each, advanced together by lockstep CG
    (one batched apply of 64 columns per iteration), FP32 MFMA + FP64 residual replacement, block
    Jacobi, against one right-hand side alone with the same solver.  Both run a fixed budget of
    args.iters iterations (block-Jacobi CG needs 10^4-10^5 iterations to reach 1e-10 on this mesh,
    more than a bounded run affords), so the comparison is throughput: right-hand-side iterations
    per second, and the residual each reached.  The 64 Green's-function columns (fault traction
    in the fault basis) are then computed in one batched launch."""
    import torch
    import tandem_amd as T
    from tandem_amd import scenario as S

    t0 = time.perf_counter()
    fr = S.FaultRectangle()
    mesh = fr.mesh()
    op = on_stream(T.DGOperator(mesh, T.ELASTICITY, 6))
    X = op.volume_points()
    lf, mf = S.lame_fields(2, SEED + 3, length=50.0)
    op.set_material(lf(X), mf(X))
    op.upload()
    ncol = op.block_jacobi_build()
    torch.cuda.synchronize()
    log("greens: setup %.1fs, %d elements, %d DoF, %d colours" % (time.perf_counter() - t0, op.E, op.n, ncol))
    total_cols = int(np.sum(op.facets()[:, 4] == 3)) * (op.degree + 1)
    nb = 64
    ld = op.n
    variant, ms = op.autotune(ncols=nb, precision="fp32")
    log("greens: autotune (64 columns, FP32)", variant, ms)
    out = {"workload": "2D P6 plane strain, %d elements, %d DoF, %d slip right-hand sides, blocks of %d"
                       % (op.E, op.n, total_cols, nb),
           "autotune_fp32_64col_ms": ms, "autotune_choice": variant, "iterations_budget": args.iters}
    budget = not args.greens_pmg_only
    if budget:
        # batched apply vs column-by-column apply (FP32)
        Xf = torch.randn((nb, ld), dtype=torch.float32, device="cuda")
        Yf = torch.empty_like(Xf)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        op.apply_batched(Xf, Yf)
        ev[0].record()
        for _ in range(5):
            op.apply_batched(Xf, Yf)
        ev[1].record()
        for _ in range(5):
            for c in range(nb):
                op.apply(Xf[c], Yf[c])
        ev[2].record()
        torch.cuda.synchronize()
        t_b, t_s = ev[0].elapsed_time(ev[1]) / 5, ev[1].elapsed_time(ev[2]) / 5
        out["apply_fp32_64col_batched_ms"] = t_b
        out["apply_fp32_64col_one_by_one_ms"] = t_s
        out["apply_batching_speedup"] = t_s / t_b
        del Xf, Yf
    # the slip basis: unit fault states in the fault adapter's basis (nodal P6 per fault facet,
    # along-fault component; up = +y, fault normal oriented to the hanging wall)
    dip = np.radians(30.0)
    nfault, nbf = op.fault_setup([0.0, 1.0], [np.sin(dip), np.cos(dip)])
    out["green_function_columns"] = nfault * nbf
    S = torch.zeros((nb, nfault * nbf), dtype=torch.float64, device="cuda")
    S[torch.arange(nb), torch.arange(nb)] = 1.0
    B = torch.empty((nb, ld), dtype=torch.float64, device="cuda")
    Xs = torch.zeros_like(B)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    fdat = op.fault_slip(S)
    op.rhs_device(B, None, fdat)
    G = None
    if budget:
        G = greens_budget_runs(op, B, Xs, fdat, nb, args, out, t1)
    if args.pmg_levels != "none":
        greens_pmg(op, mesh, lf, mf, B, fdat, G, args, out)
    return out


def greens_budget_runs(op, B, Xs, fdat, nb, args, out, t1):
    """The fixed-budget lockstep block-Jacobi CG of the 64-column block and of one column alone;
    returns the Green's-function columns of the lockstep solutions."""
    import torch

    r = op.solve_multi(B, Xs, rtol=1e-10, mode="fp32-ru", precond="block-jacobi", max_iter=args.iters)
    torch.cuda.synchronize()
    t_lock = time.perf_counter() - t1
    log("greens lockstep %.2fs" % t_lock, {k: r[k] for k in ("iterations", "rel_residual")})
    out["lockstep_seconds"] = t_lock
    out["lockstep_iterations"] = int(r["iterations"])
    out["lockstep_worst_rel_residual"] = float(np.max(r["column_rel_residual"]))
    out["lockstep_rhs_iterations_per_second"] = nb * r["iterations"] / t_lock
    # the Green's function columns: fault traction of the 64 solutions in the fault basis
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    G = op.fault_traction(Xs, fdat)
    torch.cuda.synchronize()
    out["fault_traction_64col_ms"] = (time.perf_counter() - t2) * 1e3
    out["green_block_shape"] = list(G.shape)
    # one right-hand side alone, same solver and budget
    b = B[0].clone()
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    r1 = op.solve(b, x, rtol=1e-10, mode="fp32-ru", precond="block-jacobi", max_iter=args.iters)
    torch.cuda.synchronize()
    t_one = time.perf_counter() - t1
    log("greens single column %.2fs" % t_one, {k: r1[k] for k in ("iterations", "rel_residual")})
    out["single_seconds"] = t_one
    out["single_iterations"] = int(r1["iterations"])
    out["single_rel_residual"] = float(r1["rel_residual"])
    out["single_rhs_iterations_per_second"] = r1["iterations"] / t_one
    out["batching_speedup_per_rhs_iteration"] = out["lockstep_rhs_iterations_per_second"] / out[
        "single_rhs_iterations_per_second"]
    xs = Xs[0].cpu().numpy()
    out["column0_rel_diff_lockstep_vs_single"] = float(np.linalg.norm(xs - x.cpu().numpy()) / np.linalg.norm(xs))
    return G


def greens_pmg(op, mesh, lf, mf, B, fdat, G, args, out):
    """The block of 64 right-hand sides solved to 1e-10 by flexible CG with a p-multigrid V-cycle
    (coarse levels on the same mesh and material)."""
    import torch
    import tandem_amd as T

    degs = [int(v) for v in args.pmg_levels.split(",")]
    if degs[0] != op.degree:
        raise ValueError("--pmg-levels must start at the fine degree %d" % op.degree)
    t0 = time.perf_counter()
    coarse = []
    for p in degs[1:]:
        c = on_stream(T.DGOperator(mesh, T.ELASTICITY, p))
        Xc = c.volume_points()
        c.set_material(lf(Xc), mf(Xc))
        coarse.append(c)
    M = T.PMultigrid([op] + coarse, smooth=args.pmg_smooth, coarse_max_iter=args.pmg_coarse_max_iter,
                     galerkin_penalty=args.pmg_galerkin)
    lmax = M.setup()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    log("greens p-multigrid %s: setup %.1fs, lambda_max %s" % (degs, t_setup, lmax))
    Xp = torch.zeros_like(B)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rp = M.solve_multi(B, Xp, rtol=1e-10, max_iter=args.pmg_max_iter)
    torch.cuda.synchronize()
    t_p = time.perf_counter() - t1
    log("greens p-multigrid solve %.2fs" % t_p, {k: rp[k] for k in ("iterations", "outer_iterations",
                                                                      "rel_residual", "converged")})
    Gp = op.fault_traction(Xp, fdat)
    torch.cuda.synchronize()
    out["pmg"] = {
        "levels": degs, "smooth": M.options["smooth"], "coarse_rtol": M.options["coarse_rtol"],
        "coarse_max_iter": M.options["coarse_max_iter"], "galerkin_penalty": M.options["galerkin_penalty"],
        "eig_interval": [M.options["eig_lo"], M.options["eig_hi"]], "lambda_max": [float(v) for v in lmax],
        "setup_seconds": t_setup, "seconds_to_1e-10": t_p, "iterations": int(rp["iterations"]),
        "coarse_cg_iterations": int(rp["outer_iterations"]), "converged": int(rp["converged"]),
        "worst_rel_residual": float(rp["rel_residual"]),
        "max_column_iterations": int(np.max(rp["column_iterations"])),
    }
    if G is not None:
        out["pmg"]["green_columns_rel_diff_vs_lockstep"] = float(torch.linalg.norm(Gp - G) / torch.linalg.norm(Gp))


class SurfaceSampler:
    """B: the displacement at a seeded subset of the facet quadrature points of the top face z = 1
    (the sample points are facet quadrature points, so B is the facet trace of the element basis)."""

    def __init__(self, op, count, seed):
        fct = op.facets()
        Xf = op.facet_points()
        top = np.nonzero((fct[:, 0] == fct[:, 1]) & np.all(np.abs(Xf[:, :, 2] - 1.0) < 1e-12, axis=1))[0]
        rng = np.random.default_rng(seed)
        pick = rng.choice(top.size * op.nq, size=count, replace=False)
        self.f, self.q = top[pick // op.nq], pick % op.nq
        self.el, self.loc = fct[self.f, 0], fct[self.f, 2]
        tabs = [op.host_array("E_q%d" % l).reshape(op.nq, -1) for l in range(op.dim + 1)]   # (nq, Nbf)
        self.phi = np.stack([tabs[l][q] for l, q in zip(self.loc, self.q)])                # (count, Nbf)
        self.op = op

    def apply(self, x):
        U = x.view(self.op.E, 3, -1).cpu().numpy()[self.el]                                  # (count, 3, Nbf)
        return np.einsum("snk,sk->sn", U, self.phi)

    def transpose(self, r):
        import torch
        B = np.zeros((self.op.E, 3, self.phi.shape[1]))
        np.add.at(B, self.el, r[:, :, None] * self.phi[:, None, :])
        return torch.from_numpy(B.reshape(-1)).cuda()


def run_multimesh(args):
    """32 independent fault-rectangle meshes (2D plane strain, P6 triangles) with element counts
    log-uniform in [1000, 64000] from a seed, each with a unit along-fault slip right-hand side:
    all meshes held in one operator (their disjoint union) and solved as 32 CG recursions in
    lockstep, one apply launch over every mesh's element groups per iteration (FP32 MFMA + FP64
    residual replacement, block Jacobi, towards 1e-10 within args.iters iterations, each segment
    stopping on its own); compared with the same 32 solves one mesh at a time with an operator per
    mesh and the same budget.  Every operator runs the FP32 kernel variant its autotuner picks."""
    import torch
    import tandem_amd as T
    from tandem_amd import scenario as S

    rng = np.random.default_rng(SEED + 11)
    counts = np.exp(rng.uniform(np.log(1000.0), np.log(64000.0), args.meshes)).astype(int)
    t0 = time.perf_counter()
    frs = [S.FaultRectangle.with_elements(int(n), SEED + 100 + i) for i, n in enumerate(counts)]
    meshes = [fr.mesh() for fr in frs]
    union, eoff = T.Mesh.union(meshes)
    op = on_stream(T.DGOperator(union, T.ELASTICITY, 6))
    lf, mf = S.lame_fields(2, SEED + 3, length=50.0)
    X = op.volume_points()
    op.set_material(lf(X), mf(X))
    op.upload()
    op.block_jacobi_build()
    variant, _ = op.autotune(ncols=1, precision="fp32")   # the union's fastest FP32 kernel
    bs = op.bs
    # right-hand side: each mesh's uniform slip, placed in its segment
    parts = []
    for i, fr in enumerate(frs):
        sub = T.DGOperator(meshes[i], T.ELASTICITY, 6)
        parts.append(fr.uniform_slip(sub))
    fdat = torch.from_numpy(np.concatenate(parts).reshape(-1)).cuda()
    assert fdat.numel() == op.F * 2 * op.nq, "facets of the union are the meshes' facets in order"
    b = torch.empty(op.n, dtype=torch.float64, device="cuda")
    op.rhs_device(b, None, fdat)
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    log("multimesh: setup %.1fs, %d meshes, %d elements in all, %d DoF" % (time.perf_counter() - t0, len(frs),
                                                                           op.E, op.n))
    offsets = eoff * bs
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    r = op.solve_segments(b, x, offsets, rtol=1e-10, mode="fp32-ru", precond="block-jacobi", max_iter=args.iters)
    torch.cuda.synchronize()
    t_lock = time.perf_counter() - t1
    log("multimesh lockstep %.2fs" % t_lock, {k: r[k] for k in ("iterations", "rel_residual", "converged")})
    out = {"workload": "%d fault-rectangle meshes, 2D P6 plane strain, %d..%d elements (%d in all, %d DoF), "
                       "one unit-slip right-hand side each" % (len(frs), counts.min(), counts.max(), op.E, op.n),
           "iterations_budget": args.iters,
           "element_counts": [int(m.num_elements) for m in meshes],
           "union_kernel_variant": variant,
           "lockstep_seconds": t_lock, "lockstep_iterations": int(r["iterations"]),
           "lockstep_seconds_per_iteration": t_lock / max(1, int(r["iterations"])),
           "segments_converged": int(np.sum(r["segment_rel_residual"] <= 1e-10)),
           "worst_rel_residual": float(np.max(r["segment_rel_residual"]))}
    # the same solves one mesh at a time
    seq, diffs, seq_its = 0.0, [], 0
    xs = x.cpu().numpy()
    for i, m in enumerate(meshes):
        sub = on_stream(T.DGOperator(m, T.ELASTICITY, 6))
        Xi = sub.volume_points()
        sub.set_material(lf(Xi), mf(Xi))
        sub.upload()
        sub.block_jacobi_build()
        sub.autotune(ncols=1, precision="fp32")            # each mesh's fastest FP32 kernel too
        bi = torch.empty(sub.n, dtype=torch.float64, device="cuda")
        sub.rhs_device(bi, None, torch.from_numpy(parts[i].reshape(-1)).cuda())
        xi = torch.zeros_like(bi)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        ri = sub.solve(bi, xi, rtol=1e-10, mode="fp32-ru", precond="block-jacobi", max_iter=args.iters)
        torch.cuda.synchronize()
        seq += time.perf_counter() - t2
        seq_its += int(ri["iterations"])
        ref = xi.cpu().numpy()
        seg = xs[offsets[i]:offsets[i + 1]]
        diffs.append(float(np.linalg.norm(seg - ref) / max(np.linalg.norm(ref), 1e-300)))
        del sub, bi, xi
    out["one_at_a_time_seconds"] = seq
    out["one_at_a_time_iterations"] = seq_its
    out["speedup_lockstep_vs_one_at_a_time"] = seq / t_lock
    out["max_solution_rel_diff_vs_one_at_a_time"] = max(diffs)
    return out


def run_poisson(args):
    """2D Poisson P2 on the unit square split into 32 x 32 x 2 triangles (2,048 elements), K = 1,
    u = sin(pi x) sin(pi y) with zero Dirichlet data: one apply on a seeded uniform[-1, 1] vector and
    a CG solve to 1e-10, on the GPU and with the CPU FP64 oracle (apply: warm-up then 100 applies,
    mean and min as in tandem's static.cpp; solve: scipy CG on the oracle's apply)."""
    import torch
    import tandem_amd as T
    from tandem_amd import scenario as S
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from scipy.sparse.linalg import LinearOperator, cg

    mesh = T.Mesh.box(2, 32, 0, T.BC_DIRICHLET)
    op = on_stream(T.Poisson(mesh, 2, K=1.0))
    X = op.volume_points()
    _, f = S.poisson_sine(X)
    b = op.rhs(f[:, None, :], np.zeros((op.F, 1, op.nq)))
    x = torch.from_numpy(np.random.default_rng(SEED).uniform(-1, 1, op.n)).cuda()
    y = torch.zeros_like(x)
    stream = torch.cuda.current_stream()
    t_ev, _ = time_applies(op, x, y, 100, 10, stream)
    bd = torch.from_numpy(b).cuda()
    xs = torch.zeros_like(bd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = op.solve(bd, xs, rtol=1e-10, mode="fp64")
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    # L2 error against the manufactured solution at the volume points
    U = xs.view(op.E, -1).cpu().numpy()                                   # (E, Nbf) modal
    EQ = op.host_array("E_Q").reshape(op.Nq, -1).T                        # (Nbf, Nq)
    wj = op.host_array("absDetJ").reshape(op.E, op.Nq) * op.host_array("volWeights")[None, :]
    ue, _ = S.poisson_sine(X)
    uh = U @ EQ
    err = float(np.sqrt(np.sum(wj * (uh - ue) ** 2) / np.sum(wj * ue ** 2)))
    out = {"workload": "Poisson P2, 32x32x2 triangles, %d elements, %d DoF" % (op.E, op.n),
           "gpu_apply_us": t_ev * 1e6, "gpu_apply_gdofs": op.n / t_ev * 1e-9,
           "gpu_solve_s": t_solve, "gpu_solve_iterations": int(r["iterations"]),
           "gpu_rel_residual": float(r["rel_residual"]), "l2_error_vs_manufactured": err}
    # CPU oracle
    om = O.Mesh.box(2, [32, 32], 0, [5] * 4)
    vc, _ = om.coords(2)
    oo = O.Operator(1, om, 2, np.ones(vc.shape[:2]), None)
    threads = min(os.cpu_count() or 1, 16)
    xv = x.cpu().numpy()
    for _ in range(10):
        oo.apply(xv, threads)
    ts = []
    for _ in range(100):
        t0 = time.perf_counter()
        oo.apply(xv, threads)
        ts.append(time.perf_counter() - t0)
    out["cpu_apply_us_mean"] = float(np.mean(ts)) * 1e6
    out["cpu_apply_us_min"] = float(np.min(ts)) * 1e6
    out["cpu_threads"] = threads
    A = LinearOperator((oo.n, oo.n), matvec=lambda v: oo.apply(np.ascontiguousarray(v), threads))
    t0 = time.perf_counter()
    its = [0]
    xc, info = cg(A, b, rtol=1e-10, maxiter=100000, callback=lambda _: its.__setitem__(0, its[0] + 1))
    out["cpu_solve_s"] = time.perf_counter() - t0
    out["cpu_solve_iterations"] = its[0]
    out["cpu_vs_gpu_solution_rel_diff"] = float(np.linalg.norm(xc - xs.cpu().numpy()) / np.linalg.norm(xc))
    return out


def run_adjoint(args):
    """Adjoint gradient of the misfit J(lam, mu) = 1/2 |B u - d|^2 on the P4 tet mesh (the 24^3-cube apply workload by
    default): B samples the displacement at 2,048 seeded points of the free surface z = 1 (Dirichlet
    zero on the other faces, seeded smooth body force), d comes from Lame fields perturbed by a
    seeded 10% smooth field.  Gradient per volume quadrature point from the adjoint kernel, checked
    (i) against Richardson-extrapolated central differences of psi^T A(m) u built from the FP64
    apply along 16 seeded smooth directions and (ii) against central differences of J itself with
    FP64 solves along 2 of them (at this size limited by the FP64 solver: the true residual of
    b - A x bottoms out near 2e-9, so J carries noise of that order; the exact-solve comparison
    is the small-mesh GPU test); then the FP64 and FP32 gradient kernels are timed."""
    import torch
    import tandem_amd as T
    from tandem_amd import scenario as S

    t0 = time.perf_counter()
    mesh = T.Mesh.box(3, args.cells, scheme=1, bc=[5, 5, 5, 5, 5, 1])
    nodes = S.warp(mesh.affine_nodes(4))
    op = on_stream(T.DGOperator(mesh, T.ELASTICITY, 4, nodes))
    X = op.volume_points()
    lf, mf = S.lame_fields(3, SEED)
    lam, mu = lf(X), mf(X)
    pert = S.LogNormalField(3, 1.0, 1.0, 0.3, SEED + 7)
    ptrue = 0.1 * np.tanh(np.log(pert(X)))                      # smooth seeded field, |.| < 0.1
    force = np.zeros((op.E, 3, op.Nq))
    force[:, 2, :] = -1.0 + 0.5 * np.sin(2 * np.pi * X[..., 0]) * np.cos(2 * np.pi * X[..., 1])
    force[:, 0, :] = 0.3 * np.cos(3 * X[..., 2])
    fdev = torch.from_numpy(force.ravel()).cuda()
    b = torch.empty(op.n, dtype=torch.float64, device="cuda")
    sampler = SurfaceSampler(op, 2048, SEED + 11)

    # FP64 CG run until its recursive residual is 1e-13: the true residual then sits at the FP64
    # floor of b - A x for this operator (about 2e-9 here), the most accurate solution FP64 gives
    def forward(lq, mq, rtol):
        op.set_material(lq, mq)
        op.rhs_device(b, fdev, None)
        x = torch.zeros_like(b)
        r = op.solve(b, x, rtol=rtol, mode="fp64", precond="block-jacobi", max_iter=40000)
        log("adjoint: forward solve", {k: r[k] for k in ("iterations", "rel_residual", "seconds")})
        if not r["rel_residual"] <= 1e-8:
            raise RuntimeError("forward solve did not converge: %s" % r)
        return x

    d = sampler.apply(forward(lam * (1 + ptrue), mu * (1 + ptrue), 1e-13))
    log("adjoint: setup + data %.1fs, %d elements" % (time.perf_counter() - t0, op.E))
    u = forward(lam, mu, 1e-13)
    res = sampler.apply(u) - d
    J = 0.5 * float(np.sum(res ** 2))
    psi = torch.zeros_like(u)
    rhs_adj = sampler.transpose(res)
    ra = op.solve(rhs_adj, psi, rtol=1e-13, mode="fp64", precond="block-jacobi", max_iter=40000)
    log("adjoint: adjoint solve", {k: ra[k] for k in ("iterations", "rel_residual", "seconds")})
    out_res = {"adjoint_rel_residual": ra["rel_residual"]}
    gl, gm = op.material_gradient(u, psi)
    g_lam, g_mu = -gl.cpu().numpy(), -gm.cpu().numpy()
    out = {"workload": "P4 tet elasticity, %d^3 Kuhn cubes x 6 = %d elements, 2048 surface samples" % (args.cells, op.E),
           "J": J, "adjoint_iterations": ra["iterations"]}
    out.update(out_res)
    rng = np.random.default_rng(SEED + 13)
    # (i) directional derivatives of psi^T A(m) u from the FP64 apply
    ut, pt = u.clone(), psi.clone()
    y = torch.empty_like(u)
    errs = []
    def smooth_direction(X):
        k = rng.normal(size=(3, 3)) * 2.0
        ph = rng.uniform(0, 2 * np.pi, size=3)
        return np.tanh(np.cos(X @ k.T + ph) @ rng.uniform(0.3, 1.0, size=3))

    def V(l, m):
        op.set_material(l, m)
        op.apply(ut, y)
        return float(torch.dot(pt, y))

    for k in range(16):
        # seeded smooth relative directions: a step of 1e-4 moves no penalty extremal node
        dl = smooth_direction(X) * lam
        dm = smooth_direction(X) * mu
        an = float(np.sum(g_lam * dl) + np.sum(g_mu * dm))
        h = 1e-4
        c1 = (V(lam + h * dl, mu + h * dm) - V(lam - h * dl, mu - h * dm)) / (2 * h)
        c2 = (V(lam + h / 2 * dl, mu + h / 2 * dm) - V(lam - h / 2 * dl, mu - h / 2 * dm)) / h
        fd = -(4 * c2 - c1) / 3
        errs.append(abs(fd - an) / abs(an))
    out["apply_fd_directions"] = 16
    out["apply_fd_max_rel_err"] = max(errs)
    # (ii) finite differences of J with FP64 solves
    jerr = []
    for k in range(2):
        dl = smooth_direction(X) * lam
        dm = smooth_direction(X) * mu
        an = float(np.sum(g_lam * dl) + np.sum(g_mu * dm))
        Js = []
        h = 1e-3  # Richardson-extrapolated central differences at h and h/2 (error O(h^4))
        for eps in (h, -h, h / 2, -h / 2):
            xs = forward(lam + eps * dl, mu + eps * dm, 1e-13)
            Js.append(0.5 * float(np.sum((sampler.apply(xs) - d) ** 2)))
        c1 = (Js[0] - Js[1]) / (2 * h)
        c2 = (Js[2] - Js[3]) / h
        fd = (4.0 * c2 - c1) / 3.0
        jerr.append(abs(fd - an) / abs(an))
        log("adjoint J-FD direction", k, fd, an)
    out["J_fd_directions"] = 2
    out["J_fd_max_rel_err"] = max(jerr)
    op.set_material(lam, mu)
    time_gradients(op, u, psi, out, 10)
    return out


def time_gradients(op, u, psi, out, reps):
    """Times the gradient kernels, matrix-core ("mfma", the default) and thread-per-point, in FP64 and
    FP32 (HIP events on the operator's stream, torch's current one), and records the FP32 gradients'
    error against the FP64 matrix-core gradient and the algorithmic bandwidth of each."""
    import torch

    prewarm(op, u, torch.empty_like(u))
    ref = None
    for variant in ("auto", "points"):
        op.set_variant(variant)
        vname = "mfma" if variant == "auto" else variant
        for dt, name in ((torch.float64, "fp64"), (torch.float32, "fp32")):
            uu, pp = u.to(dt), psi.to(dt)
            gl_, gm_ = op.material_gradient(uu, pp)
            torch.cuda.synchronize()
            cat = torch.cat([gl_.double(), gm_.double()])
            if ref is None:
                ref = cat
            else:
                out["gradient_rel_l2_%s_%s_vs_mfma_fp64" % (vname, name)] = float(torch.linalg.norm(cat - ref) /
                                                                                   torch.linalg.norm(ref))
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            for _ in range(reps):
                op.material_gradient(uu, pp)
            ev[1].record()
            torch.cuda.synchronize()
            ms = ev[0].elapsed_time(ev[1]) / reps
            out["gradient_kernel_ms_%s_%s" % (vname, name)] = ms
            # algorithmic bytes: u, psi, J^-1 and W|J| per volume point, M^-1 per element, facet G'
            # of both sides and the normal, the two outputs
            E, Nbf, Nq, nq = op.E, op.bs // 3, op.Nq, op.nq
            nbytes = dt.itemsize * (E * (2 * Nbf * 3 + Nq * 10 + Nbf * Nbf + 2 * Nq) + op.F * nq * (2 * 9 + 3))
            out["gradient_alg_gbs_%s_%s" % (vname, name)] = nbytes / (ms * 1e-3) / 1e9
    op.set_variant("auto")


def run_gradient(args):
    """The adjoint gradient kernels alone on the P4 tet apply workload (seeded random u and psi in
    [-1, 1]): one evaluation of each kernel and precision, then --steps timed evaluations."""
    import torch

    op, _, _, _ = build_workload(args.cells)
    on_stream(op)
    rng = np.random.default_rng(SEED + 17)
    u = torch.from_numpy(rng.uniform(-1.0, 1.0, op.n)).cuda()
    psi = torch.from_numpy(rng.uniform(-1.0, 1.0, op.n)).cuda()
    out = {"workload": "adjoint gradient kernels, P4 tets, %d^3 Kuhn cubes x 6 = %d elements" % (args.cells, op.E)}
    time_gradients(op, u, psi, out, args.steps)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=24)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default launch shapes")
    ap.add_argument("--workload", choices=["apply", "precision", "greens", "adjoint", "multimesh", "poisson", "gradient"],
                    default="apply",
                    help="apply: the headline line (default); precision: FP32-MFMA vs FP64 solve at --cells "
                         "(40 for the 384,000-element study); greens: batched Green's-function solves; "
                         "adjoint: misfit gradient with finite-difference checks; gradient: the gradient kernels alone")
    ap.add_argument("--blocks", type=int, default=2, help="greens: number of 64-column blocks timed")
    ap.add_argument("--pmg-levels", default="6,3,1",
                    help="greens: degrees of the p-multigrid levels for the solve to 1e-10 ('none': skip)")
    ap.add_argument("--pmg-smooth", type=int, default=6, help="greens: Chebyshev steps per smoothing")
    ap.add_argument("--pmg-coarse-max-iter", type=int, default=500,
                    help="greens: coarse CG iteration cap per V-cycle")
    ap.add_argument("--pmg-galerkin", type=int, default=0,
                    help="greens: 1 = coarse levels take the fine facet penalty, 0 = their own")
    ap.add_argument("--pmg-max-iter", type=int, default=300, help="greens: flexible-CG iteration cap")
    ap.add_argument("--greens-pmg-only", action="store_true",
                    help="greens: only the p-multigrid solve of the block (no fixed-budget runs)")
    ap.add_argument("--meshes", type=int, default=32, help="multimesh: number of independent meshes")
    ap.add_argument("--precond", choices=["block-jacobi", "none"], default="block-jacobi",
                    help="precision: preconditioner of both solves")
    ap.add_argument("--iters", type=int, default=2000, help="greens / multimesh: iteration budget of the solves")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the operator is not sharded across GPUs); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py: no GPU available", file=sys.stderr)
        sys.exit(2)
    if args.workload != "apply":
        os.environ.setdefault("TDG_PROGRESS_SECONDS", "30")  # the long solves report progress
        run = {"precision": run_precision, "greens": run_greens, "adjoint": run_adjoint,
               "multimesh": run_multimesh, "poisson": run_poisson, "gradient": run_gradient}[args.workload]
        with torch.cuda.stream(torch.cuda.Stream()):
            print(json.dumps(run(args)), flush=True)
        return
    t0 = time.perf_counter()
    op, mesh, nodes, fields = build_workload(args.cells)
    op.upload()
    log("setup %.1fs, %d elements, %d DoF" % (time.perf_counter() - t0, op.E, op.n))
    stream = torch.cuda.Stream()
    op.set_stream(stream)
    # the library's autotuner picks the launch shape of the two-kernel apply for each precision
    # (--no-autotune keeps the library's defaults, e.g. for counter runs of fixed kernels)
    shapes = {}
    for prec in ("fp64", "fp32"):
        if args.no_autotune:
            shapes[prec] = "library default"
            continue
        best, ms = op.autotune_shapes(1, prec)
        shapes[prec] = {"shape": best[0], "xcd_order": best[1], "ms": {"%d/%d" % k: v for k, v in ms.items()}}
        log("autotune %s: shape %d, workgroup order %d" % (prec, best[0], best[1]))
    # torch's own work (vector set-up, conversions, temporaries handed to the library) goes on the
    # operator's stream too: ordered with the library's kernels, and the caching allocator cannot
    # reuse a temporary while a library kernel on another stream still reads it
    torch.cuda.set_stream(stream)
    rng = np.random.default_rng(SEED)
    x = torch.from_numpy(rng.uniform(-1.0, 1.0, op.n)).cuda()
    y = torch.zeros_like(x)
    torch.cuda.synchronize()
    t_ev, t_wall = time_applies(op, x, y, args.steps, args.warmup, stream)
    xf, yf = x.float(), torch.zeros(op.n, dtype=torch.float32, device="cuda")
    t32, _ = time_applies(op, xf, yf, args.steps, args.warmup, stream)
    B64, F = op.apply_counts(8)
    B32, _ = op.apply_counts(4)
    achieved = F / t_ev
    roof = {
        "bound": "mfma",
        "achieved": achieved * 1e-12,
        "peak": PEAK_FP64 * 1e-12,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP64,
        "traffic": traffic_from_profile(),
        "kernel": " + ".join(KERNELS),
        "algorithmic_flops_per_launch": F,
        "algorithmic_bytes_per_launch": B64,
        "hbm_frac_algorithmic": B64 / t_ev / PEAK_HBM,
    }
    res = {
        "metric": METRIC,
        "value": op.n / t_ev * 1e-9,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_ev * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "P4 tet elasticity apply, [0,1]^3 %d^3 Kuhn cubes x 6 = %d elements, %d DoF, warped P4 "
                        "geometry, seeded log-normal lambda/mu" % (args.cells, op.E, op.n),
            "elements": op.E,
            "dofs": op.n,
            "degree": 4,
        },
        "wall_ms_per_step": t_wall * 1e3,
        "roofline": roof,
        "fp32_apply": {"value": op.n / t32 * 1e-9, "unit": "GDoF/s", "ms": t32 * 1e3,
                       "frac_fp32_peak": F / t32 / PEAK_FP32},
        "launch_shapes": shapes,
    }
    if not args.no_solve:
        res["solve"] = solve_comparison(op, mesh, nodes, fields)
    res["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline(mesh, nodes, *op.material_Q)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
