#!/usr/bin/env python3
"""Benchmark of the MI355X matrix-free DG operator on the P4 tetrahedral elasticity workload.

Workload: 3D linear elasticity, polynomial degree 4 on tetrahedra, [0,1]^3 split into 24^3 cubes of
6 Kuhn tetrahedra (82,944 elements, 35 x 3 basis functions each: 8,709,120 DoF), degree-4 geometry
warped by 0.02 sin(2 pi x) sin(2 pi y) sin(2 pi z), seeded log-normal lambda/mu (means 2 and 1,
sigma 0.3, correlation length 0.2), Dirichlet on all faces.  A step is one matrix-free apply
y = A x of a seeded uniform[-1,1] vector in FP64, inputs resident in HBM.

Prints one JSON line: GDoF/s of the FP64 apply (`value`), the roofline of the apply kernel measured
with HIP events on its stream, the CPU baseline (the CPU FP64 oracle timed on the host cores on a
smaller mesh of the same element type), the FP32 apply, and the time to a 1e-10 relative residual
of FP64 CG versus FP32 CG inside FP64 iterative refinement on the same mesh and body force.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-solve] [--no-cpu-baseline]
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
KERNEL = "elasticity_mfma_kernel<3, 4, double>"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(cells, degree=4, seed=SEED):
    import tandem_amd as T
    from tandem_amd import scenario as S

    mesh = T.Mesh.box(3, cells, scheme=1, bc=T.BC_DIRICHLET)
    nodes = S.warp(mesh.affine_nodes(degree))
    op = T.DGOperator(mesh, T.ELASTICITY, degree, nodes)
    X = op.volume_points()
    lam, mu = S.lame_fields(3, seed)
    op.set_material(lam(X), mu(X))
    return op, mesh, nodes, (lam, mu)


def time_applies(op, x, y, steps, warmup, stream):
    import torch

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


def cpu_baseline(cells=10, reps=10):
    """The CPU FP64 oracle (element-centric restatement of tandem's apply) on a Kuhn mesh of
    cells^3 cubes with the same element type, geometry and coefficients."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from tandem_amd import scenario as S

    threads = min(os.cpu_count() or 1, 16)
    om = O.Mesh.box(3, [cells] * 3, 1, [5] * 6)
    nodes = S.warp(om.affine_nodes(4))
    vc, _ = om.coords(4, nodes)
    lam, mu = S.lame_fields(3, SEED)
    op = O.Operator(0, om, 4, lam(vc), mu(vc), nodes=nodes)
    x = np.random.default_rng(SEED).uniform(-1, 1, op.n)
    op.apply(x, threads)
    t0 = time.perf_counter()
    for _ in range(reps):
        op.apply(x, threads)
    tm = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    op.apply(x, 1)
    t1 = time.perf_counter() - t0
    return {
        "value": op.n / tm * 1e-9,
        "unit": "GDoF/s",
        "cores": threads,
        "kind": "port",
        "sample": "CPU FP64 oracle apply, P4 tets on %d^3 Kuhn cubes (%d elements, %d DoF), mean of %d applies "
                  "on %d OpenMP threads; 1 thread: %.4f GDoF/s" % (cells, op.E, op.n, reps, threads, op.n / t1 * 1e-9),
    }


def traffic_from_profile():
    p = os.path.join(ROOT, "profiles", "elasticity_apply_p4_c1_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch") if d.get("kernel") == KERNEL else None


def solve_comparison(op, mesh, nodes, fields, rtol=1e-10):
    """FP64 CG vs FP32 CG + FP64 refinement on the manufactured cosine solution."""
    import torch
    from tandem_amd import scenario as S

    lam, mu = fields
    X = op.volume_points()
    f = S.elasticity_body_force(X, lam, mu)               # (E, Nq, 3)
    Xf = op.facet_points()
    g, _ = S.cosine_solution(Xf)                            # (F, nq, 3)
    b = op.rhs(np.ascontiguousarray(f.transpose(0, 2, 1)), np.ascontiguousarray(g.transpose(0, 2, 1)))
    bd = torch.from_numpy(b).cuda()
    out = {}
    xs = {}
    for mode in ("fp64", "fp32-ru"):
        x = torch.zeros_like(bd)
        r = op.solve(bd, x, rtol=rtol, mode=mode, max_iter=40000)
        log("solve", mode, r)
        out[mode] = {"seconds": r["seconds"], "iterations": r["iterations"], "fp64_replacements": r["outer_iterations"],
                     "rel_residual": r["rel_residual"], "converged": bool(r["converged"])}
        xs[mode] = x.cpu().numpy()
    out["speedup_fp32ru_vs_fp64"] = out["fp64"]["seconds"] / out["fp32-ru"]["seconds"]
    out["solution_rel_diff_fp32ru_vs_fp64"] = float(np.linalg.norm(xs["fp32-ru"] - xs["fp64"]) / np.linalg.norm(xs["fp64"]))
    # discretisation error of the FP64 solution against the manufactured solution at the volume points
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=24)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy kernel (y = -x over BYTES of float64) --steps times; "
                         "its FETCH_SIZE / WRITE_SIZE counters calibrate the HBM traffic of the apply kernel")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the operator is not sharded across GPUs); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py: no GPU available", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        n = args.copy_probe // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(args.steps):
            torch.neg(x, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": args.steps}))
        return
    t0 = time.perf_counter()
    op, mesh, nodes, fields = build_workload(args.cells)
    op.upload()
    log("setup %.1fs, %d elements, %d DoF" % (time.perf_counter() - t0, op.E, op.n))
    stream = torch.cuda.Stream()
    op.set_stream(stream)
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
        "kernel": KERNEL,
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
    }
    if not args.no_solve:
        res["solve"] = solve_comparison(op, mesh, nodes, fields)
    res["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
