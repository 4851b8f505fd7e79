"""Benchmark of the MLS transfer path on one MI355X.

A step is one pass of the hot path over one batch: build the operator for
1,000,000 sources uniform in [0,1]^3 and 1,000,000 targets uniform in
[0.05,0.95]^3 (kNN + fused setup: weights, Vandermonde, moments, Jacobi
pseudo-inverse, coefficients), then apply it to 4 scalar fields (smooth
trig polynomials + 1e-6 Gaussian noise). Inputs are resident in HBM before
the timed region. The headline value is target points per second,
N_t / (t_knn + t_setup + t_apply(4 fields)), for the plain FP64 form with
the reference's quadratic basis and k = basis size (10), Wendland C2.

Prints one JSON line (rank 0). See DESIGN.md for the roofline definitions.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def f_setup(m, k):
    """Algorithmic FP64 flops per target of the setup (fixed before
    measuring): geometry 25k, symmetric moments k m(m+1), 6 nominal Jacobi
    sweeps of m(m-1)/2 rotations at (18m + 30), pinv row 0 2m^2,
    coefficients k(2m+1)."""
    return 25 * k + k * m * (m + 1) + 6 * (m * (m - 1) // 2) * (18 * m + 30) + 2 * m * m + k * (2 * m + 1)


def b_apply(nnz, n_tgt, n_src, fields, fp32=False):
    """Algorithmic HBM bytes of one apply of `fields` fields."""
    if fp32:
        return nnz * 8 + 4 * (n_tgt + 1) + fields * 4 * (n_src + n_tgt)
    return nnz * 12 + 4 * (n_tgt + 1) + fields * 8 * (n_src + n_tgt)


def b_knn(n_src, n_tgt, k):
    return 24 * n_src + n_tgt * (24 + 4 * k)


def make_inputs(n_src, n_tgt, fields, seed=1):
    import numpy as np

    rng = np.random.default_rng(seed)
    src = rng.random((n_src, 3))
    tgt = 0.05 + 0.9 * rng.random((n_tgt, 3))
    x, y, z = src[:, 0], src[:, 1], src[:, 2]
    tp = 2 * np.pi
    funcs = [
        np.sin(tp * x) * np.cos(tp * y) + z * z,
        np.cos(tp * x + 0.3) * np.sin(2 * tp * z) + 0.5 * y,
        x * y * z + np.sin(3 * tp * y),
        np.cos(tp * (x + y + z)) + x * x - y,
    ]
    vals = np.stack([funcs[i % 4] for i in range(fields)])
    vals = vals + 1e-6 * rng.standard_normal(vals.shape)
    return src, tgt, vals


def cpu_baseline(src, tgt, vals, m, k, rbf, sample):
    """The oracle (C++ restatement of the reference, OpenMP) on the first
    `sample` targets against all sources, same stages: kNN, setup, apply."""
    import numpy as np

    import oracle

    threads = min(os.cpu_count() or 1, int(os.environ.get("DTKX_CPU_THREADS", "16")))
    oracle.set_num_threads(threads)
    t0 = time.perf_counter()
    idx, coeffs, _ = oracle.mls_create(src, tgt[:sample], m, k=k, rbf_name=rbf)
    oracle.apply(idx, coeffs, vals)
    dt = time.perf_counter() - t0
    return {
        "value": sample / dt,
        "unit": "target pts/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {sample} of {len(tgt)} targets, all {len(src)} sources, "
                  f"k-d tree kNN + setup + apply of {vals.shape[0]} fields ({dt:.1f} s)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--fields", type=int, default=4)
    ap.add_argument("--rbf", default="wendland2")
    ap.add_argument("--solver", default="classical")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--group", type=int, default=0,
                    help="lanes per system for the classical solver (0 = default)")
    ap.add_argument("--cpu-sample", type=int, default=20_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is designed, not built)"}))
        sys.exit(2)

    import numpy as np
    import torch

    import datatransferkit_amd as dtk

    dev = torch.device("cuda:0")
    m = 10
    n = args.n
    src, tgt, vals = make_inputs(n, n, args.fields)
    d_src = torch.from_numpy(src).to(dev)
    d_tgt = torch.from_numpy(tgt).to(dev)
    # the FP32-bulk form takes float values (the halved-traffic apply)
    vdt = torch.float32 if args.precision == "fp32_bulk" else torch.float64
    d_vals = torch.from_numpy(vals).to(dev).to(vdt)
    d_out = torch.zeros((args.fields, n), dtype=vdt, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    kw = dict(basis="quadratic", rbf=args.rbf, k=args.k, solver=args.solver,
              precision=args.precision, stream=stream, group_size=args.group)

    def step():
        op = dtk.MovingLeastSquaresOperator(d_src, d_tgt, **kw)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        op.apply(d_vals, d_out)
        b.record()
        b.synchronize()
        st = op.stats()
        st["ms_apply"] = a.elapsed_time(b)
        del op
        return st

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(step())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = 1000 * dt / args.steps
    value = n * args.steps / dt

    ms_knn = float(np.mean([s["ms_knn"] for s in stats]))
    ms_setup = float(np.mean([s["ms_setup"] for s in stats]))
    ms_apply = float(np.mean([s["ms_apply"] for s in stats]))
    kk = stats[0]["k"]
    nnz = stats[0]["nnz"]
    setup_tflops = f_setup(m, kk) * n / (ms_setup * 1e-3) / 1e12
    apply_gbs = b_apply(nnz, n, n, args.fields, args.precision == "fp32_bulk") / (ms_apply * 1e-3) / 1e9
    knn_gbs = b_knn(n, n, kk) / (ms_knn * 1e-3) / 1e9

    roofline = {
        "bound": "mfma",
        "kernel": "setup_classical_kernel (fused MLS setup, FP64 VALU)",
        "achieved": round(setup_tflops, 4),
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(setup_tflops / FP64_PEAK_TFLOPS, 5),
        "traffic": None,
        "algorithmic_per_unit": f"F_setup(m={m},k={kk}) = {f_setup(m, kk)} flop per target",
    }
    out = {
        "metric": "MLS setup+apply target pts/s (1M pts, quad basis): plain FP64 vs batched/FP32",
        "value": value,
        "unit": "target pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"MLS transfer {n} sources U[0,1]^3 -> {n} targets U[0.05,0.95]^3, "
                        f"quadratic basis, k={kk}, {args.rbf}, {args.fields} fields, "
                        f"{args.precision}, {args.solver} Jacobi",
            "n_src": n, "n_tgt": n, "basis": "quadratic", "k": kk, "rbf": args.rbf,
            "fields": args.fields, "precision": args.precision, "solver": args.solver,
        },
        "stages": {
            "ms_knn": ms_knn, "ms_setup": ms_setup, "ms_apply": ms_apply,
            "knn_GBps_algorithmic": knn_gbs,
            "setup_TFLOPs_algorithmic": setup_tflops,
            "apply_GBps_algorithmic": apply_gbs,
            "apply_frac_hbm": apply_gbs / HBM_PEAK_GBS,
            "rotations_per_target": stats[0]["total_rotations"] / n,
            "n_underdetermined": stats[0]["n_underdetermined"],
            "n_unconverged": stats[0]["n_unconverged"],
            "tree_height": stats[0]["tree_height"],
            "group_size": stats[0]["group_size_used"],
        },
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(src, tgt, vals, m, args.k, args.rbf, args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
