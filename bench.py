#!/usr/bin/env python3
"""Benchmark of the CPR-preconditioned Krylov solve path on one MI355X.

Workload: the 168x100x120 two-unknown storage-aquifer system (4,032,000 unknowns, 56,056,320
nonzeros, FP64 CSR), CPR (AMG pressure stage + FSAI) right-preconditioned GMRES(50) to
rtol 1e-8. A "step" is one solve from a zero initial guess with the system and the
preconditioner already resident in HBM. The FP32-preconditioner solve is the headline value;
the FP64-only solve runs beside it. The dominant kernel's roofline (the FP64 CSR SpMV with A)
is measured with hipEvents on the library stream.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n_rows, n_cols, nnz, sv, sx, sy):
    # algorithmic bytes: values + column indices + row pointers + x + y
    return nnz * (sv + 4) + 4 * (n_rows + 1) + sx * n_cols + sy * n_rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cfg", type=int, default=3)
    ap.add_argument("--spmv-iters", type=int, default=50)
    ap.add_argument("--cpu-sample", type=int, default=10, help="CPU SpMV repetitions in the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every GPU box here has one GPU); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    import deco_amd
    deco_amd.init(0)
    t0 = time.time()
    g = deco_amd.generate(args.cfg)
    n, nnz = g["n"], g["nnz"]
    log(f"generated cfg {args.cfg}: n={n} nnz={nnz} in {time.time() - t0:.1f}s")
    A = deco_amd.CSR(n, n, g["r"], g["c"], g["v"])

    # ---------------------------------------------------------------- SpMV roofline
    ms64 = A.spmv_time(deco_amd._lib.PREC_FP64, args.spmv_iters)
    ms32 = A.spmv_time(deco_amd._lib.PREC_FP32, args.spmv_iters)
    b64 = spmv_bytes(n, n, nnz, 8, 8, 8)
    b32 = spmv_bytes(n, n, nnz, 4, 4, 4)
    gbs64 = b64 / (ms64 * 1e-3) / 1e9
    gbs32 = b32 / (ms32 * 1e-3) / 1e9
    log(f"spmv fp64 {ms64 * 1e3:.1f} us {gbs64:.0f} GB/s; fp32 {ms32 * 1e3:.1f} us {gbs32:.0f} GB/s")

    result = {
        "metric": "CPR-GMRES time-to-1e-8 (FP32-precond vs FP64, 1 GPU); SpMV HBM GB/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "scaling": "weak", "vs_baseline": None, "data": "synthetic",
        "config": {"workload": f"cfg{args.cfg}: 168x100x120 grid, 2 unknowns/cell, n={n}, nnz={nnz}, "
                               "CPR(AMG+FSAI) GMRES(50) rtol 1e-8",
                   "n": n, "nnz": nnz, "parallelism": "single-gpu"},
        "spmv": {"fp64_us": ms64 * 1e3, "fp64_GBps": gbs64, "fp32_us": ms32 * 1e3, "fp32_GBps": gbs32,
                 "bytes_fp64": b64, "bytes_fp32": b32},
        "roofline": {"bound": "hbm", "achieved": gbs64, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs64 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_spmv<16,EPI_SET,double,double,double> (FP64 CSR SpMV with A)",
                     "bytes_per_launch": b64},
    }

    solve_done = False
    try:
        from deco_amd import solver as _solver  # noqa: F401
        have_solver = True
    except ImportError:
        have_solver = False
    if have_solver:
        solve_done = _solver.bench_solve(g, A, args, result, log)
    if not solve_done:
        result.update({"value": gbs64, "unit": "GB/s", "higher_is_better": True, "dtype": "f64",
                       "ms_per_step": ms64})

    # ---------------------------------------------------------------- CPU baseline
    if not args.no_cpu:
        import oracle
        M = oracle.Mat(n, n, g["r"], g["c"], g["v"])
        x = np.random.default_rng(0).uniform(-1, 1, n)
        threads = oracle.num_threads()
        oracle.spmv(M, x)
        t1 = time.perf_counter()
        for _ in range(args.cpu_sample):
            oracle.spmv(M, x)
        dt = (time.perf_counter() - t1) / args.cpu_sample
        result["cpu_baseline"] = {
            "value": b64 / dt / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{args.cpu_sample} FP64 CSR SpMVs with A of the same system (oracle/oracle.cpp "
                      f"alg::spmv restatement, OpenMP {threads} threads); {dt * 1e3:.1f} ms each",
        }
        if solve_done and "cpu_solve" in result:
            result["cpu_baseline"] = result.pop("cpu_solve")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
