#!/usr/bin/env python3
"""Benchmark of the CPR-preconditioned Krylov solve path on one MI355X.

Default workload: the 168x100x120 two-unknown storage-aquifer system (4,032,000 unknowns,
56,056,320 nonzeros, FP64 CSR), CPR (AMG pressure stage + FSAI) right-preconditioned GMRES(50)
to rtol 1e-8 (--cfg selects another generator configuration). A "step" is one solve from a zero initial guess with the system and the
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


# CPR parameters of the bench: the reference's own set for its two-unknown system
# (param_sleipner.json "prec": FSAI power 1 without filtering, HMIS, default relaxation,
# truncation 1e-3, drop 1e-4, strength 0.5, no DRS, tol_zero 1e-30).
PREC_TREE = {
    "block_size": "2", "pressure_id": "0", "seq": "3", "weights": "quasiimpes_GE", "use_drs": "0",
    "drs": {"e_dd": "0.9999", "e_ps": "0.0"}, "CPD": "1", "tol_zero": "1e-30",
    "prec_glo": {"type": "FSAI", "tol_rel": "0", "tol_abs": "0", "power": "1", "block_size": "1"},
    "prec_loc": {"tol_rel": "0", "amg_coarse_type": "10", "amg_relax_type": "-1", "amg_agg_levs": "0",
                 "amg_agg_path": "4", "amg_A_drop_tol": "1e-4", "amg_prol_trunc_fac": "1e-3",
                 "amg_SoC_alpha": "0.5"},
}


# generator configurations a bench run can select: description and Krylov method
WORKLOADS = {
    1: ("two-phase 32x32x16, 2 unknowns/cell", "rgmres", "30"),
    11: ("two-phase 32x32x32, 2 unknowns/cell", "rgmres", "30"),
    2: ("black-oil 46x112x22 (60% active, 2% fault links), 3 unknowns/cell", "bicgstab", "50"),
    3: ("storage aquifer 168x100x120 (12 layers, sigma 2), 2 unknowns/cell", "rgmres", "50"),
    5: ("anisotropic 64x118x263 (sigma 3, kz/kx 1e-3), 2 unknowns/cell", "rgmres", "50"),
}


def cpu_solve_baseline(g, M, prec_tree, sol, gpu_iters, args, log):
    """The oracle's OpenMP CPR-GMRES on the host cores, on the same system and the same AMG
    hierarchy, timed over a bounded number of iterations and scaled to the GPU's iteration
    count."""
    import deco_amd
    import oracle
    n = g["n"]
    Ao = oracle.Mat(n, n, g["r"], g["c"], g["v"])
    bo = g["b"].copy()
    t0 = time.perf_counter()
    Mo = oracle.CPR(deco_amd.cpr_params(deco_amd.PTree(prec_tree)), Ao, bo)
    levels, cinv = M.amg().export()
    lv = [dict(A=oracle.Mat(l["n"], l["n"], *l["A"]), P=oracle.Mat(l["n"], l["nc"], *l["P"]),
               R=oracle.Mat(l["nc"], l["n"], *l["R"]), d=l["d"], cf=l["cf"]) for l in levels]
    Mo.set_local(oracle.amg(lv, 0, cinv))
    setup = time.perf_counter() - t0
    k = args.cpu_iters
    sp = deco_amd.sol_params(dict(sol, itmax=str(k)))
    t1 = time.perf_counter()
    _, info = oracle.solve(sp, Ao, Mo, bo)
    dt = time.perf_counter() - t1
    per_it = dt / max(info.iter, 1)
    threads = oracle.num_threads()
    log(f"CPU oracle: {info.iter} iterations in {dt:.2f}s ({threads} threads), setup+import {setup:.1f}s")
    return {"value": per_it * gpu_iters, "unit": "s", "cores": threads, "kind": "port",
            "sample": f"oracle CPR-rgmres(50) on the same decoupled system and AMG hierarchy, {info.iter} "
                      f"iterations timed ({dt:.2f}s, OpenMP {threads} threads), scaled to the {gpu_iters} "
                      f"iterations of the GPU FP64 solve; per-iteration {per_it * 1e3:.1f} ms"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cfg", type=int, default=3)
    ap.add_argument("--spmv-iters", type=int, default=50)
    ap.add_argument("--cpu-sample", type=int, default=10, help="CPU SpMV repetitions in the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-solve", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=5, help="GMRES iterations in the CPU baseline sample")
    ap.add_argument("--export-levels", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every GPU box here has one GPU); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    if args.cfg not in WORKLOADS:
        print(f"bench.py: unknown --cfg {args.cfg}; choose one of {sorted(WORKLOADS)}", file=sys.stderr)
        return 2
    desc, method, restart = WORKLOADS[args.cfg]
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
        "config": {"workload": f"{desc}, n={n}, nnz={nnz}, CPR(AMG+FSAI) {method}({restart}) rtol 1e-8",
                   "n": n, "nnz": nnz, "parallelism": "single-gpu"},
        "spmv": {"fp64_us": ms64 * 1e3, "fp64_GBps": gbs64, "fp32_us": ms32 * 1e3, "fp32_GBps": gbs32,
                 "bytes_fp64": b64, "bytes_fp32": b32},
        "roofline": {"bound": "hbm", "achieved": gbs64, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs64 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_spmv_stream<double,double,EpiStd<EPI_SET>> (FP64 CSR SpMV with A)",
                     "bytes_per_launch": b64},
    }

    # ---------------------------------------------------------------- CPR-GMRES solves
    tree = dict(PREC_TREE, block_size=str(g["bs"]))
    prec_tree = deco_amd.PTree(tree)
    sol = {"type": method, "reltol": "1e-8", "restart": restart, "itmax": "500", "verbosity": "0"}
    b_dev = deco_amd.DeviceArray.from_numpy(g["b"])
    deco_amd.lib().deco_sync()
    t0 = time.perf_counter()
    M = deco_amd.CPR(prec_tree, A, b_dev)            # decouples A and b in place (CPR.hpp:42-172)
    deco_amd.lib().deco_sync()
    setup_s = time.perf_counter() - t0
    levels, _ = M.amg().export() if args.export_levels else ([], None)
    log(f"CPR setup {setup_s:.2f}s")
    x_ref = deco_amd.SOL(dict(sol, type="rgmres", restart="50", reltol="1e-12", itmax="2000"), A, M, b_dev)
    xr = x_ref.x.numpy()
    log(f"FP64 reference solve: iter {x_ref.iter} res {x_ref.res_rel:.2e}")

    def run(precision):
        M.set_precision(precision)
        for _ in range(args.warmup):
            deco_amd.SOL(sol, A, M, b_dev)
        deco_amd.lib().deco_sync()
        times, last = [], None
        t1 = time.perf_counter()
        for _ in range(args.steps):
            last = deco_amd.SOL(sol, A, M, b_dev)
            times.append(last.info.time_ms)
        deco_amd.lib().deco_sync()
        wall = (time.perf_counter() - t1) / args.steps
        x = last.x.numpy()
        return dict(ms=float(np.median(times)), wall_ms=wall * 1e3, iter=last.iter, res_rel=last.res_rel,
                    flag=last.flag, err_vs_fp64_oracle=float(np.linalg.norm(x - xr) / np.linalg.norm(xr)),
                    err_vs_xstar=float(np.linalg.norm(x - g["xstar"]) / np.linalg.norm(g["xstar"])))

    r64 = run(deco_amd.PREC_FP64)
    log(f"FP64-only: {r64}")
    r32 = run(deco_amd.PREC_FP32)
    log(f"FP32-precond: {r32}")
    M.set_precision(deco_amd.PREC_FP64)
    result.update({
        "value": r32["wall_ms"] / 1e3, "unit": "s", "higher_is_better": False, "dtype": "f64",
        "ms_per_step": r32["wall_ms"],
        "solve": {"fp32_precond": r32, "fp64_only": r64, "speedup_fp32_vs_fp64": r64["wall_ms"] / r32["wall_ms"],
                  "setup_s": setup_s, "reference_solve_iter": x_ref.iter},
    })

    # ---------------------------------------------------------------- CPU baseline
    if not args.no_cpu:
        import oracle
        Ao = oracle.Mat(n, n, g["r"], g["c"], g["v"])
        x = np.random.default_rng(0).uniform(-1, 1, n)
        threads = oracle.num_threads()
        oracle.spmv(Ao, x)
        t1 = time.perf_counter()
        for _ in range(args.cpu_sample):
            oracle.spmv(Ao, x)
        dt = (time.perf_counter() - t1) / args.cpu_sample
        del Ao
        result["spmv"]["cpu_GBps"] = b64 / dt / 1e9
        result["spmv"]["cpu_sample"] = (f"{args.cpu_sample} FP64 CSR SpMVs with A (oracle restatement, "
                                        f"OpenMP {threads} threads); {dt * 1e3:.1f} ms each")
        result["cpu_baseline"] = {"value": b64 / dt / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
                                  "sample": result["spmv"]["cpu_sample"]}
        if not args.no_cpu_solve:
            result["cpu_baseline"] = cpu_solve_baseline(g, M, tree, sol, r64["iter"], args, log)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
