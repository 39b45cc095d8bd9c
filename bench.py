#!/usr/bin/env python3
"""Benchmark of the CPR-preconditioned Krylov solve path on one MI355X.

Default workload: the 168x100x120 two-unknown storage-aquifer system (4,032,000 unknowns,
56,056,320 nonzeros, FP64 CSR), CPR (AMG pressure stage + FSAI) right-preconditioned GMRES(50)
to rtol 1e-8 (--cfg selects another generator configuration). A "step" is one solve from a zero
initial guess with the system and the preconditioner already resident in HBM. The
FP32-preconditioner solve is the headline value; the FP64-only solve runs beside it (both with
the preconditioner's products on cell blocks), and both again on the CSR storage. The dominant kernel's roofline (the solvers' FP64 product with A on its cell-block
shadow) is measured with hipEvents on the library stream.

Prints one JSON line.
"""
import argparse
import ctypes
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


# configuration-specific CPR keys: the anisotropic case takes the FSAI pattern from A^2 (power 2,
# device SpGEMM)
WORKLOAD_PREC = {5: {"prec_glo": {"power": "2"}}}

WORKLOAD_KEYS = {1: "two_phase_32x32x16", 11: "two_phase_32x32x32", 2: "black_oil_46x112x22",
                 3: "storage_aquifer_168x100x120", 5: "anisotropic_64x118x263"}


def cpu_solve_baseline(g, M, prec_tree, sol, gpu_iters, args, log):
    """The oracle's OpenMP CPR-GMRES on the host cores, on the same system and the same AMG
    hierarchy: the whole solve to the same tolerance (median of --cpu-reps runs), or, with
    --cpu-iters k > 0, k iterations scaled to the GPU's iteration count."""
    import deco_amd
    import oracle
    n = g["n"]
    Ao = oracle.Mat(n, n, g["r"], g["c"], g["v"])
    bo = g["b"].copy()
    t0 = time.perf_counter()
    Mo = oracle.CPR(deco_amd.cpr_params(deco_amd.PTree(prec_tree)), Ao, bo)
    # the oracle's own AMG setup (R, Galerkin products, drop rule, smoother diagonals, coarsest
    # inverse) on the device hierarchy's interpolation operators and C/F splittings
    levels, _ = M.amg().export()
    p = deco_amd.cpr_params(deco_amd.PTree(prec_tree))
    Mo.set_local(oracle.amg(Mo.pressure_matrix(), [oracle.Mat(l["n"], l["nc"], *l["P"]) for l in levels],
                            [l["cf"] for l in levels], p.amg_relax_type, p.amg_A_drop_tol, p.amg_relax_weight))
    setup = time.perf_counter() - t0
    threads = oracle.num_threads()
    k = args.cpu_iters
    desc = (f"oracle CPR-{sol['type']}({sol['restart']}) on the same decoupled system, AMG on the device's "
            f"interpolation operators")
    if k > 0:
        sp = deco_amd.sol_params(dict(sol, itmax=str(k)))
        t1 = time.perf_counter()
        _, info = oracle.solve(sp, Ao, Mo, bo)
        dt = time.perf_counter() - t1
        per_it = dt / max(info.iter, 1)
        log(f"CPU oracle: {info.iter} iterations in {dt:.2f}s ({threads} threads), setup+import {setup:.1f}s")
        return {"value": per_it * gpu_iters, "unit": "s", "cores": threads, "kind": "port",
                "sample": f"{desc}, {info.iter} iterations timed ({dt:.2f}s, OpenMP {threads} threads), scaled to "
                          f"the {gpu_iters} iterations of the GPU FP64 solve; per-iteration {per_it * 1e3:.1f} ms"}
    sp = deco_amd.sol_params(sol)
    times = []
    for _ in range(max(args.cpu_reps, 1)):
        t1 = time.perf_counter()
        _, info = oracle.solve(sp, Ao, Mo, bo)
        times.append(time.perf_counter() - t1)
    dt = float(np.median(times))
    log(f"CPU oracle: solve to rtol {sol['reltol']} in {info.iter} iterations, median {dt:.2f}s of {len(times)} "
        f"({threads} threads), setup+import {setup:.1f}s")
    # the same solve on 4 threads (the reference's run script uses 4 OpenMP threads)
    oracle.set_threads(4)
    t1 = time.perf_counter()
    _, info4 = oracle.solve(sp, Ao, Mo, bo)
    dt4 = time.perf_counter() - t1
    oracle.set_threads(threads)
    log(f"CPU oracle on 4 threads: {info4.iter} iterations in {dt4:.2f}s")
    return {"value": dt, "unit": "s", "cores": threads, "kind": "port", "value_4_threads": dt4,
            "host": cpu_model(),
            "sample": f"{desc}: the whole solve to rtol {sol['reltol']} ({info.iter} iterations; the GPU FP64 solve "
                      f"takes {gpu_iters}), median of {len(times)} runs, OpenMP {threads} threads; "
                      f"per-iteration {dt / max(info.iter, 1) * 1e3:.1f} ms; one run on 4 threads {dt4:.2f}s"}


def solve_only(deco_amd, g, A, key, log):
    """CPR setup with the bench's parameters, then in each precision one warm-up and one timed
    flexible CPR-GMRES(50) solve to rtol 1e-8, and five 512 MiB copies (the reference reading of the
    FETCH_SIZE / WRITE_SIZE counters): the whole-solve workload of the counter passes."""
    n = g["n"]
    b = deco_amd.DeviceArray.from_numpy(g["b"])
    M = deco_amd.CPR(deco_amd.PTree(dict(PREC_TREE, block_size=str(g["bs"]))), A, b)
    deco_amd.lib().deco_sync()
    log("CPR setup done")
    sol = {"type": "rgmres", "reltol": "1e-8", "restart": "50", "itmax": "500", "verbosity": "0", "flexible": "1"}
    out = {"workload": key, "n": n, "nnz": g["nnz"]}
    for prec, name in ((deco_amd.PREC_FP64, "fp64"), (deco_amd.PREC_FP32, "fp32_precond")):
        M.set_precision(prec)
        deco_amd.SOL(sol, A, M, b)
        S = deco_amd.SOL(sol, A, M, b)
        out[name] = {"iter": S.iter, "ms": S.info.time_ms, "n_prec": S.info.n_prec}
        log(f"{name} solve: {out[name]}")
    out["copy_bytes"] = copy_calibration(deco_amd)
    print(json.dumps(out), flush=True)
    return 0


def adjoint_small(deco_amd, args, log):
    """Forward CPR-GMRES(30) solve of b = A x* and the adjoint of L = w.x (w ~ N(0,1), and the
    manufactured w = A^T y* whose solution is y*): the transposed solve, lambda = W^T mu and dL/dA
    on A's pattern, to rtol 1e-8 on the 32x32x16 two-unknown system, both preconditioner
    precisions; medians of args.steps runs."""
    g = deco_amd.generate(1)
    n = g["n"]
    A = deco_amd.CSR(n, n, g["r"], g["c"], g["v"])
    b = deco_amd.DeviceArray.from_numpy(g["b"])
    M = deco_amd.CPR(deco_amd.PTree(dict(PREC_TREE, block_size=str(g["bs"]))), A, b)
    sol = {"type": "rgmres", "reltol": "1e-8", "restart": "30", "itmax": "500", "verbosity": "0"}
    w = deco_amd.DeviceArray.from_numpy(np.random.default_rng(1004).standard_normal(n))
    # a manufactured adjoint right-hand side w = A^T y* (A before decoupling): lambda = y*
    import scipy.sparse as sps
    ystar = np.random.default_rng(1006).uniform(-1.0, 1.0, n)
    wm = deco_amd.DeviceArray.from_numpy(sps.csr_matrix((g["v"], g["c"], g["r"]), shape=(n, n)).T @ ystar)
    out = {"system": f"two-phase 32x32x16, n={n}, nnz={g['nnz']}"}

    def med(fn):
        fn()
        deco_amd.lib().deco_sync()
        ts = []
        for _ in range(max(args.steps, 3)):
            t1 = time.perf_counter()
            r = fn()
            deco_amd.lib().deco_sync()
            ts.append(time.perf_counter() - t1)
        return float(np.median(ts)) * 1e3, r

    # outputs allocated once, outside the timed calls (x, mu, lambda, dL/dA)
    xf, xv, mu = (deco_amd.DeviceArray(n) for _ in range(3))
    lam, dA = deco_amd.DeviceArray(n), deco_amd.DeviceArray(g["nnz"])
    for prec, key in ((deco_amd.PREC_FP64, "fp64"), (deco_amd.PREC_FP32, "fp32_precond")):
        M.set_precision(prec)
        fw_ms, Sf = med(lambda: deco_amd.SOL(sol, A, M, b, xf))
        fwv_ms, Sv = med(lambda: deco_amd.SOL(sol, A, M, w, xv))

        def adjoint():
            St = deco_amd.SOL(sol, A, M, w, mu, transpose=True)
            M.transform_transpose(St.x, out=lam)
            deco_amd.grad_A_pattern(A, lam, Sf.x, out=dA)
            return St
        adj_ms, St = med(adjoint)

        def adjoint_man():
            St = deco_amd.SOL(sol, A, M, wm, mu, transpose=True)
            M.transform_transpose(St.x, out=lam)
            deco_amd.grad_A_pattern(A, lam, Sf.x, out=dA)
            return St
        adjm_ms, Sm = med(adjoint_man)
        err_m = float(np.linalg.norm(lam.numpy() - ystar) / np.linalg.norm(ystar))
        out[key] = {"forward_ms": fw_ms, "forward_iter": Sf.iter, "adjoint_ms": adj_ms, "adjoint_iter": St.iter,
                    "ratio_to_forward": adj_ms / fw_ms, "forward_same_rhs_ms": fwv_ms,
                    "forward_same_rhs_iter": Sv.iter, "ratio_to_forward_same_rhs": adj_ms / fwv_ms,
                    "adjoint_manufactured_ms": adjm_ms, "adjoint_manufactured_iter": Sm.iter,
                    "adjoint_manufactured_err_vs_ystar": err_m, "ratio_manufactured_to_forward": adjm_ms / fw_ms}
    M.set_precision(deco_amd.PREC_FP64)
    log(f"adjoint (32x32x16): {out}")
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return f"{line.split(':', 1)[1].strip()}, {os.cpu_count()} logical CPUs"
    except OSError:
        pass
    return f"{os.cpu_count()} logical CPUs"


def bsr_bytes(nb, mb, nnzb, R, C, sv, sx, sy, layout=0):
    # algorithmic bytes of the block product: block values + block columns (4 bytes, 2 for the
    # 16-bit offsets of layout 2, none for the uniform offsets of layout 3) + x + y, and the row
    # structure the layout reads: block row pointers (chunked, layout 0) or per-slice offsets and
    # widths of 64 block rows (sliced, layouts 1-3; their padding blocks are not counted)
    structure = 8 * ((nb + 63) // 64) if layout else 4 * (nb + 1)
    col = {2: 2, 3: 0}.get(layout, 4)
    return nnzb * (R * C * sv + col) + structure + sx * mb * C + sy * nb * R


def spmv_section(deco_amd, A, n, nnz, iters, bs):
    """Products with A timed with hipEvents on the library stream: FP64 and FP32 CSR, and for a
    system with cells the FP64 and FP32 cell-block forms (the FP64 one is the solvers' operator,
    the FP32 one the preconditioner's full residual)."""
    P64, P32 = deco_amd._lib.PREC_FP64, deco_amd._lib.PREC_FP32
    ms64 = A.spmv_time(P64, iters)
    ms32 = A.spmv_time(P32, iters)
    b64 = spmv_bytes(n, n, nnz, 8, 8, 8)
    b32 = spmv_bytes(n, n, nnz, 4, 4, 4)
    out = {"fp64_us": ms64 * 1e3, "fp64_GBps": b64 / (ms64 * 1e-3) / 1e9, "fp32_us": ms32 * 1e3,
           "fp32_GBps": b32 / (ms32 * 1e-3) / 1e9, "bytes_fp64": b64, "bytes_fp32": b32}
    if bs > 1:
        msb64, nnzb = A.spmv_time(P64, iters, block=(bs, bs))
        msb32, _ = A.spmv_time(P32, iters, block=(bs, bs))
        sliced, _, stored = A.block_info(bs, bs)
        bb64 = bsr_bytes(n // bs, n // bs, nnzb, bs, bs, 8, 8, 8, sliced)
        bb32 = bsr_bytes(n // bs, n // bs, nnzb, bs, bs, 4, 4, 4, sliced)
        out.update({"block_layout": ("chunked block CSR", "sliced (SELL, 64 block rows per slice)",
                                     "sliced (SELL, 64 block rows per slice, 16-bit block-column offsets)",
                                     "sliced (SELL, 64 block rows per slice, uniform block-column offsets: no "
                                     "stored columns)")[sliced],
                    "block_padding": stored / max(nnzb, 1) - 1.0,
                    "blocks": nnzb, "block_fp64_us": msb64 * 1e3, "block_fp64_GBps": bb64 / (msb64 * 1e-3) / 1e9,
                    "block_fp32_us": msb32 * 1e3, "block_fp32_GBps": bb32 / (msb32 * 1e-3) / 1e9,
                    "bytes_block_fp64": bb64, "bytes_block_fp32": bb32})
    return out


def copy_calibration(deco_amd, nbytes=512 << 20, reps=5):
    """Plain device-to-device copies of a known size (the reference reading for the FETCH_SIZE /
    WRITE_SIZE counters of a profiled run)."""
    a = deco_amd.DeviceArray(nbytes // 8)
    b = deco_amd.DeviceArray(nbytes // 8)
    L = deco_amd.lib()
    for _ in range(reps):
        deco_amd.check(L.deco_memcpy_d2d(b.ptr, a.ptr, nbytes))
    deco_amd.check(L.deco_sync())
    return nbytes


def copy_rate(deco_amd, nbytes=512 << 20, reps=20):
    """Measured copy bandwidth on this GPU (bytes read + written per second, GB/s): back-to-back
    512 MiB device-to-device copies between two device syncs."""
    a = deco_amd.DeviceArray(nbytes // 8)
    b = deco_amd.DeviceArray(nbytes // 8)
    L = deco_amd.lib()
    deco_amd.check(L.deco_memcpy_d2d(b.ptr, a.ptr, nbytes))
    deco_amd.check(L.deco_sync())
    t0 = time.perf_counter()
    for _ in range(reps):
        deco_amd.check(L.deco_memcpy_d2d(b.ptr, a.ptr, nbytes))
    deco_amd.check(L.deco_sync())
    return 2 * nbytes * reps / (time.perf_counter() - t0) / 1e9


def traffic_from_profiles(kind, workload_key):
    """HBM bytes per launch of an FP64 product with A measured by the FETCH_SIZE and WRITE_SIZE
    passes, calibrated against a plain copy (profiles/<kind>_<workload>_traffic.md), or None."""
    path = os.path.join(ROOT, "profiles", f"{kind}_{workload_key}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 20 timed solves (about 0.25 s each way): a host stall of a few ms in one solve's polling loop
    # moves the mean wall time by a fraction of a percent instead of several
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cfg", type=int, default=3)
    ap.add_argument("--spmv-iters", type=int, default=50)
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the SpMV timing and a 512 MiB copy (for counter passes)")
    ap.add_argument("--cpu-sample", type=int, default=10, help="CPU SpMV repetitions in the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-solve", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=0,
                    help="GMRES iterations in the CPU baseline sample (0: the whole solve)")
    ap.add_argument("--cpu-reps", type=int, default=3, help="repetitions of the whole CPU solve (median)")
    ap.add_argument("--relax", type=int, default=None,
                    help="amg_relax_type of the headline solves (default: the reference parameter set's -1)")
    ap.add_argument("--relax-weight", type=float, default=0.0, help="amg_relax_weight (relax 7)")
    ap.add_argument("--solve-only", action="store_true",
                    help="CPR setup, one warm-up and one timed solve per precision and a 512 MiB copy (for "
                         "counter passes over whole solves)")
    ap.add_argument("--no-smoother-compare", action="store_true",
                    help="skip the second CPR setup with weighted Jacobi (relax 7, weight 2/3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the solve path runs on one GPU (the multi-GPU row partition is not implemented); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2
    if args.cfg not in WORKLOADS:
        print(f"bench.py: unknown --cfg {args.cfg}; choose one of {sorted(WORKLOADS)}", file=sys.stderr)
        return 2
    desc, method, restart = WORKLOADS[args.cfg]
    key = WORKLOAD_KEYS[args.cfg]
    import deco_amd
    deco_amd.init(0)
    t0 = time.time()
    g = deco_amd.generate(args.cfg)
    n, nnz = g["n"], g["nnz"]
    log(f"generated {key}: n={n} nnz={nnz} in {time.time() - t0:.1f}s")
    t0 = time.perf_counter()
    A = deco_amd.CSR(n, n, g["r"], g["c"], g["v"])   # host CSR -> HBM (pageable host arrays)
    deco_amd.lib().deco_sync()
    upload_s = time.perf_counter() - t0
    log(f"CSR upload {upload_s * 1e3:.1f} ms ({(12 * nnz + 4 * (n + 1)) / upload_s / 1e9:.1f} GB/s)")

    # ---------------------------------------------------------------- SpMV roofline
    if args.solve_only:
        return solve_only(deco_amd, g, A, key, log)
    spmv = spmv_section(deco_amd, A, n, nnz, args.spmv_iters, g["bs"])
    log(f"spmv fp64 {spmv['fp64_us']:.1f} us {spmv['fp64_GBps']:.0f} GB/s; "
        f"fp32 {spmv['fp32_us']:.1f} us {spmv['fp32_GBps']:.0f} GB/s")
    if "blocks" in spmv:
        log(f"block spmv fp64 {spmv['block_fp64_us']:.1f} us {spmv['block_fp64_GBps']:.0f} GB/s; "
            f"fp32 {spmv['block_fp32_us']:.1f} us {spmv['block_fp32_GBps']:.0f} GB/s")
    if args.spmv_only:
        cb = copy_calibration(deco_amd)
        print(json.dumps({"workload": key, "n": n, "nnz": nnz, "spmv": spmv, "copy_bytes": cb}), flush=True)
        return 0

    result = {
        "metric": "CPR-GMRES time-to-1e-8 (FP32-precond vs FP64, 1 GPU); SpMV HBM GB/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "scaling": "weak", "vs_baseline": None, "data": "synthetic",
        "config": {"workload": f"{desc}, n={n}, nnz={nnz}, CPR(AMG+FSAI) {method}({restart}) rtol 1e-8",
                   "n": n, "nnz": nnz, "parallelism": "single-gpu"},
        "spmv": spmv,
        "upload_s": upload_s,
    }
    copy_GBps = copy_rate(deco_amd)
    rd = ctypes.c_double()
    deco_amd.check(deco_amd.lib().deco_read_bandwidth(4 << 30, 10, ctypes.byref(rd)))   # 4 GiB: HBM, not cache
    read_GBps = rd.value
    log(f"device copy {copy_GBps:.0f} GB/s (read + write); streaming read {read_GBps:.0f} GB/s")
    # the dominant kernel: the solvers' FP64 product with A (cell-block form when A has cells)
    if "blocks" in spmv:
        bs = g["bs"]
        result["roofline"] = {
            "bound": "hbm", "achieved": spmv["block_fp64_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": spmv["block_fp64_GBps"] / HBM_PEAK_GBS, "traffic": traffic_from_profiles("spmv_block_fp64", key),
            "kernel": ((f"k_sell_stream<double,{bs},{bs},double,EpiOuter,"
                        f"{'DiaCols' if 'uniform' in spmv['block_layout'] else 'short' if '16-bit' in spmv['block_layout'] else 'int'},"
                        f"{7 if bs == 2 else 4}>")
                       if spmv["block_layout"].startswith("sliced")
                       else f"k_bsr_stream<double,{bs},{bs},double,EpiOuter>")
                      + f" (FP64 {bs}x{bs} cell-block product with A, {spmv['block_layout']})",
            "bytes_per_launch": spmv["bytes_block_fp64"]}
        result["roofline"].update(copy_GBps=copy_GBps, read_GBps=read_GBps,
                                  frac_of_read=spmv["block_fp64_GBps"] / read_GBps)
    else:
        result["roofline"] = {
            "bound": "hbm", "achieved": spmv["fp64_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": spmv["fp64_GBps"] / HBM_PEAK_GBS, "traffic": traffic_from_profiles("spmv_fp64", key),
            "kernel": "k_spmv_stream<double,double,EpiOuter> (FP64 CSR SpMV with A; the column-sorted chunk form where its chunks would need 32-bit columns)",
            "bytes_per_launch": spmv["bytes_fp64"]}
        result["roofline"].update(copy_GBps=copy_GBps, read_GBps=read_GBps,
                                  frac_of_read=spmv["fp64_GBps"] / read_GBps)

    # ---------------------------------------------------------------- in-cell reorder
    # AD::block_reorder, the driver's step before CPR (cpr.cpp:41), on a device copy of A
    order = "21" if g["bs"] == 2 else "213"
    Ac = A.copy()
    bc = deco_amd.DeviceArray.from_numpy(g["b"])
    deco_amd.lib().deco_sync()
    t0 = time.perf_counter()
    Ac.block_reorder(g["bs"], order, order, bc)                  # returns after a device sync
    dt = time.perf_counter() - t0
    rb_bytes = 2 * (12 * nnz + 4 * (n + 1)) + 16 * n             # CSR read + written, b read + written
    result["block_reorder"] = {"ms": dt * 1e3, "order": order, "bytes": rb_bytes, "GBps": rb_bytes / dt / 1e9}
    log(f"block_reorder {order}: {dt * 1e3:.2f} ms")
    del Ac, bc

    # ---------------------------------------------------------------- CPR-Krylov solves
    tree = dict(PREC_TREE, block_size=str(g["bs"]))
    for k, sub in WORKLOAD_PREC.get(args.cfg, {}).items():
        tree[k] = dict(tree[k], **sub)
    if args.relax is not None:
        tree["prec_loc"] = dict(tree["prec_loc"], amg_relax_type=str(args.relax))
    if args.relax_weight > 0:
        tree["prec_loc"] = dict(tree["prec_loc"], amg_relax_weight=str(args.relax_weight))
    prec_tree = deco_amd.PTree(tree)
    sol = {"type": method, "reltol": "1e-8", "restart": restart, "itmax": "500", "verbosity": "0"}
    b_dev = deco_amd.DeviceArray.from_numpy(g["b"])
    deco_amd.lib().deco_sync()
    t0 = time.perf_counter()
    M = deco_amd.CPR(prec_tree, A, b_dev)            # decouples A and b in place (CPR.hpp:42-172)
    deco_amd.lib().deco_sync()
    setup_s = time.perf_counter() - t0
    log(f"CPR setup {setup_s:.2f}s")
    tight = deco_amd.SOL(dict(sol, type="rgmres", restart="50", reltol="1e-12", itmax="2000"), A, M, b_dev)
    xt = tight.x.numpy()
    log(f"FP64 reference solve (rtol 1e-12): iter {tight.iter} res {tight.res_rel:.2e}")
    import scipy.sparse as sps
    A0 = sps.csr_matrix((g["v"], g["c"], g["r"]), shape=(n, n))   # the system before decoupling
    err = lambda x, ref: float(np.linalg.norm(x - ref) / np.linalg.norm(ref))

    def run(Mp, Ap, bp, precision, flexible=0, xref=None, xtight=None):
        """warm-up, then args.steps timed solves; the last solve's x and, from one more solve with
        the library's byte count on, the solve's algorithmic bytes"""
        Mp.set_precision(precision)
        sp_ = dict(sol, flexible=str(flexible))
        xbuf = deco_amd.DeviceArray(n)   # the solution's storage, allocated outside the timed solves
        for _ in range(args.warmup):
            deco_amd.SOL(sp_, Ap, Mp, bp, xbuf)
        deco_amd.lib().deco_sync()
        times, last = [], None
        t1 = time.perf_counter()
        for _ in range(args.steps):
            last = deco_amd.SOL(sp_, Ap, Mp, bp, xbuf)
            times.append(last.info.time_ms)
        deco_amd.lib().deco_sync()
        wall = (time.perf_counter() - t1) / args.steps
        x = last.x.numpy()
        deco_amd.count_bytes(True)
        deco_amd.SOL(sp_, Ap, Mp, bp, xbuf)
        nbytes, nlaunch = deco_amd.counted_bytes()
        deco_amd.count_bytes(False)
        # one more solve with events around each product with A: the roofline kernel's in-solve time
        deco_amd.time_outer_products(True)
        deco_amd.SOL(sp_, Ap, Mp, bp, xbuf)
        otot, ocnt = deco_amd.outer_product_times()
        deco_amd.time_outer_products(False)
        ms = float(np.median(times))
        r = dict(ms=ms, wall_ms=wall * 1e3, iter=last.iter, res_rel=last.res_rel, flag=last.flag,
                 n_prec=last.info.n_prec, alg_bytes=nbytes, launches=nlaunch,
                 alg_GBps=nbytes / (ms * 1e-3) / 1e9, frac_of_peak=nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                 outer_products=ocnt, outer_product_us=(otot / ocnt * 1e3) if ocnt else None)
        if xref is not None:
            r["err_vs_fp64_same_rtol"] = err(x, xref)
        if xtight is not None:
            r["err_vs_fp64_rtol1e-12"] = err(x, xtight)
        r["err_vs_xstar"] = err(x, g["xstar"])
        r["res_rel_original"] = float(np.linalg.norm(g["b"] - A0 @ x) / np.linalg.norm(g["b"]))
        return last, r

    S64, r64 = run(M, A, b_dev, deco_amd.PREC_FP64, 0, None, xt)
    x64 = S64.x.numpy()
    r64["err_vs_fp64_same_rtol"] = 0.0
    _, r64f = run(M, A, b_dev, deco_amd.PREC_FP64, 1, x64, xt)
    S32, r32 = run(M, A, b_dev, deco_amd.PREC_FP32, 0, x64, xt)
    M.set_precision(deco_amd.PREC_FP64)
    log(f"FP64-only (reference GMRES form): {r64}")
    log(f"FP64-only (flexible): {r64f}")
    log(f"FP32-precond: {r32}")
    # Both precisions run the preconditioner's products on cell-block shadows. The same pair of
    # solves on the CSR arrays (the reference's storage and summation order) shows what the
    # precision gains there.
    M.set_layout(deco_amd.LAYOUT_CSR)
    Sa, ra = run(M, A, b_dev, deco_amd.PREC_FP64, 1)
    xa = Sa.x.numpy()
    Sb, rb = run(M, A, b_dev, deco_amd.PREC_FP32, 0)
    xb = Sb.x.numpy()
    csr = {"fp64_ms": ra["wall_ms"], "fp64_iter": ra["iter"], "fp64_n_prec": ra["n_prec"], "fp32_ms": rb["wall_ms"],
           "fp32_iter": rb["iter"], "fp32_n_prec": rb["n_prec"], "speedup_fp32_vs_fp64": ra["wall_ms"] / rb["wall_ms"],
           "fp64_err_vs_block_fp64": err(xa, x64), "fp32_err_vs_fp64_same_rtol": err(xb, xa),
           "fp32_err_vs_fp64_rtol1e-12": err(xb, xt)}
    log(f"CSR storage: {csr}")
    M.set_layout(deco_amd.LAYOUT_BLOCK)
    M.set_precision(deco_amd.PREC_FP64)
    relax = int(tree["prec_loc"]["amg_relax_type"])
    result.update({
        "value": r32["wall_ms"] / 1e3, "unit": "s", "higher_is_better": False, "dtype": "f64",
        "ms_per_step": r32["wall_ms"],
        "solve": {"fp32_precond": r32, "fp64_only": r64f, "fp64_only_reference_form": r64,
                  "speedup_fp32_vs_fp64": r64f["wall_ms"] / r32["wall_ms"],
                  "speedup_fp32_vs_fp64_reference_form": r64["wall_ms"] / r32["wall_ms"],
                  "note": ("fp64_only runs flexible right GMRES, as the FP32-preconditioned solve does, so both "
                           "make the same preconditioner applications; fp64_only_reference_form keeps "
                           "rgmres.tpp's final M^-1(V y) application"),
                  "amg_relax_type": relax, "csr_storage": csr, "setup_s": setup_s, "fp64_rtol1e-12_iter": tight.iter},
        "solve_roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                           "fp32_precond": {"alg_bytes": r32["alg_bytes"], "achieved": r32["alg_GBps"],
                                            "frac": r32["frac_of_peak"]},
                           "fp64_only": {"alg_bytes": r64f["alg_bytes"], "achieved": r64f["alg_GBps"],
                                         "frac": r64f["frac_of_peak"]},
                           "definition": ("library byte count of one whole solve (every product's stored values, "
                                          "indices and row data, one pass over each vector it reads or writes; "
                                          "Gram-Schmidt sweeps; vector kernels, copies and fills) / the solve's "
                                          "device time")},
    })

    # the roofline kernel inside the solves (events around each product with A, one solve each):
    # FP64 directions in the flexible FP64 solve, f32 directions (x read as 4 bytes) in the FP32 one
    rl = result["roofline"]
    if r64f["outer_product_us"]:
        rl["in_solve_us"] = r64f["outer_product_us"]
        rl["in_solve_GBps"] = rl["bytes_per_launch"] / (r64f["outer_product_us"] * 1e-6) / 1e9
        rl["in_solve_frac"] = rl["in_solve_GBps"] / HBM_PEAK_GBS
        rl["back_to_back_us"] = rl["bytes_per_launch"] / (rl["achieved"] * 1e9) * 1e6
    if r32["outer_product_us"]:
        b32 = rl["bytes_per_launch"] - 4.0 * n
        rl["in_solve_fp32_directions"] = {"us": r32["outer_product_us"], "bytes_per_launch": b32,
                                          "GBps": b32 / (r32["outer_product_us"] * 1e-6) / 1e9}

    # the same workload with the explicit weighted-Jacobi smoother (relax 7, weight 2/3) against
    # the reference parameter set's default (-1: two-stage Gauss-Seidel here)
    if not args.no_smoother_compare and relax != 7:
        A2 = deco_amd.CSR(n, n, g["r"], g["c"], g["v"])
        b2 = deco_amd.DeviceArray.from_numpy(g["b"])
        t2 = dict(tree, prec_loc=dict(tree["prec_loc"], amg_relax_type="7", amg_relax_weight=str(2.0 / 3.0)))
        t0 = time.perf_counter()
        M2 = deco_amd.CPR(deco_amd.PTree(t2), A2, b2)
        deco_amd.lib().deco_sync()
        setup2 = time.perf_counter() - t0
        _, j64 = run(M2, A2, b2, deco_amd.PREC_FP64, 1, x64, xt)
        _, j32 = run(M2, A2, b2, deco_amd.PREC_FP32, 0, x64, xt)
        result["solve"]["smoothers"] = {
            str(relax): {"fp64_ms": r64f["wall_ms"], "fp64_iter": r64f["iter"], "fp32_ms": r32["wall_ms"],
                         "fp32_iter": r32["iter"], "setup_s": setup_s},
            "7 (weight 2/3)": {"fp64_ms": j64["wall_ms"], "fp64_iter": j64["iter"], "fp32_ms": j32["wall_ms"],
                               "fp32_iter": j32["iter"], "setup_s": setup2,
                               "speedup_fp32_vs_fp64": j64["wall_ms"] / j32["wall_ms"]}}
        log(f"smoothers: {result['solve']['smoothers']}")
        del M2, A2, b2
    # a manufactured adjoint right-hand side, as b = A x* is for the forward solve: w = A^T y*
    # (the system before decoupling), so the adjoint solution is lambda = y*
    ystar = np.random.default_rng(1006).uniform(-1.0, 1.0, n)
    w_man = A0.T @ ystar
    del A0

    # ---------------------------------------------------------------- adjoint
    # transposed CPR solve (W A)^T mu = w, lambda = W^T mu, dL/dA on A's pattern; and, because
    # iteration counts depend on the right-hand side (w ~ N(0,1) is harder than b = A x*), the
    # forward solve with the same w as the like-for-like cost. Both preconditioner precisions.
    if not args.no_adjoint and method == "rgmres":
        w = deco_amd.DeviceArray.from_numpy(np.random.default_rng(1004).standard_normal(n))
        fwd = {deco_amd.PREC_FP64: r64, deco_amd.PREC_FP32: r32}
        result["adjoint"] = {}
        # outputs allocated once, outside the timed calls (mu, lambda, dL/dA on A's pattern, x)
        mu, lam, xw = (deco_amd.DeviceArray(n) for _ in range(3))
        dA = deco_amd.DeviceArray(g["nnz"])
        for prec, key in ((deco_amd.PREC_FP64, "fp64"), (deco_amd.PREC_FP32, "fp32_precond")):
            M.set_precision(prec)
            St = deco_amd.SOL(sol, A, M, w, mu, transpose=True)  # warm-up (builds A^T and the transposes)
            deco_amd.lib().deco_sync()
            ts, tf = [], []
            for _ in range(args.steps):
                t1 = time.perf_counter()
                St = deco_amd.SOL(sol, A, M, w, mu, transpose=True)
                M.transform_transpose(St.x, out=lam)
                deco_amd.grad_A_pattern(A, lam, S64.x, out=dA)
                deco_amd.lib().deco_sync()
                ts.append(time.perf_counter() - t1)
            for _ in range(args.steps):
                t1 = time.perf_counter()
                Sf = deco_amd.SOL(sol, A, M, w, xw)
                deco_amd.lib().deco_sync()
                tf.append(time.perf_counter() - t1)
            adj_ms, fw_ms = float(np.median(ts)) * 1e3, float(np.median(tf)) * 1e3
            result["adjoint"][key] = {
                "wall_ms": adj_ms, "iter": St.iter, "flag": St.flag, "res_rel": St.res_rel,
                "ratio_to_forward_b": adj_ms / fwd[prec]["wall_ms"],
                "forward_same_rhs_ms": fw_ms, "forward_same_rhs_iter": Sf.iter,
                "ratio_to_forward_same_rhs": adj_ms / fw_ms,
                "per_iteration_ratio": (adj_ms / max(St.iter, 1)) / (fw_ms / max(Sf.iter, 1))}
            log(f"adjoint {key}: {result['adjoint'][key]}")
        # the manufactured right-hand side w = A^T y*: the like-for-like counterpart of the forward
        # solve of b = A x*, and lambda = y* checks the adjoint solution
        wm = deco_amd.DeviceArray.from_numpy(w_man)
        result["adjoint_manufactured"] = {"rhs": "w = A^T y*, y* ~ U[-1,1] (seed 1006), A before decoupling"}
        for prec, key in ((deco_amd.PREC_FP64, "fp64"), (deco_amd.PREC_FP32, "fp32_precond")):
            M.set_precision(prec)
            St = deco_amd.SOL(sol, A, M, wm, mu, transpose=True)
            deco_amd.lib().deco_sync()
            ts = []
            for _ in range(args.steps):
                t1 = time.perf_counter()
                St = deco_amd.SOL(sol, A, M, wm, mu, transpose=True)
                M.transform_transpose(St.x, out=lam)
                deco_amd.grad_A_pattern(A, lam, S64.x, out=dA)
                deco_amd.lib().deco_sync()
                ts.append(time.perf_counter() - t1)
            adj_ms = float(np.median(ts)) * 1e3
            lam_h = lam.numpy()
            result["adjoint_manufactured"][key] = {
                "wall_ms": adj_ms, "iter": St.iter, "flag": St.flag, "res_rel": St.res_rel,
                "err_vs_ystar": float(np.linalg.norm(lam_h - ystar) / np.linalg.norm(ystar)),
                "ratio_to_forward_b": adj_ms / fwd[prec]["wall_ms"], "forward_b_iter": fwd[prec]["iter"]}
            log(f"adjoint, manufactured rhs, {key}: {result['adjoint_manufactured'][key]}")
        M.set_precision(deco_amd.PREC_FP64)

    # adjoint-to-forward time on the 32x32x16 two-unknown system (the adjoint case's timing size)
    if not args.no_adjoint:
        result["adjoint_small"] = adjoint_small(deco_amd, args, log)

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
        result["spmv"]["cpu_GBps"] = spmv["bytes_fp64"] / dt / 1e9
        result["spmv"]["cpu_sample"] = (f"{args.cpu_sample} FP64 CSR SpMVs with A (oracle restatement, "
                                        f"OpenMP {threads} threads); {dt * 1e3:.1f} ms each")
        result["cpu_baseline"] = {"value": spmv["bytes_fp64"] / dt / 1e9, "unit": "GB/s", "cores": threads,
                                  "kind": "port", "sample": result["spmv"]["cpu_sample"]}
        if not args.no_cpu_solve:
            result["cpu_baseline"] = cpu_solve_baseline(g, M, tree, sol, r64["iter"], args, log)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
