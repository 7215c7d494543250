#!/usr/bin/env python3
"""Single-GPU SpMV benchmark: the all-formats single-matrix workload.

Workload (one matrix per section, FP64 unless stated, int32 indices):
  * lognormal: M = N = 4,194,304, per-row nnz lognormal (mean 16, sigma 1,
    capped at 4096), 70 % of columns uniform over [0, N) and 30 % within +-64 of
    the diagonal, values U[-1,1), x = U[-1,1) (seed 4); FP64 (headline) and FP32;
  * stencil27: 27-point stencil on a 160^3 grid (diag 26, off-diagonals -1) for
    the DIA and BSR paths.
One step = one y = A*x with the matrix and vectors resident in HBM.

Prints ONE JSON line. `value` is the effective bandwidth (algorithmic bytes:
values + indices + x read once + y written) of the fastest kernel (least time)
on the FP64 lognormal matrix, divided by the wall time per step. `serving`
(256 resident matrices, k uniform in 1..32, saturation and Poisson arrivals)
and `batched` (1,024 matrices, k in {1, 4, 8, 16, 32}, saturation) come from
sspmv_amd/bin/serve_bench: requests/s and p50 / p99 latency of the packed
persistent kernel against per-call CSR (the plain kernel and the server's own
row engine, synchronous and asynchronous). `roofline` compares the
same bytes with that kernel's device time (hipEvents on its stream) against
MI355X's 8 TB/s. `cpu_baseline` times the reference's load-balanced OpenMP
CSR SpMV (kernel_flag=2), built from the reference's sources into oracle/_ref
(or, where that build is absent, the oracle's restatement of it) on this
host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
LOGNORMAL = [4194304, 4194304, 16.0, 1.0, 4096, 0.3, 64]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernels(lib, section):
    """(label, fmt, variant, params) measured per section."""
    sell = lambda C, s: lib.make_params(chunk=C, sigma=s, alignment=1)
    if section == "stencil27":
        return [("CSR-vector", lib.CSR, lib.VARIANT_VECTOR, None), ("CSR-merge", lib.CSR, lib.VARIANT_MERGE, None),
                ("SELL-64-sigma4096", lib.SELL_C_SIGMA, 0,
                                                                    sell(64, 4096)),
                ("DIA", lib.DIA, lib.VARIANT_VECTOR, lib.make_params(alignment=1, max_diags=10240)),
                ("DIA-lds", lib.DIA, lib.VARIANT_MERGE, lib.make_params(alignment=1, max_diags=10240)),
                ("BSR-4x4", lib.BSR, 0, lib.make_params(block_rows=4, block_cols=4))]
    ks = [("CSR-vector", lib.CSR, lib.VARIANT_VECTOR, None), ("CSR-merge", lib.CSR, lib.VARIANT_MERGE, None),
          ("CSR-panel8", lib.CSR, lib.VARIANT_PANEL, None), ("CSR-panel16", lib.CSR, lib.VARIANT_PANEL16, None),
          ("CSR-panel4", lib.CSR, lib.VARIANT_PANEL4, None),
          ("SELL-64-sigma4096", lib.SELL_C_SIGMA, 0, sell(64, 4096)), ("SELL-64-R", lib.SELL_C_R, 0, sell(64, 0)),
          ("COO", lib.COO, 0, None), ("CSR5-64x16", lib.CSR5, 0, lib.make_params(omega=64, sigma=16))]
    return ks if section == "lognormal" else ks[:6]


def traffic_from_profiles(label):
    """HBM bytes per launch from rocprofv3 --pmc (profiles/traffic.json)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get(label)
    except Exception:
        return None


def measure(lib, M, dx, dy, section, wanted, warmup):
    results = []
    for label, fmt, variant, params in kernels(lib, section):
        if wanted and label not in wanted:
            continue
        try:
            if fmt != lib.CSR:
                M.convert(fmt, params)
            conv_ms = M.convert_ms(fmt) if fmt != lib.CSR else 0.0
        except lib.LespmvError as e:
            log(f"[{section}] {label}: skipped ({e})")
            continue
        eff = M.effective_bytes(fmt, 1, False)
        try:  # the reference's own formula (include/spmv_benchmark.h:26-160), for comparison only
            ref_bytes = M.reference_bytes(fmt)
        except lib.LespmvError:
            ref_bytes = None
        M.time_spmv(fmt, 1.0, dx, 0.0, dy, max(warmup, 1), variant)
        ms = M.time_spmv(fmt, 1.0, dx, 0.0, dy, 50, variant)
        results.append(dict(label=label, fmt=fmt, variant=variant, params=params, ms=ms, gbs=eff / ms / 1e6,
                            effective_bytes=eff, convert_ms=conv_ms, format_bytes=M.format_bytes(fmt),
                            ref_gbs=ref_bytes / ms / 1e6 if ref_bytes else None))
        log(f"[{section}] {label}: {ms * 1e3:.1f} us  {eff / ms / 1e6:.0f} GB/s  convert {conv_ms:.1f} ms")
        if fmt != lib.CSR:
            M.release(fmt)
    return results


def summary(results):
    return {r["label"]: {"us": round(r["ms"] * 1e3, 2), "GBs": round(r["gbs"], 1), "frac": round(r["gbs"] * 1e9 /
                                                                                                  HBM_PEAK, 4),
                         "convert_ms": round(r["convert_ms"], 2), "format_bytes": r["format_bytes"],
                         "GBs_reference_formula": round(r["ref_gbs"], 1) if r["ref_gbs"] else None}
            for r in results}


def reference_csr_bench():
    """The reference's own LeSpMV_csr (kernel_flag 2) built by oracle/build_ref.sh
    into oracle/_ref/, at the instruction level this host runs (x86-64-v4 with
    AVX-512, else x86-64-v3); None when it was not built."""
    import ctypes as C
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
    except OSError:
        flags = ""
    v4 = all(t in flags for t in ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"))
    level = "x86-64-v4" if v4 else "x86-64-v3"
    path = os.path.join(ROOT, "oracle", "_ref", f"libref_csr_bench_{level}.so")
    if not os.path.exists(path):
        return None, level
    lib = C.CDLL(path)
    p = np.ctypeslib.ndpointer
    lib.refbench_csr_lb.argtypes = [C.c_int, C.c_int, C.c_int, p(np.int32), p(np.int32), p(np.float64),
                                    p(np.float64), p(np.float64), C.c_int, C.c_int]
    lib.refbench_csr_lb.restype = C.c_double
    return lib, level


def cpu_baseline(A, x, threads, max_seconds):
    """ms per CPU SpMV, iterations, y, kind and a description of the kernel:
    the reference's own code where oracle/_ref holds it, else the oracle's
    restatement of the same kernel."""
    ref, level = reference_csr_bench()
    rp = np.ascontiguousarray(A.rp, dtype=np.int32)
    ci = np.ascontiguousarray(A.ci, dtype=np.int32)
    v = np.ascontiguousarray(A.v, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(A.m)
    if ref is not None:
        def time(iters):
            return ref.refbench_csr_lb(A.m, A.n, A.nnz, rp, ci, v, x, y, threads, iters)
        kind, what = "reference", f"the reference's LeSpMV_csr kernel_flag=2 built from its sources ({level})"
    else:
        from oracle.oracle import Oracle
        O = Oracle()

        def time(iters):
            return O.time_csr_lb(A.m, A.n, rp, ci, v, x, y, threads, iters)
        kind, what = "port", "kernel_flag=2 restated, OpenMP"
    # reference iteration policy (include/spmv_benchmark.h:185-195), bounded by max_seconds
    t_one = time(1)
    iters = 1000 if t_one < 20.0 else int(min(1000, max(20, 20000.0 / t_one)))
    iters = max(1, min(iters, int(max_seconds * 1000.0 / max(t_one, 1e-3))))
    ms = time(iters)
    return ms, iters, y, kind, what


def serving(matrices, requests, rates, ks=None):
    """The batched serving workload (sspmv_amd/bin/serve_bench): resident
    matrices (256 for serving, 1,024 for the batched launch), requests with k in
    1..32 (or the given list), packed persistent kernel vs per-call CSR."""
    import subprocess
    exe = os.path.join(ROOT, "sspmv_amd", "bin", "serve_bench")
    cmd = [exe, f"--matrices={matrices}", f"--requests={requests}", f"--rates={rates}"]
    if ks:
        cmd.append(f"--ks={ks}")
    log("[serving] " + " ".join(cmd[1:]))
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    for line in out.stderr.splitlines():
        log("[serving] " + line)
    if out.returncode != 0:
        raise RuntimeError(f"serve_bench failed ({out.returncode}): {out.stderr[-2000:]}")
    r = json.loads(out.stdout.strip().splitlines()[-1])
    sat = r["server_saturation"]["req_per_s"]
    r["speedup_vs_percall_async"] = round(sat / r["percall_async"]["req_per_s"], 3)
    r["speedup_vs_percall_sync"] = round(sat / r["percall_sync"]["req_per_s"], 3)
    r["speedup_vs_percall_engine_async"] = round(sat / r["percall_engine_async"]["req_per_s"], 3)
    r["speedup_vs_percall_engine_sync"] = round(sat / r["percall_engine_sync"]["req_per_s"], 3)
    return r


def spmm_section(lib, steps):
    """Multi-vector products on a 4x4-block random matrix (1M rows, 8 blocks per
    block row, the batched workload's block family at full size): BSR 4x4 on
    the matrix cores, BSR scalar, the plain CSR kernel and the CSR row engine
    (4, 8, 16 lanes per row); GFLOP/s and effective GB/s."""
    A = lib.generate(lib.GEN_BLOCKRAND, [262144, 262144, 4, 4, 8], 11)
    M = lib.Matrix.from_host(A)
    M.convert(lib.BSR, lib.make_params(block_rows=4, block_cols=4))
    res = {}
    for k in (4, 16, 32):
        X = lib.DeviceArray.from_numpy(lib.fill_uniform(A.n * k, 12))
        Y = lib.DeviceArray.empty(A.m * k, np.float64)
        row = {}
        for label, fmt, var in (("BSR-4x4-mfma", lib.BSR, lib.VARIANT_AUTO), ("BSR-4x4-scalar", lib.BSR,
                                                                              lib.VARIANT_VECTOR),
                                ("CSR", lib.CSR, lib.VARIANT_AUTO), ("CSR-engine-4", lib.CSR, 18),
                                ("CSR-engine-8", lib.CSR, 19), ("CSR-engine-16", lib.CSR, 20)):
            M.time_spmm(fmt, k, 1.0, X, k, 0.0, Y, k, 3, var)
            ms = M.time_spmm(fmt, k, 1.0, X, k, 0.0, Y, k, steps, var)
            eff = M.effective_bytes(fmt, k, False)
            row[label] = {"us": round(ms * 1e3, 2), "GFLOPs": round(2.0 * A.nnz * k / (ms * 1e-3) / 1e9, 1),
                          "GBs": round(eff / ms / 1e6, 1)}
            log(f"[spmm k={k}] {label}: {ms * 1e3:.1f} us")
        res[f"k{k}"] = row
    return res


def copy_bandwidth(lib, nbytes=1 << 30, iters=20):
    src = lib.DeviceArray(nbytes)
    dst = lib.DeviceArray(nbytes)
    src.fill_bytes(1)
    lib.time_copy(dst, src, nbytes, 2)
    ms = lib.time_copy(dst, src, nbytes, iters)
    return 2.0 * nbytes / (ms * 1e-3) / 1e9, ms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="bound on the CPU baseline's timed work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--formats", default="all", help="comma list of kernel labels to measure, or 'all'")
    ap.add_argument("--sections", default="lognormal,lognormal_fp32,stencil27,spmm,copy,serving,batched")
    ap.add_argument("--serve-matrices", type=int, default=256)
    ap.add_argument("--serve-requests", type=int, default=4000)
    ap.add_argument("--serve-rates", default="5000,20000,50000")
    ap.add_argument("--copy-only", action="store_true", help="only run the plain copy kernel (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build runs the single-matrix "
              "SpMV workload on one MI355X and does not shard it.", file=sys.stderr)
        sys.exit(2)

    from sspmv_amd import lib
    lib.check(lib.load().lespmv_init(0))
    if args.copy_only:
        gbs, ms = copy_bandwidth(lib, iters=5)
        print(json.dumps({"copy_GBs": round(gbs, 1), "copy_bytes_per_launch": 2 * (1 << 30), "copy_ms": ms}))
        return
    sections = set(args.sections.split(","))
    wanted = None if args.formats == "all" else set(args.formats.split(","))

    t0 = time.time()
    A = lib.generate(lib.GEN_LOGNORMAL, LOGNORMAL, 2, np.float64)
    x = lib.fill_uniform(A.n, 4, np.float64)
    log(f"generated lognormal M={A.m} nnz={A.nnz} in {time.time() - t0:.1f}s")
    M = lib.Matrix.from_host(A)
    dx = lib.DeviceArray.from_numpy(x)
    dy = lib.DeviceArray.empty(A.m, np.float64)
    results = measure(lib, M, dx, dy, "lognormal", wanted, args.warmup)
    best = min(results, key=lambda r: r["ms"])  # fastest kernel on this matrix
    fmt, variant = best["fmt"], best["variant"]
    if fmt != lib.CSR:
        M.convert(fmt, best["params"])

    # timed region: K steps bracketed by device syncs; the kernel's own device
    # time comes from hipEvents recorded on the stream it runs on
    M.time_spmv(fmt, 1.0, dx, 0.0, dy, max(args.warmup, 1), variant)
    lib.synchronize()
    t0 = time.perf_counter()
    dev_ms = M.time_spmv(fmt, 1.0, dx, 0.0, dy, args.steps, variant)
    lib.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    eff = best["effective_bytes"]
    y_gpu = dy.numpy()

    out = {
        "metric": "SpMV GB/s & %HBM peak; batched req/s + p99 latency vs plain CSR, same GPU",
        "value": round(eff / (wall_ms * 1e-3) / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"single-matrix SpMV, lognormal M=N={A.m} nnz={A.nnz} (row nnz lognormal mean 16 sigma 1 "
                        "cap 4096; 70% cols uniform, 30% within +-64 of the diagonal), FP64 values, int32 indices",
            "format": best["label"], "m": A.m, "nnz": A.nnz,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(eff / (dev_ms * 1e-3) / 1e9, 2),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(eff / (dev_ms * 1e-3) / HBM_PEAK, 4),
            "traffic": traffic_from_profiles(best["label"]),
            "algorithmic_bytes": eff,
            "kernel_ms": round(dev_ms, 5),
        },
        "formats": {"lognormal_f64": summary(results)},
    }
    if "copy" in sections:
        gbs, _ = copy_bandwidth(lib)
        out["copy_GBs"] = round(gbs, 1)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ms, iters, y_cpu, kind, what = cpu_baseline(A, x, threads, args.cpu_seconds)
        out["cpu_baseline"] = {
            "value": round(M.effective_bytes(lib.CSR) / (ms * 1e-3) / 1e9, 3),
            "unit": "GB/s",
            "cores": threads,
            "kind": kind,
            "sample": f"the full lognormal matrix, {iters} CSR SpMVs ({what}, OpenMP), {ms:.2f} ms each",
        }
        nb = np.linalg.norm(y_cpu)
        out["parity_rel2_vs_cpu"] = float(np.linalg.norm(y_gpu - y_cpu) / (nb if nb else 1.0))
    del M, dx, dy
    if "lognormal_fp32" in sections:
        A32 = lib.generate(lib.GEN_LOGNORMAL, LOGNORMAL, 2, np.float32)
        M = lib.Matrix.from_host(A32)
        dx = lib.DeviceArray.from_numpy(x.astype(np.float32))
        dy = lib.DeviceArray.empty(A32.m, np.float32)
        out["formats"]["lognormal_f32"] = summary(measure(lib, M, dx, dy, "lognormal_fp32", wanted, args.warmup))
        del M, dx, dy, A32
    if "stencil27" in sections:
        S = lib.generate(lib.GEN_STENCIL27, [160, 26.0, -1.0], 0, np.float64)
        M = lib.Matrix.from_host(S)
        dx = lib.DeviceArray.from_numpy(lib.fill_uniform(S.n, 5))
        dy = lib.DeviceArray.empty(S.m, np.float64)
        out["formats"]["stencil27_f64"] = summary(measure(lib, M, dx, dy, "stencil27", wanted, args.warmup))
        del M, dx, dy, S
    if "spmm" in sections:
        out["spmm_blockrand4x4_1M"] = spmm_section(lib, 20)
    if "serving" in sections:
        out["serving"] = serving(args.serve_matrices, args.serve_requests, args.serve_rates)
    if "batched" in sections:
        out["batched"] = serving(1024, 4096, "", "1,4,8,16,32")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
