#!/usr/bin/env python3
"""Single-GPU SpMV benchmark: the all-formats single-matrix workload.

Workload (one matrix per section, FP64 unless stated, int32 indices):
  * lognormal: M = N = 4,194,304, per-row nnz lognormal (mean 16, sigma 1,
    capped at 4096), 70 % of columns uniform over [0, N) and 30 % within +-64 of
    the diagonal, values U[-1,1), x = U[-1,1) (seed 4); FP64 (headline) and FP32;
  * stencil27: 27-point stencil on a 160^3 grid (diag 26, off-diagonals -1) for
    the DIA and BSR paths, FP64 and FP32;
  * laplace2d: the 5-point Laplacian on a 1024^2 grid (FP64), timed rotating
    over four device copies so that every call reads HBM, with the reference's
    CPU kernel on the same matrix beside it.
One step = one y = A*x with the matrix and vectors resident in HBM. Every
kernel's y is compared with the reference's FP64 CPU result on the same
inputs (rel2_vs_fp64_reference; FP32 kernels against the FP64 result).

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
import re
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
    csr = [("CSR-auto", lib.CSR, lib.VARIANT_AUTO, None), ("CSR-vector", lib.CSR, lib.VARIANT_VECTOR, None),
           ("CSR-merge", lib.CSR, lib.VARIANT_MERGE, None), ("CSR-rowtiles", lib.CSR, lib.VARIANT_ROW_TILES, None)]
    if section.startswith("stencil27"):
        return csr + [("SELL-64-sigma4096", lib.SELL_C_SIGMA, 0, sell(64, 4096)),
                      ("DIA", lib.DIA, lib.VARIANT_VECTOR, lib.make_params(alignment=1, max_diags=10240)),
                      ("DIA-lds", lib.DIA, lib.VARIANT_MERGE, lib.make_params(alignment=1, max_diags=10240)),
                      ("BSR-4x4", lib.BSR, 0, lib.make_params(block_rows=4, block_cols=4))]
    if section == "laplace2d":
        return csr + [("SELL-64-sigma4096", lib.SELL_C_SIGMA, 0, sell(64, 4096)),
                      ("DIA", lib.DIA, lib.VARIANT_VECTOR, lib.make_params(alignment=1, max_diags=10240))]
    # lognormal (FP64 and FP32): ELL and S-ELL are refused beyond 8 x nnz stored slots (the refusal
    # carries the padded size); DIA does not apply (more than 10,240 diagonals)
    return csr + [("CSR-panel8", lib.CSR, lib.VARIANT_PANEL, None), ("CSR-panel16", lib.CSR, lib.VARIANT_PANEL16, None),
                  ("CSR-panel4", lib.CSR, lib.VARIANT_PANEL4, None),
                  ("ELL", lib.ELL, 0, lib.make_params(chunk=64, alignment=1, max_fill=8)),
                  ("S-ELL-64", lib.SELL, 0, lib.make_params(chunk=64, sigma=1, alignment=1, max_fill=8)),
                  ("SELL-64-sigma4096", lib.SELL_C_SIGMA, 0, sell(64, 4096)), ("SELL-64-R", lib.SELL_C_R, 0,
                                                                                sell(64, 0)),
                  ("COO", lib.COO, 0, None), ("CSR5-64x16", lib.CSR5, 0, lib.make_params(omega=64, sigma=16))]


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


def rel2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb else 1.0))


def measure(lib, M, dx, dy, section, wanted, warmup, yref=None):
    """Each kernel of the section: convert (or prepare the CSR variant's layout),
    time 50 calls with hipEvents, and (with yref, the FP64 reference y) the
    relative 2-norm error of its y. A conversion the library refuses is recorded
    with its reason (for ELL / S-ELL the padded slot count)."""
    results, refused = [], {}
    for label, fmt, variant, params in kernels(lib, section):
        if wanted and label not in wanted:
            continue
        try:
            if fmt != lib.CSR:
                M.convert(fmt, params)
                conv_ms = M.convert_ms(fmt)
            else:
                conv_ms = M.prepare(fmt, variant)  # partition / column-panel layout, timed
        except lib.LespmvError as e:
            msg = str(e)
            log(f"[{section}] {label}: refused ({msg})")
            slots = re.search(r"(\d+) stored slots", msg)
            refused[label] = {"refused": msg, "code": e.code}
            if slots:
                vb = 4 if section.endswith("fp32") else 8
                refused[label]["padded_slots"] = int(slots.group(1))
                refused[label]["padded_bytes"] = int(slots.group(1)) * (4 + vb)
            continue
        eff = M.effective_bytes(fmt, 1, False)
        try:  # the reference's own formula (include/spmv_benchmark.h:26-160), for comparison only
            ref_bytes = M.reference_bytes(fmt)
        except lib.LespmvError:
            ref_bytes = None
        M.time_spmv(fmt, 1.0, dx, 0.0, dy, max(warmup, 1), variant)
        ms = M.time_spmv(fmt, 1.0, dx, 0.0, dy, 50, variant)
        err = rel2(dy.numpy(), yref) if yref is not None else None  # y = A x of the last timed call
        results.append(dict(label=label, fmt=fmt, variant=variant, params=params, ms=ms, gbs=eff / ms / 1e6,
                            effective_bytes=eff, convert_ms=conv_ms, format_bytes=M.format_bytes(fmt),
                            ref_gbs=ref_bytes / ms / 1e6 if ref_bytes else None, rel2=err))
        log(f"[{section}] {label}: {ms * 1e3:.1f} us  {eff / ms / 1e6:.0f} GB/s  convert {conv_ms:.1f} ms"
            + (f"  rel2 {err:.2e}" if err is not None else ""))
        M.release(fmt)  # CSR: frees the panel layout (the CSR arrays stay)
    return results, refused


def summary(results, refused=None):
    out = {r["label"]: {"us": round(r["ms"] * 1e3, 2), "GBs": round(r["gbs"], 1),
                        "frac": round(r["gbs"] * 1e9 / HBM_PEAK, 4), "convert_ms": round(r["convert_ms"], 2),
                        "format_bytes": r["format_bytes"],
                        "GBs_reference_formula": round(r["ref_gbs"], 1) if r["ref_gbs"] else None,
                        **({"rel2_vs_fp64_reference": r["rel2"]} if r.get("rel2") is not None else {})}
           for r in results}
    out.update(refused or {})
    return out


def host_info():
    """CPU model and topology of this host (lscpu), and the OpenMP thread setting."""
    info = {"omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except Exception as e:  # lscpu absent: keep what /proc/cpuinfo says
        info["lscpu_error"] = str(e)
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        info["model"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return info


def laplace_section(lib, steps, warmup, cpu, cpu_seconds, threads, copies=4):
    """Config 1: 5-point Laplacian on a 1024 x 1024 grid (N = 1,048,576,
    nnz = 5,238,784, diag 4, off-diagonal -1), FP64, x = U[-1,1) seed 1. One
    matrix moves 83.84 MB per SpMV, which the 256 MB Infinity Cache would hold,
    so the HBM figure rotates over `copies` device copies (matrix, x and y each
    their own, 335 MB together); the single-copy (cache-resident) time is
    reported beside it. The CPU side is the reference's LeSpMV_csr
    kernel_flag=2 (benchmark_spmv_csr.cpp:109 -> test_spmv_csr.cpp:14-87) on
    the same matrix, on this host's threads."""
    A = lib.generate(lib.GEN_LAPLACE2D, [1024, 1024, 4.0, -1.0], 0, np.float64)
    x = lib.fill_uniform(A.n, 1, np.float64)
    mats = [lib.Matrix.from_host(A) for _ in range(copies)]
    xs = [lib.DeviceArray.from_numpy(x) for _ in range(copies)]
    ys = [lib.DeviceArray.empty(A.m, np.float64) for _ in range(copies)]
    out = {"workload": f"5-point Laplacian 1024x1024 (N={A.m}, nnz={A.nnz}), FP64, int32 indices",
           "rotation_copies": copies, "kernels": {}}
    yref = None
    if cpu:
        ms, iters, y_cpu, kind, what = cpu_baseline(A, x, threads, cpu_seconds)
        yref = y_cpu
        eff = mats[0].effective_bytes(lib.CSR)
        refb = mats[0].reference_bytes(lib.CSR)
        out["cpu"] = {"ms": round(ms, 5), "GFLOPs": round(2.0 * A.nnz / (ms * 1e-3) / 1e9, 2),
                      "GBs": round(eff / (ms * 1e-3) / 1e9, 2), "GBs_reference_formula": round(refb / (ms * 1e-3) / 1e9, 2),
                      "iterations": iters, "threads": threads, "kind": kind, "kernel": what, "host": host_info()}
        log(f"[laplace2d] CPU {what}: {ms:.3f} ms  {eff / (ms * 1e-3) / 1e9:.1f} GB/s")
    for label, fmt, variant, params in kernels(lib, "laplace2d"):
        if fmt != lib.CSR:
            for M in mats:
                M.convert(fmt, params)
        M0 = mats[0]
        eff = M0.effective_bytes(fmt, 1, False)
        lib.time_spmv_rotating(mats, fmt, 1.0, xs, 0.0, ys, max(1, warmup // 4), variant)
        ms_hbm = lib.time_spmv_rotating(mats, fmt, 1.0, xs, 0.0, ys, steps, variant)
        ms_cache = M0.time_spmv(fmt, 1.0, xs[0], 0.0, ys[0], steps, variant)
        row = {"us": round(ms_hbm * 1e3, 2), "GBs": round(eff / ms_hbm / 1e6, 1),
               "frac": round(eff / ms_hbm / 1e6 * 1e9 / HBM_PEAK, 4),
               "us_cache_resident": round(ms_cache * 1e3, 2), "GBs_cache_resident": round(eff / ms_cache / 1e6, 1)}
        if yref is not None:
            row["rel2_vs_fp64_reference"] = rel2(ys[0].numpy(), yref)
        out["kernels"][label] = row
        log(f"[laplace2d] {label}: {ms_hbm * 1e3:.1f} us rotating ({eff / ms_hbm / 1e6:.0f} GB/s), "
            f"{ms_cache * 1e3:.1f} us one copy")
        for M in mats:
            M.release(fmt)
    if yref is not None and out["kernels"]:
        best = min(out["kernels"].items(), key=lambda kv: kv[1]["us"])
        out["gpu_over_cpu"] = round(out["cpu"]["ms"] * 1e3 / best[1]["us"], 1)
    return out


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


def headline(lib, args, A, M, dx, dy, results, refused, y_cpu, cpu, threads):
    """The headline line: the fastest kernel on the FP64 lognormal matrix, timed
    over the K steps (wall clock between device syncs, and hipEvents on the
    kernel's stream), with its roofline and the CPU baseline."""
    best = min(results, key=lambda r: r["ms"])  # fastest kernel on this matrix
    fmt, variant = best["fmt"], best["variant"]
    if fmt != lib.CSR:
        M.convert(fmt, best["params"])
    else:
        M.prepare(fmt, variant)  # the layout is built before the timed region

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
        "formats": {"lognormal_f64": summary(results, refused)},
    }
    if cpu is not None:
        ms, iters, kind, what = cpu
        out["cpu_baseline"] = {
            "value": round(M.effective_bytes(lib.CSR) / (ms * 1e-3) / 1e9, 3),
            "unit": "GB/s",
            "cores": threads,
            "kind": kind,
            "sample": f"the full lognormal matrix, {iters} CSR SpMVs ({what}, OpenMP), {ms:.2f} ms each",
            "host": host_info(),
        }
        out["parity_rel2_vs_cpu"] = rel2(y_gpu, y_cpu)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="bound on the CPU baseline's timed work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--formats", default="all", help="comma list of kernel labels to measure, or 'all'")
    ap.add_argument("--sections",
                    default="lognormal,lognormal_fp32,stencil27,stencil27_fp32,laplace2d,spmm,copy,serving,batched")
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

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.time()
    A = lib.generate(lib.GEN_LOGNORMAL, LOGNORMAL, 2, np.float64)
    x = lib.fill_uniform(A.n, 4, np.float64)
    log(f"generated lognormal M={A.m} nnz={A.nnz} in {time.time() - t0:.1f}s")
    # CPU baseline first: its y (the reference's FP64 kernel) is the reference
    # every GPU kernel's error is stated against, FP32 ones included
    y_cpu, cpu = None, None
    if not args.no_cpu:
        ms, iters, y_cpu, kind, what = cpu_baseline(A, x, threads, args.cpu_seconds)
        cpu = (ms, iters, kind, what)
    M = lib.Matrix.from_host(A)
    dx = lib.DeviceArray.from_numpy(x)
    dy = lib.DeviceArray.empty(A.m, np.float64)
    results, refused = measure(lib, M, dx, dy, "lognormal", wanted, args.warmup, y_cpu)
    if not results:  # a --formats selection without lognormal kernels: the other sections only
        out = {"formats": {}, "note": "no lognormal kernel selected: no headline"}
    else:
        out = headline(lib, args, A, M, dx, dy, results, refused, y_cpu, cpu, threads)
    del M, dx, dy
    if "copy" in sections:
        gbs, _ = copy_bandwidth(lib)
        out["copy_GBs"] = round(gbs, 1)
    if "lognormal_fp32" in sections:
        A32 = lib.generate(lib.GEN_LOGNORMAL, LOGNORMAL, 2, np.float32)
        M = lib.Matrix.from_host(A32)
        dx = lib.DeviceArray.from_numpy(x.astype(np.float32))
        dy = lib.DeviceArray.empty(A32.m, np.float32)
        out["formats"]["lognormal_f32"] = summary(*measure(lib, M, dx, dy, "lognormal_fp32", wanted, args.warmup,
                                                           y_cpu))
        del M, dx, dy, A32
    del A, y_cpu
    for dt, key in ((np.float64, "stencil27"), (np.float32, "stencil27_fp32")):
        if key not in sections:
            continue
        S = lib.generate(lib.GEN_STENCIL27, [160, 26.0, -1.0], 0, dt)
        M = lib.Matrix.from_host(S)
        dx = lib.DeviceArray.from_numpy(lib.fill_uniform(S.n, 5, dt))
        dy = lib.DeviceArray.empty(S.m, dt)
        name = "stencil27_f64" if dt == np.float64 else "stencil27_f32"
        out["formats"][name] = summary(*measure(lib, M, dx, dy, key, wanted, args.warmup))
        del M, dx, dy, S
    if "laplace2d" in sections:
        out["laplace2d_1024"] = laplace_section(lib, max(args.steps, 50), args.warmup, not args.no_cpu,
                                                args.cpu_seconds, threads)
    if "spmm" in sections:
        out["spmm_blockrand4x4_1M"] = spmm_section(lib, 20)
    if "serving" in sections:
        out["serving"] = serving(args.serve_matrices, args.serve_requests, args.serve_rates)
    if "batched" in sections:
        out["batched"] = serving(1024, 4096, "", "1,4,8,16,32")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
