"""Write tests/golden/lespmv_golden.npz from the compiled reference.

Run in the build container only (needs oracle/_ref/libref_lespmv.so, built by
oracle/build_ref.sh from the reference sources in place). The fixture holds
inputs and the reference's outputs: every conversion's arrays and scalars at
the reference's own parameters (plus GPU-relevant ones) with the reference's
bytes_per_spmv of each converted struct, and y vectors from
LeSpMV_<fmt> with the reference harness's alpha=0.8, beta=0.7 and y0=0
(include/sparse_operation.h:124-138), plus y0 != 0 runs of the serial kernels.

Quirks recorded rather than reproduced:
* CSR5 tile_desc_offset: the last slot of every tile with empty rows is never
  written by the reference (memaligned, uninitialised); a mask marks the
  written slots.
* DIA diag_data is stored only up to 20k entries (diag_offsets always).
* DIA serial bounds rows by num_rows on both sides (src/spmv_dia.cpp:32-33), so
  its y is only meaningful for square matrices; y is stored only for those.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import (BSR, COO, CSR, CSR5, DIA, ELL, SCALARS, SELL, SELL_C_R, SELL_C_SIGMA,  # noqa: E402
                           Reference, array_dtypes, reference_params)
from tests.matrices import feature_cases, suite  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lespmv_golden.npz")
FEATURES_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "features_golden.npz")


# FP32 fixtures are kept for a subset to keep the file small; FP64 covers all.
F32_NAMES = {"lap8x8", "rand_empty", "unsorted", "blocks4", "rect_wide"}


def param_sets(fmt, dt):
    ps = [reference_params(fmt, dt)]
    if fmt == SELL_C_SIGMA:
        ps += [[4, 8, 1, 0], [4, 64, 8, 0], [64, 128, 1, 0], [64, 256, 1, 0]]
    elif fmt == SELL:
        ps += [[64, 0, 1, 0]]
    elif fmt == SELL_C_R:
        ps += [[64, 0, 1, 0]]
    elif fmt == ELL:
        ps += [[1, 0, 0, 0]]
    elif fmt == BSR:
        ps += [[4, 4, 0, 0], [1, 1, 0, 0], [3, 2, 0, 0], [2, 8, 0, 0]]
    elif fmt == DIA:
        ps += [[0, 0, 1, 10240]]
    return ps


MTX_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mtx")


def write_mtx_fixtures():
    """Small MatrixMarket files covering the reader's branches (written here)."""
    os.makedirs(MTX_DIR, exist_ok=True)
    r = np.random.default_rng(99)
    files = {}
    # general real with duplicates, comments and an unsorted entry order
    ent = [(int(r.integers(1, 13)), int(r.integers(1, 10)), float(np.round(r.uniform(-3, 3), 3))) for _ in range(40)]
    ent += [(5, 5, 1.5), (5, 5, -1.5), (7, 2, 0.25), (7, 2, 0.5)]
    files["general_real.mtx"] = ("%%MatrixMarket matrix coordinate real general\n% a comment\n% another\n"
                                 f"12 9 {len(ent)}\n" + "".join(f"{a} {b} {c}\n" for a, b, c in ent))
    # symmetric real: lower triangle with diagonal
    ent = [(i, j, float(np.round(r.uniform(-2, 2), 3))) for i in range(1, 11) for j in range(1, i + 1)
           if r.random() < 0.4 or i == j]
    files["symmetric_real.mtx"] = ("%%MatrixMarket matrix coordinate real symmetric\n"
                                   f"10 10 {len(ent)}\n" + "".join(f"{a} {b} {c}\n" for a, b, c in ent))
    ent = [(int(r.integers(1, 8)), int(r.integers(1, 16))) for _ in range(25)]
    files["pattern_general.mtx"] = ("%%MatrixMarket matrix coordinate pattern general\n"
                                    f"7 15 {len(ent)}\n" + "".join(f"{a} {b}\n" for a, b in ent))
    ent = [(i, j, int(r.integers(-5, 6))) for i in range(1, 9) for j in range(1, i + 1) if r.random() < 0.5]
    files["integer_symmetric.mtx"] = ("%%MatrixMarket matrix coordinate integer symmetric\n"
                                      f"8 8 {len(ent)}\n" + "".join(f"{a} {b} {c}\n" for a, b, c in ent))
    for name, text in files.items():
        with open(os.path.join(MTX_DIR, name), "w") as f:
            f.write(text)
    return sorted(files)


def key(*parts):
    return "/".join(str(p) for p in parts)


def main():
    ref = Reference()
    ref.set_threads(4)
    out = {}
    rng = np.random.default_rng(20240521)
    for dt in (np.float64, np.float32):
        dn = np.dtype(dt).name
        for name, A in suite(dt).items():
            if dt == np.float32 and name not in F32_NAMES:
                continue
            x = rng.uniform(-1.0, 1.0, A.n).astype(dt)
            y0 = rng.uniform(-1.0, 1.0, A.m).astype(dt)
            out[key(dn, name, "m")] = np.int64(A.m)
            out[key(dn, name, "n")] = np.int64(A.n)
            out[key(dn, name, "rp")] = A.rp
            out[key(dn, name, "ci")] = A.ci
            out[key(dn, name, "v")] = A.v
            out[key(dn, name, "x")] = x
            out[key(dn, name, "y0")] = y0
            for flag in (0, 1, 2):
                y = np.zeros(A.m, dt)
                _, res = ref.run(A.m, A.n, A.rp, A.ci, A.v, CSR, [0, 0, 0, 0], flag, 0.8, x, 0.7, y)
                out[key(dn, name, "csr", flag, "y")] = y
            # the reference's bytes_per_spmv of the CSR struct (include/spmv_benchmark.h:43-52)
            out[key(dn, name, "csr", "bytes_per_spmv")] = np.int64(res.scalar("bytes_per_spmv"))
            y = y0.copy()
            ref.run(A.m, A.n, A.rp, A.ci, A.v, CSR, [0, 0, 0, 0], 0, 0.8, x, 0.7, y)
            out[key(dn, name, "csr", 0, "y_beta")] = y
            for fmt in (COO, DIA, ELL, SELL, SELL_C_SIGMA, SELL_C_R, CSR5, BSR):
                if fmt == CSR5 and dt != np.float64:
                    continue  # the reference instantiates LeSpMV_csr5 for double only
                # The reference reads x[-1] for SELL pads (src/spmv_s_ell.cpp:44-48) and x past N in
                # the last BSR block column (src/spmv_bsr.cpp:39-43); zeros around x make those reads
                # deterministic (they are multiplied by zero padding).
                xpad = np.zeros(A.n + 2 + 64, dt)
                xpad[1:A.n + 1] = x
                xv = xpad[1:]
                for pi, prm in enumerate(param_sets(fmt, dt)):
                    y = np.zeros(A.m, dt)
                    rc, res = ref.run(A.m, A.n, A.rp, A.ci, A.v, fmt, prm, 0, 0.8, xv, 0.7, y)
                    base = key(dn, name, fmt, pi)
                    out[key(base, "params")] = np.asarray(prm, np.int64)
                    out[key(base, "rc")] = np.int64(rc)
                    if rc:
                        if fmt == DIA:
                            out[key(base, "complete_ndiags")] = np.int64(res.scalar("complete_ndiags"))
                        continue
                    for arr, adt in array_dtypes(fmt, dt).items():
                        a = res.get(arr, adt)
                        if arr == "diag_data" and a.size > 20000:
                            continue  # keep the fixture small: offsets still pin the diagonal order
                        out[key(base, arr)] = a
                    for s in SCALARS[fmt] + ["bytes_per_spmv"]:
                        out[key(base, s)] = np.int64(res.scalar(s))
                    if fmt == CSR5:
                        tp = res.get("tile_ptr", np.uint32)
                        dop = res.get("tile_desc_offset_ptr", np.int32)
                        mask = np.ones(res.scalar("num_offsets"), dtype=bool)
                        for t in range(res.scalar("p") - 1):
                            if (tp[t] >> 31) & 1 and dop[t + 1] > dop[t]:
                                mask[dop[t + 1] - 1] = False
                        out[key(base, "tile_desc_offset_written")] = mask
                    if fmt == DIA and A.m != A.n:
                        continue
                    out[key(base, "y")] = y
                    if fmt not in (CSR5,):
                        yb = y0.copy()
                        ref.run(A.m, A.n, A.rp, A.ci, A.v, fmt, prm, 0, 0.8, xv, 0.7, yb)
                        out[key(base, "y_beta")] = yb
    # coo_to_csr with duplicates, compact and not (include/sparse_conversion.h:29-91)
    r2 = np.random.default_rng(7)
    m, n, nnz = 40, 30, 300
    ri = r2.integers(0, m, nnz).astype(np.int32)
    ci = r2.integers(0, n, nnz).astype(np.int32)
    v = np.round(r2.uniform(-2, 2, nnz), 1)
    v[r2.random(nnz) < 0.1] = 0.0
    out["coo2csr/m"], out["coo2csr/n"] = np.int64(m), np.int64(n)
    out["coo2csr/ri"], out["coo2csr/ci"], out["coo2csr/v"] = ri, ci, v
    for compact in (0, 1):
        res = ref.coo_to_csr(m, n, ri, ci, v, bool(compact))
        for arr, adt in (("row_offset", np.int32), ("col_index", np.int32), ("values", np.float64)):
            out[key("coo2csr", compact, arr)] = res.get(arr, adt)
    # MatrixMarket inputs through read_csr_matrix (utils/sparse_io.cpp:26-204)
    for name in write_mtx_fixtures():
        path = os.path.join(MTX_DIR, name)
        for compact in (0, 1):
            res = ref.read_mtx(path, np.float64, bool(compact))
            for arr, adt in (("row_offset", np.int32), ("col_index", np.int32), ("values", np.float64)):
                out[key("mtx", name, compact, arr)] = res.get(arr, adt)
            for sc in ("num_rows", "num_cols", "num_nnzs"):
                out[key("mtx", name, compact, sc)] = np.int64(res.scalar(sc))
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {len(out)} arrays, {os.path.getsize(OUT) / 1e6:.2f} MB")


def write_features():
    """Get_features (MTX<int, double>) on the MatrixMarket texts of
    tests.matrices.feature_cases(), which the tests regenerate byte for byte.
    Stores every feature scalar and every image array."""
    import tempfile
    ref = Reference()
    ref.set_threads(4)
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, text in feature_cases().items():
            path = os.path.join(d, name + ".mtx")
            with open(path, "w") as f:
                f.write(text)
            scal, img = ref.features(path)
            names = sorted(scal)
            out[key(name, "scalar_names")] = np.array(names)
            out[key(name, "scalar_values")] = np.array([scal[k] for k in names], dtype=np.float64)
            for k, arr in img.items():
                out[key(name, "img", k)] = arr
    np.savez_compressed(FEATURES_OUT, **out)
    print(f"wrote {FEATURES_OUT}: {len(out)} arrays, {os.path.getsize(FEATURES_OUT) / 1e6:.2f} MB")


if __name__ == "__main__":
    if "--features-only" not in sys.argv:
        main()
    write_features()
