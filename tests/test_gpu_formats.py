"""Every non-CSR format on the GPU: conversion parity and SpMV/SpMM parity.

Conversions run on the GPU and are read back in the reference layout; every
index and value array must equal the oracle's (itself pinned to the compiled
reference by tests/test_oracle_golden.py) bit for bit, at the reference's
parameters and at GPU-native ones. y must match the FP64 CSR oracle to 1e-12
(FP64) / 1e-5 (FP32) relative 2-norm.
"""
import numpy as np
import pytest

from oracle.oracle import BSR, COO, CSR, CSR5, DIA, ELL, SELL, SELL_C_R, SELL_C_SIGMA
from tests.conftest import TOL, rel2
from tests.fmt_helpers import compare_exports, lib_params, param_sets
from tests.matrices import HostCsr, suite

pytestmark = pytest.mark.gpu

FORMATS = [ELL, SELL, SELL_C_SIGMA, SELL_C_R, COO, DIA, BSR, CSR5]


def oracle_y(oracle, A, alpha, x, beta, y0):
    y = y0.astype(np.float64).copy()
    oracle.spmv_csr(A.m, A.n, A.rp, A.ci, A.v.astype(np.float64), alpha, x.astype(np.float64), beta, y)
    return y


def available(gpu, fmt):
    A = suite()["lap8x8"]
    M = gpu.Matrix.from_host(A)
    try:
        M.convert(fmt, None)
        return True
    except gpu.LespmvError as e:
        if e.code == gpu.ERR_UNSUPPORTED:
            return False
        raise


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_format_conversion_and_spmv(gpu, oracle, fmt, dt):
    if not available(gpu, fmt):
        pytest.skip(f"format {fmt} not built in this library")
    rng = np.random.default_rng(fmt * 10 + (dt == np.float32))
    worst = 0.0
    for name, A in suite(dt).items():
        M = gpu.Matrix.from_host(A)
        x = rng.uniform(-1, 1, A.n).astype(dt)
        y0 = rng.uniform(-1, 1, A.m).astype(dt)
        ref = oracle_y(oracle, A, 0.8, x, 0.7, y0)
        ref0 = oracle_y(oracle, A, 1.0, x, 0.0, y0)
        for prm in param_sets(fmt, dt):
            ctx = (name, np.dtype(dt).name, fmt, list(prm))
            rc, res = oracle.convert(A.m, A.n, A.rp, A.ci, A.v, fmt, prm)
            try:
                M.convert(fmt, lib_params(gpu, fmt, prm))
                grc = 0
            except gpu.LespmvError as e:
                grc = e.code
            assert grc == rc, (ctx, grc, rc, res.error())
            if rc:
                continue
            mask = None
            if fmt == CSR5:
                mask = np.ones(res.scalar("num_offsets"), bool)  # every slot is defined by the oracle
            compare_exports(M, res, fmt, dt, ctx, mask)
            dx = gpu.DeviceArray.from_numpy(x)
            for variant in (0, 1, 2) if fmt in (DIA, COO) else (0,):
                dy = gpu.DeviceArray.from_numpy(y0)
                M.spmv(fmt, 0.8, dx, 0.7, dy, variant=variant)
                err = rel2(dy.numpy(), ref)
                worst = max(worst, err)
                assert err <= TOL[np.dtype(dt)], (ctx, variant, err)
                dy = gpu.DeviceArray.from_numpy(np.full(A.m, np.nan, dt))
                M.spmv(fmt, 1.0, dx, 0.0, dy, variant=variant)
                y = dy.numpy()
                assert np.all(np.isfinite(y)), (ctx, variant)
                assert rel2(y, ref0) <= TOL[np.dtype(dt)], (ctx, variant)
            M.release(fmt)
    print(f"format {fmt} {np.dtype(dt).name}: worst rel2 vs FP64 oracle {worst:.3e}")


@pytest.mark.parametrize("fmt", FORMATS)
def test_format_matches_reference_fixture(gpu, golden, fmt):
    """GPU conversion vs the compiled reference's arrays stored in the fixture."""
    if not available(gpu, fmt):
        pytest.skip(f"format {fmt} not built in this library")
    from oracle.oracle import SCALARS, array_dtypes
    n = 0
    for k in golden.files:
        p = k.split("/")
        if len(p) != 5 or p[4] != "params" or int(p[2]) != fmt:
            continue
        dn, name, _, pi = p[:4]
        base = f"{dn}/{name}/{fmt}/{pi}"
        if int(golden[f"{base}/rc"]) != 0:
            continue
        g = lambda s: golden[f"{dn}/{name}/{s}"]
        A = HostCsr(int(g("m")), int(g("n")), g("rp"), g("ci"), g("v"))
        M = gpu.Matrix.from_host(A)
        M.convert(fmt, lib_params(gpu, fmt, golden[f"{base}/params"]))
        for arr, adt in array_dtypes(fmt, A.v.dtype).items():
            if f"{base}/{arr}" not in golden.files:
                continue
            got, exp = M.export(fmt, arr, adt), golden[f"{base}/{arr}"]
            if arr == "tile_desc_offset":
                w = golden[f"{base}/tile_desc_offset_written"]
                got, exp = got[w], exp[w]
            assert got.shape == exp.shape and np.array_equal(got, exp), (base, arr)
            n += 1
        for s in SCALARS[fmt]:
            assert M.export_scalar(fmt, s) == int(golden[f"{base}/{s}"]), (base, s)
        if f"{base}/y" in golden.files:
            dx = gpu.DeviceArray.from_numpy(g("x"))
            dy = gpu.DeviceArray.from_numpy(np.zeros(A.m, A.v.dtype))
            M.spmv(fmt, 0.8, dx, 0.7, dy)
            assert rel2(dy.numpy(), golden[f"{base}/y"]) <= TOL[A.v.dtype], base
    assert n > 0


def test_sell_sigma_must_divide(gpu):
    A = suite()["rand200"]
    M = gpu.Matrix.from_host(A)
    with pytest.raises(gpu.LespmvError) as e:
        M.convert(SELL_C_SIGMA, gpu.make_params(chunk=4, sigma=6, alignment=8))
    assert e.value.code == gpu.ERR_UNSUPPORTED


@pytest.mark.parametrize("chunk", [64, 32, 4])
def test_sell_pads_never_multiply_x(gpu, chunk):
    """Pads (col -1, value 0) add nothing even where x holds inf: only row 0
    uses column 0 and x[0] = inf, so every other row stays finite and exact."""
    rng = np.random.default_rng(chunk)
    m = n = 300
    rows = [[0, 1]] + [sorted(set([i] + list(rng.integers(1, n, i % 7)))) for i in range(1, m)]
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate([np.array(r, np.int32) for r in rows])
    v = rng.uniform(0.5, 1.5, len(ci))
    A = HostCsr(m, n, rp, ci, v)
    x = rng.uniform(-1, 1, n)
    x[0] = np.inf
    ref = np.array([sum(v[j] * x[ci[j]] for j in range(rp[i], rp[i + 1])) for i in range(m)])
    M = gpu.Matrix.from_host(A)
    for fmt, prm in ((SELL_C_SIGMA, gpu.make_params(chunk=chunk, sigma=4 * chunk, alignment=1)),
                     (SELL_C_R, gpu.make_params(chunk=chunk, alignment=1))):
        M.convert(fmt, prm)
        dx = gpu.DeviceArray.from_numpy(x)
        dy = gpu.DeviceArray.from_numpy(np.zeros(m))
        M.spmv(fmt, 1.0, dx, 0.0, dy)
        y = dy.numpy()
        assert np.isposinf(y[0]), (fmt, chunk, y[0])
        assert np.all(np.isfinite(y[1:])), (fmt, chunk)
        assert rel2(y[1:], ref[1:]) <= 1e-12, (fmt, chunk)
        M.release(fmt)


def test_not_converted_is_an_error(gpu):
    A = suite()["rand200"]
    M = gpu.Matrix.from_host(A)
    dx = gpu.DeviceArray.from_numpy(np.ones(A.n))
    dy = gpu.DeviceArray.from_numpy(np.ones(A.m))
    with pytest.raises(gpu.LespmvError) as e:
        M.spmv(SELL_C_SIGMA, 1.0, dx, 0.0, dy)
    assert e.value.code == gpu.ERR_NOT_CONVERTED


@pytest.mark.parametrize("fmt", [ELL, SELL, SELL_C_SIGMA, SELL_C_R, COO, DIA, BSR, CSR5])
@pytest.mark.parametrize("k", [1, 4, 8, 16, 32, 33, 64])
def test_format_spmm_columns_equal_spmv(gpu, oracle, fmt, k):
    if not available(gpu, fmt):
        pytest.skip(f"format {fmt} not built in this library")
    A = suite()["banded"] if fmt == DIA else suite()["rand_long"]
    M = gpu.Matrix.from_host(A)
    M.convert(fmt, None)
    rng = np.random.default_rng(k + 100 * fmt)
    X = rng.uniform(-1, 1, (A.n, k))
    Y0 = rng.uniform(-1, 1, (A.m, k))
    dX, dY = gpu.DeviceArray.from_numpy(X), gpu.DeviceArray.from_numpy(Y0)
    M.spmm(fmt, k, 0.8, dX, k, 0.7, dY, k)
    Y = dY.numpy()
    for j in range(k):
        ref = oracle_y(oracle, A, 0.8, X[:, j].copy(), 0.7, Y0[:, j].copy())
        assert rel2(Y[:, j], ref) <= 1e-12, (fmt, k, j)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_csr5_spmm_columns_bitwise(gpu, dt):
    """CSR5 SpMM runs the tile kernels once per right-hand side: column j of Y is
    bit for bit the CSR5 SpMV of column j, at the reference's omega and at
    omega = 64, with a leading dimension beyond k whose extra column stays
    untouched; tiles that span rows and tiles of empty rows included."""
    rng = np.random.default_rng(5)
    for name in ("rand_long", "rand_empty", "rect_tall"):
        A = suite(dt)[name]
        M = gpu.Matrix.from_host(A)
        for omega in (64 // np.dtype(dt).itemsize, 64):
            M.convert(CSR5, gpu.make_params(omega=omega, sigma=16))
            for k in (1, 3, 8):
                ld = k + 1
                X = rng.uniform(-1, 1, (A.n, ld)).astype(dt)
                Y0 = rng.uniform(-1, 1, (A.m, ld)).astype(dt)
                dX, dY = gpu.DeviceArray.from_numpy(X), gpu.DeviceArray.from_numpy(Y0)
                M.spmm(CSR5, k, 0.8, dX, ld, 0.7, dY, ld)
                Y = dY.numpy()
                assert np.array_equal(Y[:, k:], Y0[:, k:])
                for j in range(k):
                    dx = gpu.DeviceArray.from_numpy(np.ascontiguousarray(X[:, j]))
                    dy = gpu.DeviceArray.from_numpy(np.ascontiguousarray(Y0[:, j]))
                    M.spmv(CSR5, 0.8, dx, 0.7, dy)
                    np.testing.assert_array_equal(Y[:, j], dy.numpy(), err_msg=f"{name} omega {omega} k {k} j {j}")
            M.release(CSR5)


@pytest.mark.parametrize("fmt", [SELL_C_SIGMA, SELL_C_R])
def test_generated_lognormal_sliced(gpu, oracle, fmt):
    """A 200k-row lognormal matrix (config-2 family): GPU sigma sort vs oracle,
    at the reference parameters and at C=64."""
    A = gpu.generate(gpu.GEN_LOGNORMAL, [200000, 200000, 16.0, 1.0, 4096, 0.3, 64], 2)
    M = gpu.Matrix.from_host(A)
    x = gpu.fill_uniform(A.n, 4)
    ref = oracle_y(oracle, A, 1.0, x, 0.0, np.zeros(A.m))
    for prm in ([4, 16384, 8, 0], [64, 4096, 1, 0]):
        rc, res = oracle.convert(A.m, A.n, A.rp, A.ci, A.v, fmt, prm)
        M.convert(fmt, lib_params(gpu, fmt, prm))
        assert np.array_equal(M.export(fmt, "reorder", np.int32), res.get("reorder", np.int32))
        assert np.array_equal(M.export(fmt, "chunk_len", np.int32), res.get("chunk_len", np.int32))
        dx = gpu.DeviceArray.from_numpy(x)
        dy = gpu.DeviceArray.empty(A.m, np.float64)
        M.spmv(fmt, 1.0, dx, 0.0, dy)
        assert rel2(dy.numpy(), ref) <= 1e-12
        M.release(fmt)


def closed_form_bytes(fmt, A, res, sv, k=1, beta_nz=False):
    """Algorithmic bytes of one product (values + indices describing the matrix,
    x read once, y written, y read when beta != 0), from the counts of the
    oracle's conversion: the definitions the benchmark's roofline uses."""
    si = 4
    vec = k * (A.n + A.m * (2 if beta_nz else 1)) * sv
    nnz = len(A.ci)
    if fmt == CSR:
        return nnz * (sv + si) + (A.m + 1) * si + vec
    if fmt == COO:
        return nnz * (sv + 2 * si) + vec
    if fmt == ELL:
        return nnz * (sv + si) + vec
    if fmt == SELL:
        return nnz * (sv + si) + res.scalar("chunk_num") * si + vec
    if fmt in (SELL_C_SIGMA, SELL_C_R):
        return nnz * (sv + si) + res.scalar("validchunkNum") * si + A.m * si + vec
    if fmt == DIA:
        return nnz * sv + res.scalar("complete_ndiags") * 8 + vec
    if fmt == BSR:
        return nnz * sv + res.scalar("nnzb") * si + (res.scalar("mb") + 1) * si + vec
    if fmt == CSR5:
        p = res.scalar("p")
        return (nnz * (sv + si) + (p + 1) * 4 + p * res.scalar("omega") * res.scalar("num_packets") * 4
                + res.scalar("num_offsets") * si + vec)
    raise ValueError(fmt)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_byte_formulas(gpu, oracle, golden, dt):
    """lespmv_reference_bytes equals the compiled reference's bytes_per_spmv
    (include/spmv_benchmark.h:26-150, stored in the fixture for every matrix,
    format and parameter set); lespmv_effective_bytes equals the closed forms
    above for k = 1 and 4, beta zero and not."""
    dn = np.dtype(dt).name
    sv = np.dtype(dt).itemsize
    names = sorted({k.split("/")[1] for k in golden.files if k.startswith(dn + "/") and k.endswith("/m")})
    checked = 0
    for name in names:
        g = lambda s: golden[f"{dn}/{name}/{s}"]
        A = HostCsr(int(g("m")), int(g("n")), g("rp"), g("ci"), g("v"))
        M = gpu.Matrix.from_host(A)
        assert M.reference_bytes(CSR) == int(g("csr/bytes_per_spmv")), name
        for k, bz in ((1, False), (1, True), (4, True)):
            assert M.effective_bytes(CSR, k, bz) == closed_form_bytes(CSR, A, None, sv, k, bz), (name, k, bz)
        for fmt in FORMATS:
            pi = 0
            while f"{dn}/{name}/{fmt}/{pi}/params" in golden.files:
                base = f"{dn}/{name}/{fmt}/{pi}"
                pi += 1
                if int(golden[f"{base}/rc"]) != 0:
                    continue
                prm = golden[f"{base}/params"]
                rc, res = oracle.convert(A.m, A.n, A.rp, A.ci, A.v, fmt, prm)
                assert rc == 0, base
                M.convert(fmt, lib_params(gpu, fmt, prm))
                assert M.reference_bytes(fmt) == int(golden[f"{base}/bytes_per_spmv"]), base
                for k, bz in ((1, False), (1, True), (4, False)):
                    want = closed_form_bytes(fmt, A, res, sv, k, bz)
                    assert M.effective_bytes(fmt, k, bz) == want, (base, k, bz)
                M.release(fmt)
                checked += 1
    assert checked > 50
    print(f"{dn}: byte formulas checked on {checked} conversions")
