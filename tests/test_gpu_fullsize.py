"""Parity at the benchmark's full sizes, on the GPU against the CPU oracle.

Matrices (the single-matrix workload):
  * lognormal: M = N = 4,194,304, row nnz lognormal (mean 16, sigma 1, cap
    4096), 70 % of columns uniform, 30 % within +-64 of the diagonal
    (nnz = 67,101,604), FP64 and FP32;
  * stencil: the 27-point stencil on a 160^3 grid (M = N = 4,096,000,
    nnz = 109,215,352), FP64.
Every conversion the GPU makes is read back in the reference's layout and
compared bit for bit with the oracle's conversion of the same matrix (the
oracle's O(nnz log nnz) conversions finish in seconds at these sizes); every
kernel's y = 0.8 A x + 0.7 y0 is compared with the FP64 CSR oracle within
1e-12 (FP64) / 1e-5 (FP32) relative 2-norm, and the FP32 error is printed.
Layouts the reference itself refuses (DIA beyond 10,240 diagonals) must be
refused with the same code; ELL of the lognormal matrix would store
8,317,304,832 slots (124 x nnz) and is refused by the fill limit.
"""
import numpy as np
import pytest

from oracle.oracle import BSR, COO, CSR, CSR5, DIA, ELL, SELL, SELL_C_R, SELL_C_SIGMA, reference_params
from tests.conftest import TOL, rel2
from tests.fmt_helpers import compare_exports, lib_params

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

LOGNORMAL = [4194304, 4194304, 16.0, 1.0, 4096, 0.3, 64]


class Case:
    """A full-size matrix on the host and the device, its inputs and the oracle's y."""

    def __init__(self, gpu, oracle, family, params, seed, dt, name):
        self.name, self.dt = name, np.dtype(dt)
        self.A = gpu.generate(family, params, seed, dt)
        A = self.A
        self.x = gpu.fill_uniform(A.n, 4, dt)
        self.y0 = gpu.fill_uniform(A.m, 9, dt)
        self.ref = self.y0.astype(np.float64)
        oracle.spmv_csr(A.m, A.n, A.rp, A.ci, A.v.astype(np.float64), 0.8, self.x.astype(np.float64), 0.7,
                        self.ref)
        self.M = gpu.Matrix.from_host(A)
        self.dx = gpu.DeviceArray.from_numpy(self.x)
        print(f"{name} {self.dt.name}: m={A.m} nnz={A.nnz}", flush=True)

    def check(self, gpu, fmt, variant, ctx):
        dy = gpu.DeviceArray.from_numpy(self.y0)
        self.M.spmv(fmt, 0.8, self.dx, 0.7, dy, variant=variant)
        err = rel2(dy.numpy(), self.ref)
        print(f"  {ctx}: rel2 {err:.3e}", flush=True)
        assert err <= TOL[self.dt], (ctx, err)
        return err


@pytest.fixture(scope="module", params=[np.float64, np.float32], ids=["f64", "f32"])
def lognormal(request, gpu, oracle):
    return Case(gpu, oracle, gpu.GEN_LOGNORMAL, LOGNORMAL, 2, request.param, "lognormal")


@pytest.fixture(scope="module")
def stencil(gpu, oracle):
    return Case(gpu, oracle, gpu.GEN_STENCIL27, [160, 26.0, -1.0], 0, np.float64, "stencil27")


def test_lognormal_csr_kernels(gpu, lognormal):
    c = lognormal
    if c.dt == np.float64:
        # values and indices 805,219,248 + row pointers 16,777,220 + x and y 2 x 33,554,432
        assert c.M.effective_bytes(CSR) == 889105332
    for var in (gpu.VARIANT_AUTO, gpu.VARIANT_VECTOR, gpu.VARIANT_MERGE, gpu.VARIANT_PANEL, gpu.VARIANT_PANEL16,
                gpu.VARIANT_PANEL4, 18, 20):
        c.check(gpu, CSR, var, f"CSR variant {var}")
        c.M.release(CSR)


def convert_both(gpu, oracle, c, fmt, prm):
    A = c.A
    rc, res = oracle.convert(A.m, A.n, A.rp, A.ci, A.v, fmt, prm)
    try:
        c.M.convert(fmt, lib_params(gpu, fmt, prm))
        grc = 0
    except gpu.LespmvError as e:
        grc = e.code
    assert grc == rc, (fmt, prm, grc, rc, res.error())
    return rc, res


LOGNORMAL_FORMATS = [
    (COO, [0, 0, 0, 0]),
    (SELL, "ref"), (SELL, [64, 0, 1, 0]),
    (SELL_C_SIGMA, "ref"), (SELL_C_SIGMA, [64, 4096, 1, 0]),
    (SELL_C_R, "ref"), (SELL_C_R, [64, 0, 1, 0]),
    (CSR5, "ref"), (CSR5, [64, 16, 0, 0]),
    (BSR, [4, 4, 0, 0]),
    (DIA, "ref"),  # more than 10,240 diagonals: refused by both
]


@pytest.mark.parametrize("fmt,prm", LOGNORMAL_FORMATS,
                         ids=[f"{f}-{p if p == 'ref' else '_'.join(map(str, p))}" for f, p in LOGNORMAL_FORMATS])
def test_lognormal_formats(gpu, oracle, lognormal, fmt, prm):
    c = lognormal
    if prm == "ref":
        prm = reference_params(fmt, c.dt)
    ctx = (c.name, c.dt.name, fmt, list(prm))
    rc, res = convert_both(gpu, oracle, c, fmt, prm)
    if fmt == DIA:
        assert rc == gpu.ERR_MAX_DIAGS, ctx
        return
    assert rc == 0, ctx
    mask = np.ones(res.scalar("num_offsets"), bool) if fmt == CSR5 else None
    compare_exports(c.M, res, fmt, c.dt, ctx, mask)
    del res
    variants = (0, 3) if fmt == COO else (0, 1) if fmt == BSR else (0,)  # COO 3: atomics; BSR 1: row kernel
    for var in variants:
        c.check(gpu, fmt, var, f"{ctx} variant {var}")
    c.M.release(fmt)


def test_lognormal_ell_refused(gpu, lognormal):
    c = lognormal
    with pytest.raises(gpu.LespmvError) as e:
        c.M.convert(ELL, gpu.make_params(chunk=64, alignment=1, max_fill=8))
    assert e.value.code == gpu.ERR_UNSUPPORTED
    assert "8317304832 stored slots" in str(e.value)


def test_stencil_csr_kernels(gpu, stencil):
    for var in (gpu.VARIANT_AUTO, gpu.VARIANT_VECTOR, gpu.VARIANT_MERGE, gpu.VARIANT_PANEL, 19):
        stencil.check(gpu, CSR, var, f"CSR variant {var}")
        stencil.M.release(CSR)


STENCIL_FORMATS = [
    (DIA, "ref", (1, 2)), (DIA, [0, 0, 1, 10240], (1, 2)),
    (BSR, [4, 4, 0, 0], (0, 1)), (BSR, "ref", (0,)),
    (SELL_C_SIGMA, [64, 4096, 1, 0], (0,)), (SELL_C_SIGMA, "ref", (0,)),
    (SELL, [64, 0, 1, 0], (0,)), (COO, [0, 0, 0, 0], (0, 3)), (CSR5, [64, 16, 0, 0], (0,)),
]


@pytest.mark.parametrize("fmt,prm,variants", STENCIL_FORMATS,
                         ids=[f"{f}-{p if p == 'ref' else '_'.join(map(str, p))}" for f, p, _ in STENCIL_FORMATS])
def test_stencil_formats(gpu, oracle, stencil, fmt, prm, variants):
    c = stencil
    if prm == "ref":
        prm = reference_params(fmt, c.dt)
    ctx = (c.name, fmt, list(prm))
    rc, res = convert_both(gpu, oracle, c, fmt, prm)
    assert rc == 0, ctx
    mask = np.ones(res.scalar("num_offsets"), bool) if fmt == CSR5 else None
    compare_exports(c.M, res, fmt, c.dt, ctx, mask)
    del res
    for var in variants:
        c.check(gpu, fmt, var, f"{ctx} variant {var}")
    c.M.release(fmt)
