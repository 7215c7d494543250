"""CSR with 64-bit row offsets (lespmv_create_csr64) on the GPU: the products
of the 32-bit path within 1e-12 (FP64) / 1e-5 (FP32) of the FP64 oracle,
SpMM column by column with strided X and Y, the refusal of every call such a
handle does not run, and one matrix beyond 2^31 - 1 nonzeros checked exactly
against a closed form."""
import numpy as np
import pytest

from tests.conftest import TOL, rel2

pytestmark = pytest.mark.gpu


def small_matrices(lib, dt):
    return [lib.generate(lib.GEN_LOGNORMAL, [3000, 2800, 16.0, 1.0, 4096, 0.3, 64], 2, dt),
           lib.generate(lib.GEN_LAPLACE2D, [60, 50, 4.0, -1.0], 0, dt),
           lib.generate(lib.GEN_POWERLAW, [5000, 2.0, 8.0, 2000], 4, dt),
           lib.generate(lib.GEN_STENCIL27, [14, 26.0, -1.0], 5, dt)]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_csr64_matches_oracle(gpu, oracle, dt):
    lib = gpu
    rng = np.random.default_rng(3)
    for A in small_matrices(lib, dt):
        M = lib.Matrix.csr64(A.m, A.n, A.rp.astype(np.int64), A.ci, A.v)
        x = rng.uniform(-1, 1, A.n).astype(dt)
        y0 = rng.uniform(-1, 1, A.m).astype(dt)
        dx, dy = lib.DeviceArray.from_numpy(x), lib.DeviceArray.from_numpy(y0)
        ref = y0.astype(np.float64)
        oracle.spmv_csr(A.m, A.n, A.rp, A.ci, A.v.astype(np.float64), 0.8, x.astype(np.float64), 0.7, ref)
        # every CSR variant runs on the row blocks: auto, vector, merge, vector with 4 lanes a row
        for variant in (lib.VARIANT_AUTO, lib.VARIANT_VECTOR, lib.VARIANT_MERGE, 18):
            dy = lib.DeviceArray.from_numpy(y0)
            M.spmv(lib.CSR, 0.8, dx, 0.7, dy, variant=variant)
            assert rel2(dy.numpy(), ref) <= TOL[np.dtype(dt)], (A.m, A.nnz, variant)
        # SpMM (the row engine, 2 lanes) with k = 3, ldx = 5, ldy = 4: each column, padding untouched
        k, ldx, ldy = 3, 5, 4
        X = rng.uniform(-1, 1, (A.n, ldx)).astype(dt)
        Y0 = rng.uniform(-1, 1, (A.m, ldy)).astype(dt)
        dX, dY = lib.DeviceArray.from_numpy(X), lib.DeviceArray.from_numpy(Y0)
        M.spmm(lib.CSR, k, -0.5, dX, ldx, 1.5, dY, ldy, variant=17)
        Y = dY.numpy()
        np.testing.assert_array_equal(Y[:, k:], Y0[:, k:])
        for j in range(k):
            ref = Y0[:, j].astype(np.float64)
            oracle.spmv_csr(A.m, A.n, A.rp, A.ci, A.v.astype(np.float64), -0.5, X[:, j].astype(np.float64), 1.5,
                            ref)
            assert rel2(Y[:, j], ref) <= TOL[np.dtype(dt)], (A.m, j)


def test_csr64_empty_rows_and_refusals(gpu, oracle):
    lib = gpu
    rng = np.random.default_rng(5)
    # empty rows at the start, in the middle and at the end
    rp = np.array([0, 0, 2, 2, 2, 5, 6, 6], dtype=np.int64)
    ci = np.array([1, 4, 0, 2, 5, 3], dtype=np.int32)
    v = rng.uniform(-1, 1, 6)
    M = lib.Matrix.csr64(7, 6, rp, ci, v)
    x = rng.uniform(-1, 1, 6)
    y0 = rng.uniform(-1, 1, 7)
    dx, dy = lib.DeviceArray.from_numpy(x), lib.DeviceArray.from_numpy(y0)
    M.spmv(lib.CSR, 0.8, dx, 0.7, dy)
    ref = y0.copy()
    oracle.spmv_csr(7, 6, rp.astype(np.int32), ci, v, 0.8, x, 0.7, ref)
    assert rel2(dy.numpy(), ref) <= 1e-12
    # byte counts: 8-byte row offsets
    assert M.effective_bytes(lib.CSR) == 6 * (8 + 4) + 8 * 8 + (6 + 7) * 8
    # every call this handle does not run is refused
    for call in (lambda: M.convert(lib.SELL_C_SIGMA, lib.make_params(chunk=64, sigma=64, alignment=1)),
                 lambda: M.convert(lib.COO, None),
                 lambda: M.prepare(lib.CSR, lib.VARIANT_PANEL),
                 lambda: M.export(lib.CSR, "row_offset", np.int32),
                 lambda: M.features(),
                 lambda: M.reference_bytes(lib.CSR)):
        with pytest.raises(lib.LespmvError) as e:
            call()
        assert e.value.code == lib.ERR_UNSUPPORTED
    # sizes beyond 32-bit rows are refused at creation
    with pytest.raises(lib.LespmvError) as e:
        lib.Matrix.csr64(1 << 31, 4, np.zeros(2, np.int64), np.zeros(0, np.int32), np.zeros(0))
    assert e.value.code in (lib.ERR_UNSUPPORTED, lib.ERR_ARG)


@pytest.mark.slow
def test_csr64_beyond_2_31_nonzeros(gpu):
    """A periodic band of 17 diagonals (offsets -8..8, wrapping round the
    columns) on 126,322,568 rows: 2,147,483,656 nonzeros, nine more than
    32-bit row offsets hold. Values are 1, 2 or 3 by diagonal and x[c] =
    c mod 7, so every product and sum is a small integer that FP32 holds
    exactly: y must equal the integer closed form y[r] = sum_d w_d x[(r + d)
    mod n] bit for bit."""
    lib = gpu
    m = n = 126_322_568
    offs = np.arange(-8, 9, dtype=np.int64)
    w = (1 + (offs + 8) % 3).astype(np.float32)
    nnz = 17 * m
    assert nnz > (1 << 31) - 1
    rp = np.arange(m + 1, dtype=np.int64) * 17
    ci = np.empty(nnz, dtype=np.int32)
    step = 1 << 23
    for a in range(0, m, step):
        b = min(m, a + step)
        rows = np.arange(a, b, dtype=np.int64)
        ci[17 * a:17 * b] = ((rows[:, None] + offs[None, :]) % n).astype(np.int32).ravel()
    v = np.tile(w, m)
    print(f"csr64: m={m} nnz={nnz} host arrays built", flush=True)
    M = lib.Matrix.csr64(m, n, rp, ci, v)
    del ci, v
    xi = np.arange(n, dtype=np.int64) % 7
    dx = lib.DeviceArray.from_numpy(xi.astype(np.float32))
    dy = lib.DeviceArray.empty(m, np.float32)
    M.spmv(lib.CSR, 1.0, dx, 0.0, dy)
    y = dy.numpy()
    print("csr64: SpMV done", flush=True)
    expect = np.zeros(m, dtype=np.int64)
    for d, wd in zip(offs, (1 + (offs + 8) % 3)):
        expect += int(wd) * np.roll(xi, -int(d))
    np.testing.assert_array_equal(y, expect.astype(np.float32))
