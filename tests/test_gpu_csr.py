"""CSR SpMV / SpMM on the GPU through the C ABI, against the CPU oracle.

Tolerances (relative 2-norm vs the FP64 oracle on the same inputs): 1e-12 for
FP64, 1e-5 for FP32; each FP32 case also records its measured error.
"""
import numpy as np
import pytest

from oracle.oracle import CSR
from tests.conftest import TOL, rel2
from tests.matrices import HostCsr, random_csr, suite

pytestmark = pytest.mark.gpu

# auto, vector (tuned), merge, panels 8 / 16 / 4, merge tiles cut at row starts, vector L = 2..64
VARIANTS = [0, 1, 2, 4, 5, 6, 7, 17, 18, 19, 20, 21, 22]


def oracle_y(oracle, A, alpha, x, beta, y0):
    y = y0.astype(np.float64).copy()
    oracle.spmv_csr(A.m, A.n, A.rp, A.ci, A.v.astype(np.float64), alpha, x.astype(np.float64), beta, y)
    return y


def gpu_y(gpu, M, fmt, alpha, x, beta, y0, variant=0):
    dx = gpu.DeviceArray.from_numpy(x)
    dy = gpu.DeviceArray.from_numpy(y0)
    M.spmv(fmt, alpha, dx, beta, dy, variant=variant)
    return dy.numpy()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_csr_suite_all_variants(gpu, oracle, dt):
    rng = np.random.default_rng(11)
    for name, A in suite(dt).items():
        M = gpu.Matrix.from_host(A)
        x = rng.uniform(-1, 1, A.n).astype(dt)
        y0 = rng.uniform(-1, 1, A.m).astype(dt)
        ref = oracle_y(oracle, A, 0.8, x, 0.7, y0)
        ref0 = oracle_y(oracle, A, 1.0, x, 0.0, y0)
        for var in VARIANTS:
            y = gpu_y(gpu, M, CSR, 0.8, x, 0.7, y0, var)
            assert rel2(y, ref) <= TOL[np.dtype(dt)], (name, var, rel2(y, ref))
            # beta = 0 must not read y: start from NaN
            y = gpu_y(gpu, M, CSR, 1.0, x, 0.0, np.full(A.m, np.nan, dt), var)
            assert np.all(np.isfinite(y)), (name, var)
            assert rel2(y, ref0) <= TOL[np.dtype(dt)], (name, var)


def test_csr_matches_reference_fixture(gpu, golden):
    """y of the compiled reference's LeSpMV_csr (flag 0, alpha .8, beta .7, y0 0)."""
    names = sorted({k.split("/")[1] for k in golden.files if k.startswith("float64/") and k.endswith("/m")})
    for dn in ("float64", "float32"):
        for name in names:
            if f"{dn}/{name}/m" not in golden.files:
                continue
            g = lambda s: golden[f"{dn}/{name}/{s}"]
            A = HostCsr(int(g("m")), int(g("n")), g("rp"), g("ci"), g("v"))
            M = gpu.Matrix.from_host(A)
            for var in (1, 2):
                y = gpu_y(gpu, M, CSR, 0.8, g("x"), 0.7, np.zeros(A.m, A.v.dtype), var)
                assert rel2(y, g("csr/0/y")) <= TOL[A.v.dtype], (dn, name, var)


def test_csr_edge_shapes(gpu, oracle):
    rng = np.random.default_rng(3)
    cases = {
        "all_empty": HostCsr(100, 50, np.zeros(101, np.int32), np.zeros(0, np.int32), np.zeros(0)),
        "one_by_one": HostCsr(1, 1, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([2.5])),
        "huge_row": random_csr(40, 120000, seed=5, mean=3.0, long_rows=2, long_len=100000),
        "many_empty": random_csr(30000, 30000, seed=6, mean=1.0, empty_frac=0.9),
    }
    for name, A in cases.items():
        M = gpu.Matrix.from_host(A)
        x = rng.uniform(-1, 1, A.n)
        y0 = rng.uniform(-1, 1, A.m)
        ref = oracle_y(oracle, A, 0.8, x, 0.7, y0)
        for var in (1, 2, 4, 5, 6, 7, 17, 19, 21, 22):
            y = gpu_y(gpu, M, CSR, 0.8, x, 0.7, y0, var)
            assert rel2(y, ref) <= 1e-12, (name, var)


@pytest.mark.parametrize("longest", [511, 512])
def test_row_tiles_longest_row(gpu, oracle, longest):
    """Merge-path tiles cut at row starts take rows of up to 511 entries (a
    tile of 2048 items holds its nominal 2048 - 512 items plus one whole
    row); one row of 512 sends the call to the merge path with its fix-up.
    Rows of the limit's length sit back to back, after runs of empty rows
    and at the matrix end, so tile cuts land inside them."""
    rng = np.random.default_rng(longest)
    m, n = 6000, 9000
    lens = rng.poisson(5.0, m)
    lens[rng.random(m) < 0.3] = 0
    lens[1000:1040] = longest
    lens[3000:3400] = 0
    lens[rng.choice(m, 25, replace=False)] = longest
    lens[-3:] = longest
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate([np.sort(rng.choice(n, int(L), replace=False)) for L in lens]).astype(np.int32)
    A = HostCsr(m, n, rp, ci, rng.uniform(-1, 1, int(rp[-1])))
    M = gpu.Matrix.from_host(A)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    ref = oracle_y(oracle, A, 0.8, x, 0.7, y0)
    for beta in (0.7, 0.0):
        want = ref if beta else oracle_y(oracle, A, 0.8, x, 0.0, y0)
        y = gpu_y(gpu, M, CSR, 0.8, x, beta, y0, gpu.VARIANT_ROW_TILES)
        assert rel2(y, want) <= 1e-12, (longest, beta)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_csr_generated_configs(gpu, oracle, dt):
    """Config-1 Laplacian (1024^2) and a 300k-row lognormal matrix of the config-2 family."""
    for fam, prm, seed in ((gpu.GEN_LAPLACE2D, [1024, 1024, 4.0, -1.0], 0),
                           (gpu.GEN_LOGNORMAL, [300000, 300000, 16.0, 1.0, 4096, 0.3, 64], 2)):
        A = gpu.generate(fam, prm, seed, dt)
        M = gpu.Matrix.from_host(A)
        x = gpu.fill_uniform(A.n, 4, dt)
        y0 = gpu.fill_uniform(A.m, 9, dt)
        ref = oracle_y(oracle, A, 0.8, x, 0.7, y0)
        for var in (0, 1, 2, 4, 5, 6, 7):
            y = gpu_y(gpu, M, CSR, 0.8, x, 0.7, y0, var)
            err = rel2(y, ref)
            print(f"{np.dtype(dt).name} fam={fam} variant={var} rel2={err:.3e}")
            assert err <= TOL[np.dtype(dt)]


def test_laplacian_closed_form(gpu):
    """5-point Laplacian times the constant vector vanishes on interior rows and
    equals the number of missing neighbours on the boundary."""
    A = gpu.generate(gpu.GEN_LAPLACE2D, [257, 129, 4.0, -1.0], 0)
    M = gpu.Matrix.from_host(A)
    for var in (1, 2):
        y = gpu_y(gpu, M, CSR, 1.0, np.ones(A.n), 0.0, np.zeros(A.m), var)
        missing = 5 - np.diff(A.rp)
        assert np.array_equal(y, missing.astype(np.float64))


@pytest.mark.parametrize("k", [1, 4, 8, 16, 32])
def test_csr_spmm_columns_equal_spmv(gpu, oracle, k):
    rng = np.random.default_rng(k)
    A = random_csr(3000, 2500, seed=k, mean=9.0, long_rows=3, long_len=900)
    M = gpu.Matrix.from_host(A)
    X = rng.uniform(-1, 1, (A.n, k))
    Y0 = rng.uniform(-1, 1, (A.m, k))
    dX, dY = gpu.DeviceArray.from_numpy(X), gpu.DeviceArray.from_numpy(Y0)
    M.spmm(CSR, k, 0.8, dX, k, 0.7, dY, k)
    Y = dY.numpy()
    for j in range(k):
        ref = oracle_y(oracle, A, 0.8, X[:, j].copy(), 0.7, Y0[:, j].copy())
        assert rel2(Y[:, j], ref) <= 1e-12, j


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_csr_spmm_row_engine(gpu, oracle, dt):
    """The row engine (SpMM variants 16 + log2 L, the serving kernel's rows):
    every column against the oracle's SpMV, for k giving one, two and four
    right-hand sides per lane, with vector access (leading dimension k) and
    with scalar access (leading dimension k + 1, whose extra column must stay
    untouched)."""
    rng = np.random.default_rng(5)
    mats = [random_csr(3000, 2500, seed=21, mean=9.0, long_rows=3, long_len=900, dtype=dt),
            suite(dt)["rand_empty"], suite(dt)["single_row"], suite(dt)["rect_tall"]]
    worst = 0.0
    for A in mats:
        M = gpu.Matrix.from_host(A)
        # 17, 20, 32, 33 and 64: the row's 8 or 16 lanes share its index and value loads (lanes past k too)
        for k in (1, 2, 3, 4, 5, 8, 17, 20, 32, 33, 64):
            for extra in (0, 1):
                ld = k + extra
                X = rng.uniform(-1, 1, (A.n, ld)).astype(dt)
                Y0 = rng.uniform(-1, 1, (A.m, ld)).astype(dt)
                refs = [oracle_y(oracle, A, 0.8, X[:, j].copy(), 0.7, Y0[:, j].copy()) for j in range(k)]
                for var in (16, 18, 22):
                    dX, dY = gpu.DeviceArray.from_numpy(X), gpu.DeviceArray.from_numpy(Y0)
                    M.spmm(CSR, k, 0.8, dX, ld, 0.7, dY, ld, variant=var)
                    Y = dY.numpy()
                    for j in range(k):
                        err = rel2(Y[:, j], refs[j])
                        worst = max(worst, err)
                        assert err <= TOL[np.dtype(dt)], (A.m, k, extra, var, j, err)
                    if extra:
                        assert np.array_equal(Y[:, k:], Y0[:, k:])
    print(f"row engine {np.dtype(dt).name}: worst rel2 {worst:.2e}")


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_csr_merge_wide_rows(gpu, oracle, dt):
    """Merge-path tiles whose rows average 16+ entries add each row with four
    threads (27-point stencil, 4x4 blocks with 8 blocks per block row) and
    shorter ones with one (5-point Laplacian): all against the oracle."""
    rng = np.random.default_rng(7)
    mats = [gpu.generate(gpu.GEN_STENCIL27, [22, 26.0, -1.0], 3, dt),
            gpu.generate(gpu.GEN_BLOCKRAND, [3000, 2900, 4, 4, 8], 4, dt),
            gpu.generate(gpu.GEN_LAPLACE2D, [97, 61, 4.0, -1.0], 5, dt)]
    for A in mats:
        M = gpu.Matrix.from_host(A)
        x = rng.uniform(-1, 1, A.n).astype(dt)
        y0 = rng.uniform(-1, 1, A.m).astype(dt)
        ref = oracle_y(oracle, A, 0.8, x, 0.7, y0)
        for var in (gpu.VARIANT_MERGE, gpu.VARIANT_ROW_TILES):
            got = gpu_y(gpu, M, CSR, 0.8, x, 0.7, y0, var)
            assert rel2(got, ref) <= TOL[np.dtype(dt)], (A.m, A.nnz, var)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_csr_panels_wide_and_compact_columns(gpu, oracle, dt):
    """Column panels on a matrix of 9M columns: in FP64, 16 panels (562,500
    columns each) keep panel-local columns in 16 bits plus a nibble, and 8
    and 4 panels (1,125,000 and 2,250,000 columns, beyond 2^20) keep them in
    32 bits; FP32 panels always keep 32 bits. All against the oracle."""
    rng = np.random.default_rng(21)
    m, n = 3000, 9_000_000
    lens = rng.poisson(12, m)
    lens[::50] = 0
    rows = [np.unique(rng.integers(0, n, int(k))) for k in lens]
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows).astype(np.int32)
    A = HostCsr(m, n, rp, ci, rng.uniform(-1, 1, len(ci)).astype(dt))
    M = gpu.Matrix.from_host(A)
    x = rng.uniform(-1, 1, n).astype(dt)
    y0 = rng.uniform(-1, 1, m).astype(dt)
    ref = oracle_y(oracle, A, 0.8, x, 0.7, y0)
    for var in (gpu.VARIANT_PANEL16, gpu.VARIANT_PANEL, gpu.VARIANT_PANEL4):
        got = gpu_y(gpu, M, CSR, 0.8, x, 0.7, y0, var)
        assert rel2(got, ref) <= TOL[np.dtype(dt)], (np.dtype(dt).name, var, rel2(got, ref))


def test_scratch_calls_across_streams(gpu, oracle):
    """A fresh handle's MERGE and PANEL calls (lazily analysed, layout built on
    first use) on a caller's stream, then on the handle's own stream, then back:
    each call's y against the oracle. The handle's scratch (partition, carries,
    panel partials) must be complete and not shared by overlapping calls."""
    import ctypes as C
    rng = np.random.default_rng(77)
    A = gpu.generate(gpu.GEN_LOGNORMAL, [150000, 150000, 16.0, 1.0, 4096, 0.3, 64], 13)
    x = rng.uniform(-1, 1, A.n)
    y0 = rng.uniform(-1, 1, A.m)
    ref = oracle_y(oracle, A, 0.8, x, 0.7, y0)
    L = gpu.load()
    s = C.c_void_p()
    gpu.check(L.lespmv_stream_create(C.byref(s)))
    try:
        for var in (gpu.VARIANT_MERGE, gpu.VARIANT_ROW_TILES, gpu.VARIANT_PANEL, gpu.VARIANT_PANEL4):
            M = gpu.Matrix.from_host(A)
            dx = gpu.DeviceArray.from_numpy(x)
            ys = [gpu.DeviceArray.from_numpy(y0) for _ in range(3)]
            M.spmv(CSR, 0.8, dx, 0.7, ys[0], variant=var, stream=s)   # caller's stream first
            M.spmv(CSR, 0.8, dx, 0.7, ys[1], variant=var)             # then the handle's own
            M.spmv(CSR, 0.8, dx, 0.7, ys[2], variant=var, stream=s)   # and back
            gpu.check(L.lespmv_stream_synchronize(s))
            gpu.synchronize()
            for i, dy in enumerate(ys):
                err = rel2(dy.numpy(), ref)
                assert err <= 1e-12, (var, i, err)
    finally:
        L.lespmv_stream_destroy(s)


def test_panels_compact_column_boundaries(gpu, oracle):
    """FP64 panels of exactly 2^20 columns (4 panels of a 4,194,304-column
    matrix) and of 2^19 (8 panels) keep panel-local columns in 16 bits plus a
    nibble; entries at each panel's first column p * 2^20 and last column
    p * 2^20 + 2^20 - 1 (local 0xFFFFF) must gather the right x."""
    n = 4194304
    rng = np.random.default_rng(9)
    rows = []
    for p in range(4):
        rows.append([p << 20, (p << 20) + (1 << 20) - 1])
        rows.append([(p << 20) + (1 << 19) - 1, (p << 20) + (1 << 19)])
    rows.append(list(range(0, n, 1 << 16)))
    rows.append([n - 1])
    rows.append([])
    m = len(rows)
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate([np.array(r, np.int32) for r in rows if r]).astype(np.int32)
    A = HostCsr(m, n, rp, ci, rng.uniform(-1, 1, len(ci)))
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    ref = oracle_y(oracle, A, 0.8, x, 0.7, y0)
    M = gpu.Matrix.from_host(A)
    for var in (gpu.VARIANT_PANEL4, gpu.VARIANT_PANEL, gpu.VARIANT_PANEL16):
        M.prepare(CSR, var)
        got = gpu_y(gpu, M, CSR, 0.8, x, 0.7, y0, var)
        assert rel2(got, ref) <= 1e-12, (var, rel2(got, ref))


def test_auto_picks_panels_on_scattered_columns(gpu, oracle):
    """AUTO takes the column panels on a matrix whose x spans 32 MB (FP64) with
    most columns far off the diagonal, and the vector / merge kernels on banded
    ones; lespmv_prepare builds the layout ahead (its time and bytes reported)
    and lespmv_release_format(CSR) frees it again."""
    A = gpu.generate(gpu.GEN_LOGNORMAL, [4194304, 4194304, 2.0, 1.0, 64, 0.3, 64], 17)
    M = gpu.Matrix.from_host(A)
    base = M.format_bytes(CSR)
    ms = M.prepare(CSR, gpu.VARIANT_AUTO)
    grown = M.format_bytes(CSR)
    print(f"prepare (AUTO) {ms:.2f} ms, CSR bytes {base} -> {grown}")
    assert ms > 0 and grown > base + A.nnz * 8  # the panel layout exists
    assert M.csr_plan()[0] == gpu.VARIANT_PANEL
    x = gpu.fill_uniform(A.n, 4)
    y0 = gpu.fill_uniform(A.m, 9)
    ref = oracle_y(oracle, A, 0.8, x, 0.7, y0)
    assert rel2(gpu_y(gpu, M, CSR, 0.8, x, 0.7, y0, gpu.VARIANT_AUTO), ref) <= 1e-12
    M.release(CSR)
    assert M.format_bytes(CSR) < grown
    assert rel2(gpu_y(gpu, M, CSR, 0.8, x, 0.7, y0, gpu.VARIANT_AUTO), ref) <= 1e-12  # rebuilt on first use
    B = gpu.generate(gpu.GEN_LAPLACE2D, [2048, 2048, 4.0, -1.0], 0)
    MB = gpu.Matrix.from_host(B)
    MB.prepare(CSR, gpu.VARIANT_AUTO)
    assert MB.format_bytes(CSR) < B.nnz * 12 + (B.m + 1) * 4 + 64 * 1024 * 1024  # no panels for a banded matrix
    assert MB.csr_plan()[0] != gpu.VARIANT_PANEL


def test_auto_plan(gpu):
    """The CSR variant AUTO runs (lespmv_csr_plan): column panels when x spans
    16 MB or more with 0.3 far x lines per entry or more (random 2x2 blocks:
    half a line per entry), not for random 4x4 blocks (a quarter line per
    entry), which at 285 MB and 4 entries a row take the merge-path tiles (cut
    at row starts, as every row fits a quarter tile), as does a 27-point
    stencil beyond the 256 MB last-level cache; the vector kernel for a
    stencil inside it."""
    cases = [
        (gpu.GEN_LOGNORMAL, [4194304, 4194304, 2.0, 1.0, 64, 0.3, 64], gpu.VARIANT_PANEL, (0.3, 1.0)),
        (gpu.GEN_BLOCKRAND, [2097152, 2097152, 2, 2, 4], gpu.VARIANT_PANEL, (0.4, 0.55)),
        (gpu.GEN_BLOCKRAND, [1048576, 1048576, 4, 4, 1], gpu.VARIANT_ROW_TILES, (0.2, 0.3)),  # 285 MB, 4 per row
        (gpu.GEN_STENCIL27, [110, 26.0, -1.0], gpu.VARIANT_ROW_TILES, (0.0, 0.01)),
        (gpu.GEN_STENCIL27, [40, 26.0, -1.0], gpu.VARIANT_VECTOR, (0.0, 0.01)),
    ]
    for gen, params, want, (lo, hi) in cases:
        A = gpu.generate(gen, params, 5)
        M = gpu.Matrix.from_host(A)
        variant, lanes, far = M.csr_plan()
        print(f"generator {gen} {params}: variant {variant}, lanes {lanes}, far lines per entry {far:.3f}")
        assert variant == want, (gen, params, variant)
        assert lo <= far < hi, (gen, params, far)
        assert 2 <= lanes <= 64
        del M
