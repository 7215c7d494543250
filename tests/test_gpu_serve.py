"""Batched serving on the GPU (lespmv_server_*): the packed persistent kernel
replayed from a HIP graph against the per-call kernel of the format each
request runs on (CSR row engine, sliced layouts, BSR 4x4 on the matrix cores)
and the oracle, results identical alone and in a mix, repeated graph replays
with other work in between, and admission / eviction by HBM footprint."""
import numpy as np
import pytest

from tests.conftest import TOL, rel2

pytestmark = pytest.mark.gpu


def matrices(lib):
    return [lib.generate(lib.GEN_LAPLACE2D, [60, 50, 4.0, -1.0], 0),
            lib.generate(lib.GEN_LOGNORMAL, [3000, 2800, 16.0, 1.0, 4096, 0.3, 64], 2),
            lib.generate(lib.GEN_BLOCKRAND, [400, 380, 4, 4, 8], 3),
            lib.generate(lib.GEN_POWERLAW, [5000, 2.0, 8.0, 2000], 4),
            lib.generate(lib.GEN_STENCIL27, [14, 26.0, -1.0], 5)]


def dense_inputs(lib, A, k, seed):
    X = lib.fill_uniform(A.n * k, seed).reshape(A.n, k)
    Y0 = lib.fill_uniform(A.m * k, seed + 1).reshape(A.m, k)
    return X, Y0


def run(lib, S, mids, As, reqs):
    """Submit reqs = [(matrix index, k, alpha, beta, seed)] together; returns Y arrays."""
    bufs = []
    for i, k, alpha, beta, seed in reqs:
        X, Y0 = dense_inputs(lib, As[i], k, seed)
        dX, dY = lib.DeviceArray.from_numpy(X), lib.DeviceArray.from_numpy(Y0)
        S.submit(mids[i], k, alpha, dX, k, beta, dY, k)
        bufs.append((dX, dY))
    S.drain()
    S.poll()
    return [dY.numpy() for _, dY in bufs]


def percall(lib, S, mid, A, k, alpha, dX, ldx, beta, Y0, ldy):
    """The request's product from the per-call kernel of the format the server
    runs it on (which gives bitwise the server's result), or None where that
    kernel cannot take the request's layout (a sliced k = 1 request whose x or y
    is strided)."""
    fmt = S.format(mid, k)
    dY = lib.DeviceArray.from_numpy(Y0)
    if fmt == lib.CSR:
        M = lib.Matrix.from_host(A)
        M.spmm(lib.CSR, k, alpha, dX, ldx, beta, dY, ldy, variant=16 + S.lanes(mid, k).bit_length() - 1)
    elif fmt == lib.BSR:
        S.matrix(mid, A).spmm(lib.BSR, k, alpha, dX, ldx, beta, dY, ldy)
    else:
        assert fmt in (lib.SELL, lib.SELL_C_SIGMA) and k == 1
        if ldx != 1 or ldy != 1:
            return None
        S.matrix(mid, A).spmv(fmt, alpha, dX, beta, dY)
    return dY.numpy()


REQS = [(0, 1, 1.0, 0.0, 10), (1, 3, 0.8, 0.7, 11), (2, 8, 1.0, 0.0, 12), (3, 32, -0.5, 2.0, 13),
        (4, 5, 1.0, 0.0, 14), (1, 1, 1.0, 0.0, 15), (3, 17, 1.0, 0.0, 16), (0, 32, 0.3, 0.0, 17)]


def test_packed_kernel_matches_percall_and_oracle(gpu, oracle):
    lib = gpu
    As = matrices(lib)
    S = lib.Server(np.float64, 1 << 30, max_batch=64, slots=2)
    mids = [S.add(A) for A in As]
    assert all(S.lanes(m) in (2, 4, 8, 16, 32, 64) for m in mids)
    allowed = {1: (lib.CSR, lib.SELL, lib.SELL_C_SIGMA)}
    for m in mids:
        for k in (1, 4, 5, 9, 16, 17, 32):
            assert S.format(m, k) in allowed.get(k, (lib.CSR, lib.BSR)), (m, k)
    for (i, k, alpha, beta, seed), Y in zip(REQS, run(lib, S, mids, As, REQS)):
        A = As[i]
        X, Y0 = dense_inputs(lib, A, k, seed)
        M = lib.Matrix.from_host(A)
        dX, dY = lib.DeviceArray.from_numpy(X), lib.DeviceArray.from_numpy(Y0)
        M.spmm(lib.CSR, k, alpha, dX, k, beta, dY, k)
        assert rel2(Y, dY.numpy()) <= 1e-12, (i, k)
        # the per-call kernel of the request's format gives bitwise the server's result
        np.testing.assert_array_equal(Y, percall(lib, S, mids[i], A, k, alpha, dX, k, beta, Y0, k))
        # column 0 against the oracle's CSR SpMV
        ref = oracle.spmv_csr(A.m, A.n, A.rp, A.ci, A.v, alpha, np.ascontiguousarray(X[:, 0]), beta,
                              np.ascontiguousarray(Y0[:, 0]))
        assert rel2(Y[:, 0], ref) <= 1e-12, (i, k)


def test_same_result_alone_and_in_mix(gpu):
    """Each request alone in a batch and inside the mix gives bitwise the same Y."""
    lib = gpu
    As = matrices(lib)
    S = lib.Server(np.float64, 1 << 30, max_batch=64, slots=3)
    mids = [S.add(A) for A in As]
    alone = [run(lib, S, mids, As, [r])[0] for r in REQS]
    mixed = run(lib, S, mids, As, REQS)
    for a, b in zip(alone, mixed):
        np.testing.assert_array_equal(a, b)
    # a small batch limit splits the mix over several graph launches; lanes
    # given explicitly apply to every k
    S2 = lib.Server(np.float64, 1 << 30, max_batch=3, slots=2)
    mids2 = [S2.add(A, lanes=S.lanes(m)) for A, m in zip(As, mids)]
    assert all(S2.lanes(m2, k) == S.lanes(m) for m, m2 in zip(mids, mids2) for k in (1, 2, 16))
    alone2 = [run(lib, S2, mids2, As, [r])[0] for r in REQS]
    for a, b in zip(alone2, run(lib, S2, mids2, As, REQS)):
        np.testing.assert_array_equal(a, b)
    # k = 1 requests use the same lanes on both servers (where the first runs them on CSR too)
    for (i, k, *_), a, b in zip(REQS, alone, alone2):
        if k == 1 and S.format(mids[i], 1) == lib.CSR:
            np.testing.assert_array_equal(a, b)
    assert S2.stats()["batches"] >= 3


def test_graph_replays_with_eager_work_between(gpu):
    """The same requests replayed five times through the captured graphs, with
    eager library calls and device synchronisations in between: every replay
    equals the first."""
    lib = gpu
    As = matrices(lib)
    S = lib.Server(np.float64, 1 << 30, max_batch=64, slots=2)
    mids = [S.add(A) for A in As]
    first = run(lib, S, mids, As, REQS)
    M = lib.Matrix.from_host(As[1])
    x = lib.DeviceArray.from_numpy(lib.fill_uniform(As[1].n, 3))
    y = lib.DeviceArray.empty(As[1].m, np.float64)
    for rep in range(5):
        M.spmv(lib.CSR, 1.0, x, 0.0, y)
        lib.synchronize()
        again = run(lib, S, mids, As, REQS)
        for a, b in zip(first, again):
            np.testing.assert_array_equal(a, b)


def test_admission_by_footprint(gpu):
    lib = gpu
    As = matrices(lib)
    sizes = [(A.m + 1) * 4 + A.nnz * 12 for A in As]
    # room for matrices 0 and 1 (plus their chunk tables), not for 2 as well
    budget = sizes[0] + sizes[1] + sizes[2] // 2 + 65536
    S = lib.Server(np.float64, budget, max_batch=16, slots=2)
    m0, m1 = S.add(As[0], lanes=4), S.add(As[1], lanes=8)
    # a request on matrix 0 in flight: admitting matrix 2 must evict matrix 1
    X, _ = dense_inputs(lib, As[0], 1, 1)
    dX, dY = lib.DeviceArray.from_numpy(X), lib.DeviceArray.empty((As[0].m, 1), np.float64)
    S.submit(m0, 1, 1.0, dX, 1, 0.0, dY, 1)
    m2 = S.add(As[2], lanes=4)
    st = S.stats()
    assert st["resident"] == 2 and st["bytes_used"] <= budget
    with pytest.raises(lib.LespmvError):
        S.submit(m1, 1, 1.0, dX, 1, 0.0, dY, 1)  # evicted
    S.drain()
    ids, lat = S.poll()
    assert len(ids) == 1 and lat[0] > 0
    assert m2 == 2  # ids are not reused
    # larger than the whole budget
    big = lib.generate(lib.GEN_LOGNORMAL, [20000, 20000, 16.0, 1.0, 4096, 0.3, 64], 9)
    with pytest.raises(lib.LespmvError) as e:
        S.add(big)
    assert e.value.code == lib.ERR_OOM


def test_fp32_server_and_padded_leading_dimensions(gpu, oracle):
    """FP32 and FP64 servers with requests whose leading dimensions exceed k
    (scalar X / Y access instead of vectors): each result equals the row engine
    launched on its own bitwise, leaves the padding columns untouched and is
    within the dtype's tolerance of the FP64 oracle, column by column."""
    lib = gpu
    rng = np.random.default_rng(17)
    cases = [(0, 3, 5, 4), (1, 8, 11, 9), (0, 17, 17, 20), (1, 32, 33, 32), (0, 1, 2, 3), (1, 4, 4, 4)]
    for dt in (np.float32, np.float64):
        As = [lib.generate(lib.GEN_LOGNORMAL, [3000, 2800, 16.0, 1.0, 4096, 0.3, 64], 2, dt),
              lib.generate(lib.GEN_STENCIL27, [14, 26.0, -1.0], 5, dt)]
        S = lib.Server(dt, 1 << 30, max_batch=16, slots=2)
        mids = [S.add(A) for A in As]
        bufs = []
        for i, k, ldx, ldy in cases:
            A = As[i]
            X = rng.uniform(-1, 1, (A.n, ldx)).astype(dt)
            Y0 = rng.uniform(-1, 1, (A.m, ldy)).astype(dt)
            dX, dY = lib.DeviceArray.from_numpy(X), lib.DeviceArray.from_numpy(Y0)
            S.submit(mids[i], k, 0.9, dX, ldx, 0.5, dY, ldy)
            bufs.append((X, Y0, dX, dY))
        S.drain()
        S.poll()
        for (i, k, ldx, ldy), (X, Y0, dX, dY) in zip(cases, bufs):
            A = As[i]
            Y = dY.numpy()
            same = percall(lib, S, mids[i], A, k, 0.9, dX, ldx, 0.5, Y0, ldy)
            if same is not None:
                np.testing.assert_array_equal(Y, same)
            np.testing.assert_array_equal(Y[:, k:], Y0[:, k:])
            for j in range(k):
                ref = Y0[:, j].astype(np.float64)
                oracle.spmv_csr(A.m, A.n, A.rp, A.ci, A.v.astype(np.float64), 0.9, X[:, j].astype(np.float64), 0.5,
                                ref)
                assert rel2(Y[:, j], ref) <= TOL[np.dtype(dt)], (np.dtype(dt).name, i, k, j)


def test_sliced_and_bsr_requests(gpu, oracle):
    """Resident matrices set to sliced layouts (k = 1; S-ELL and SELL-C-sigma,
    C 32 and 64) and to 4x4 BSR on the matrix cores (k >= 5) by
    lespmv_server_use: every request equals the per-call kernel of its format
    bitwise, is the same alone and in the mix, and is within 1e-12 (FP64) or
    1e-5 (FP32) of the FP64 oracle in every column, strided x included; the
    layouts' bytes count against the budget, and the refusals give their codes."""
    lib = gpu
    rng = np.random.default_rng(5)
    for dt in (np.float64, np.float32):
        As = [lib.generate(lib.GEN_LAPLACE2D, [60, 50, 4.0, -1.0], 0, dt),
              lib.generate(lib.GEN_STENCIL27, [14, 26.0, -1.0], 5, dt),
              lib.generate(lib.GEN_BLOCKRAND, [400, 380, 4, 4, 8], 3, dt),
              lib.generate(lib.GEN_POWERLAW, [5000, 2.0, 8.0, 2000], 4, dt)]
        S = lib.Server(dt, 1 << 30, max_batch=64, slots=2)
        mids = [S.add(A, lanes=4) for A in As]
        csr_bytes = S.stats()["bytes_used"]
        assert all(S.format(m, k) == lib.CSR for m in mids for k in (1, 8, 32))
        S.use(mids[0], lib.SELL_C_SIGMA, lib.make_params(chunk=64, sigma=256, alignment=1))
        S.use(mids[1], lib.SELL, lib.make_params(chunk=32, alignment=1))
        S.use(mids[2], lib.SELL_C_SIGMA, lib.make_params(chunk=32, sigma=4096, alignment=1))
        with pytest.raises(lib.LespmvError) as e:  # rows of up to 2000 nonzeros
            S.use(mids[3], lib.SELL, lib.make_params(chunk=64, alignment=1))
        assert e.value.code == lib.ERR_UNSUPPORTED and S.format(mids[3], 1) == lib.CSR
        with pytest.raises(lib.LespmvError) as e:
            S.use(mids[0], lib.SELL, lib.make_params(chunk=16, alignment=1))
        assert e.value.code == lib.ERR_ARG
        if dt == np.float64:
            S.use(mids[2], lib.BSR, None, 5)
            with pytest.raises(lib.LespmvError) as e:
                S.use(mids[1], lib.BSR, None, 4)
            assert e.value.code == lib.ERR_ARG
        else:
            with pytest.raises(lib.LespmvError) as e:
                S.use(mids[2], lib.BSR, None, 5)
            assert e.value.code == lib.ERR_UNSUPPORTED
        assert S.stats()["bytes_used"] > csr_bytes
        assert S.format(mids[0], 1) == lib.SELL_C_SIGMA and S.format(mids[1], 1) == lib.SELL
        assert S.format(mids[0], 2) == lib.CSR
        bsr = lib.BSR if dt == np.float64 else lib.CSR
        assert S.format(mids[2], 4) == lib.CSR and S.format(mids[2], 5) == bsr and S.format(mids[2], 32) == bsr
        # (matrix, k, ldx, ldy, alpha, beta)
        reqs = [(0, 1, 1, 1, 1.0, 0.0), (1, 1, 1, 1, 0.8, 0.7), (2, 1, 1, 1, -1.0, 0.5), (3, 1, 1, 1, 1.0, 0.0),
                (0, 1, 3, 2, 0.5, 1.5), (1, 1, 4, 1, 1.0, 0.0), (2, 8, 8, 8, 1.0, 0.0), (2, 16, 16, 17, 0.7, -0.3),
                (2, 32, 32, 32, 1.0, 0.0), (2, 5, 7, 5, 2.0, 0.0), (2, 21, 21, 24, 1.0, 1.0), (2, 4, 4, 4, 1.0, 0.0),
                (1, 9, 9, 9, 1.0, 0.0), (0, 3, 3, 3, 1.0, 0.0)]
        data = []
        for i, k, ldx, ldy, a, b in reqs:
            A = As[i]
            X = rng.uniform(-1, 1, (A.n, ldx)).astype(dt)
            Y0 = rng.uniform(-1, 1, (A.m, ldy)).astype(dt)
            data.append((X, Y0))

        def serve(which):
            out = []
            for r in which:
                i, k, ldx, ldy, a, b = reqs[r]
                X, Y0 = data[r]
                dX, dY = lib.DeviceArray.from_numpy(X), lib.DeviceArray.from_numpy(Y0)
                S.submit(mids[i], k, a, dX, ldx, b, dY, ldy)
                out.append((dX, dY))
            return out

        mixed = serve(range(len(reqs)))
        S.drain()
        S.poll()
        tol = TOL[np.dtype(dt)]
        for r, (dX, dY) in enumerate(mixed):
            i, k, ldx, ldy, a, b = reqs[r]
            A = As[i]
            X, Y0 = data[r]
            Y = dY.numpy()
            alone = serve([r])
            S.drain()
            S.poll()
            np.testing.assert_array_equal(Y, alone[0][1].numpy(), err_msg=f"request {r}")
            same = percall(lib, S, mids[i], A, k, a, dX, ldx, b, Y0, ldy)
            if same is not None:
                np.testing.assert_array_equal(Y, same, err_msg=f"request {r}")
            np.testing.assert_array_equal(Y[:, k:], Y0[:, k:])
            for j in range(k):
                ref = Y0[:, j].astype(np.float64)
                oracle.spmv_csr(A.m, A.n, A.rp, A.ci, A.v.astype(np.float64), a, X[:, j].astype(np.float64), b, ref)
                assert rel2(Y[:, j], ref) <= tol, (r, j)
        # back to the row engine only: the layouts leave the budget
        for m in mids:
            S.use(m, lib.CSR)
        assert S.stats()["bytes_used"] == csr_bytes
        assert all(S.format(m, k) == lib.CSR for m in mids for k in (1, 8, 32))
