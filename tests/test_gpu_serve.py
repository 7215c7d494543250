"""Batched serving on the GPU (lespmv_server_*): the packed persistent kernel
replayed from a HIP graph against per-call CSR SpMM and the oracle, results
identical alone and in a mix, repeated graph replays with other work in
between, and admission / eviction by HBM footprint."""
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


REQS = [(0, 1, 1.0, 0.0, 10), (1, 3, 0.8, 0.7, 11), (2, 8, 1.0, 0.0, 12), (3, 32, -0.5, 2.0, 13),
        (4, 5, 1.0, 0.0, 14), (1, 1, 1.0, 0.0, 15), (3, 17, 1.0, 0.0, 16), (0, 32, 0.3, 0.0, 17)]


def test_packed_kernel_matches_percall_and_oracle(gpu, oracle):
    lib = gpu
    As = matrices(lib)
    S = lib.Server(np.float64, 1 << 30, max_batch=64, slots=2)
    mids = [S.add(A) for A in As]
    assert all(S.lanes(m) in (2, 4, 8, 16, 32, 64) for m in mids)
    for (i, k, alpha, beta, seed), Y in zip(REQS, run(lib, S, mids, As, REQS)):
        A = As[i]
        X, Y0 = dense_inputs(lib, A, k, seed)
        M = lib.Matrix.from_host(A)
        dX, dY = lib.DeviceArray.from_numpy(X), lib.DeviceArray.from_numpy(Y0)
        M.spmm(lib.CSR, k, alpha, dX, k, beta, dY, k)
        assert rel2(Y, dY.numpy()) <= 1e-12, (i, k)
        # the row engine launched on its own (same lanes) gives bitwise the server's result
        lanes = S.lanes(mids[i], k)
        dY2 = lib.DeviceArray.from_numpy(Y0)
        M.spmm(lib.CSR, k, alpha, dX, k, beta, dY2, k, variant=16 + lanes.bit_length() - 1)
        np.testing.assert_array_equal(Y, dY2.numpy())
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
    # k = 1 requests use the same lanes on both servers
    for (i, k, *_), a, b in zip(REQS, alone, alone2):
        if k == 1:
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
            M = lib.Matrix.from_host(A)
            dY2 = lib.DeviceArray.from_numpy(Y0)
            M.spmm(lib.CSR, k, 0.9, dX, ldx, 0.5, dY2, ldy, variant=16 + S.lanes(mids[i], k).bit_length() - 1)
            np.testing.assert_array_equal(Y, dY2.numpy())
            np.testing.assert_array_equal(Y[:, k:], Y0[:, k:])
            for j in range(k):
                ref = Y0[:, j].astype(np.float64)
                oracle.spmv_csr(A.m, A.n, A.rp, A.ci, A.v.astype(np.float64), 0.9, X[:, j].astype(np.float64), 0.5,
                                ref)
                assert rel2(Y[:, j], ref) <= TOL[np.dtype(dt)], (np.dtype(dt).name, i, k, j)
