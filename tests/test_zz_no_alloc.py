"""The allocation-free CSR SpMV (lespmv_csr_spmv_noalloc) through the C ABI.

It runs the merge-path kernels of CSR variant MERGE with their partition and
carries in a fixed device array, on the caller's own device CSR arrays. Two
calls in a row (the second on a smaller matrix, so the scratch holds the first
call's larger contents) must each equal, bitwise, lespmv_spmv(CSR, MERGE) on a
handle made from the same arrays, and lie within 1e-12 (FP64) / 1e-5 (FP32)
relative 2-norm of the FP64 oracle.
"""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import CSR
from tests.conftest import TOL, rel2
from tests.matrices import random_csr

pytestmark = pytest.mark.gpu


def noalloc_spmv(gpu, A, alpha, x, beta, y0, stream=None):
    """y from the allocation-free entry point; the CSR arrays live in the caller's device buffers."""
    drp = gpu.DeviceArray.from_numpy(A.rp.astype(np.int32))
    dci = gpu.DeviceArray.from_numpy(A.ci.astype(np.int32))
    dv = gpu.DeviceArray.from_numpy(A.v)
    dx = gpu.DeviceArray.from_numpy(x)
    dy = gpu.DeviceArray.from_numpy(y0)
    rc = gpu.load().lespmv_csr_spmv_noalloc(gpu.dtype_id(A.v.dtype), A.m, A.n, A.nnz, drp.ptr, dci.ptr, dv.ptr,
                                            alpha, dx.ptr, beta, dy.ptr, stream)
    gpu.check(rc)
    gpu.synchronize()
    return dy.numpy()


def handle_spmv(gpu, A, alpha, x, beta, y0):
    M = gpu.Matrix.from_host(A)
    dx = gpu.DeviceArray.from_numpy(x)
    dy = gpu.DeviceArray.from_numpy(y0)
    M.spmv(CSR, alpha, dx, beta, dy, variant=gpu.VARIANT_MERGE)
    return dy.numpy()


def oracle_y(oracle, A, alpha, x, beta, y0):
    y = y0.astype(np.float64).copy()
    oracle.spmv_csr(A.m, A.n, A.rp, A.ci, A.v.astype(np.float64), alpha, x.astype(np.float64), beta, y)
    return y


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_noalloc_two_calls_match_merge_entry(gpu, oracle, dt):
    rng = np.random.default_rng(41)
    big = gpu.generate(gpu.GEN_LOGNORMAL, [400000, 400000, 16.0, 1.0, 4096, 0.3, 64], 7, dt)
    small = random_csr(9000, 7000, seed=42, mean=11.0, long_rows=4, long_len=3000, dtype=dt)
    for name, A in (("lognormal400k", big), ("random9k", small)):
        x = rng.uniform(-1, 1, A.n).astype(dt)
        y0 = rng.uniform(-1, 1, A.m).astype(dt)
        print(f"noalloc {np.dtype(dt).name} {name}: m={A.m} nnz={A.nnz}", flush=True)
        for alpha, beta, start in ((0.8, 0.7, y0), (1.0, 0.0, np.full(A.m, np.nan, dt))):
            got = noalloc_spmv(gpu, A, alpha, x, beta, start)
            want = handle_spmv(gpu, A, alpha, x, beta, start)
            assert np.array_equal(got, want), (name, alpha, beta)
            ref = oracle_y(oracle, A, alpha, x, beta, np.nan_to_num(start))
            err = rel2(got, ref)
            print(f"  alpha={alpha} beta={beta}: bitwise equal to the MERGE entry, rel2 vs oracle {err:.2e}")
            assert err <= TOL[np.dtype(dt)], (name, alpha, beta, err)


def test_noalloc_limits_and_arguments(gpu):
    L = gpu.load()
    limit = L.lespmv_csr_noalloc_max_items()
    assert limit == 131072 * 2048
    # beyond the fixed scratch: refused before any device access
    rc = L.lespmv_csr_spmv_noalloc(gpu.F64, limit, limit, 1, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), 1.0,
                                   C.c_void_p(16), 0.0, C.c_void_p(16), None)
    assert rc == gpu.ERR_UNSUPPORTED
    assert b"scratch limit" in L.lespmv_last_error()
    rc = L.lespmv_csr_spmv_noalloc(gpu.F64, 10, 10, 5, None, None, None, 1.0, None, 0.0, None, None)
    assert rc == gpu.ERR_ARG
    # an empty matrix (m = 0) is a no-op; an all-empty one writes beta * y
    A = random_csr(500, 400, seed=3, mean=0.0, empty_frac=1.0)
    y = noalloc_spmv(gpu, A, 2.0, np.ones(A.n), 0.5, np.full(A.m, 4.0))
    assert np.array_equal(y, np.full(A.m, 2.0))
