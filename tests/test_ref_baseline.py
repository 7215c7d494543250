"""The CPU baseline library (the reference's own LeSpMV_csr, kernel_flag 2,
built by oracle/build_ref.sh) against the oracle's restatement of the same
kernel, so the baseline bench.py times computes the same y. Skipped where the
reference was not built."""
import numpy as np
import pytest

import bench
from tests.matrices import random_csr


def test_reference_baseline_matches_oracle(oracle):
    ref, level = bench.reference_csr_bench()
    if ref is None:
        pytest.skip(f"oracle/_ref/libref_csr_bench_{level}.so not built")
    A = random_csr(3000, 2600, seed=31, mean=12.0, long_rows=4, long_len=500)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, A.n)
    y = np.zeros(A.m)
    rp, ci = A.rp.astype(np.int32), A.ci.astype(np.int32)
    v = A.v.astype(np.float64)
    ms = ref.refbench_csr_lb(A.m, A.n, A.nnz, rp, ci, v, x, y, 4, 3)
    assert ms > 0
    want = np.zeros(A.m)
    oracle.spmv_csr(A.m, A.n, A.rp, A.ci, v, 1.0, x, 0.0, want, flag=2, threads=4)
    # the reference sums rows with `omp simd reduction`, which may reassociate
    assert np.linalg.norm(y - want) <= 1e-13 * np.linalg.norm(want)
