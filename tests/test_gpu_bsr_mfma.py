"""BSR 4x4 SpMM on the matrix cores (v_mfma_f64_4x4x4f64) against the scalar
BSR kernel and CSR SpMM, including ragged edges (rows and columns not a
multiple of 4), k not a multiple of 4, and k beyond 32 (up to four groups of
16 right-hand sides per block)."""
import numpy as np
import pytest

from tests.conftest import rel2
from tests.matrices import random_csr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [2, 3, 4, 5, 16, 17, 31, 32, 33, 48, 63, 64])
def test_bsr44_mfma_spmm(gpu, k):
    lib = gpu
    cases = [lib.generate(lib.GEN_BLOCKRAND, [300, 280, 4, 4, 8], 7),
             random_csr(1001, 998, seed=4, mean=9.0, long_rows=3, long_len=600)]
    for A in cases:
        M = lib.Matrix.from_host(A)
        M.convert(lib.BSR, lib.make_params(block_rows=4, block_cols=4))
        X = lib.fill_uniform(A.n * k, 3).reshape(A.n, k)
        Y0 = lib.fill_uniform(A.m * k, 4).reshape(A.m, k)
        out = {}
        for fmt, var in ((lib.BSR, lib.VARIANT_AUTO), (lib.BSR, lib.VARIANT_VECTOR), (lib.CSR, lib.VARIANT_AUTO)):
            dX, dY = lib.DeviceArray.from_numpy(X), lib.DeviceArray.from_numpy(Y0)
            M.spmm(fmt, k, 0.9, dX, k, -0.4, dY, k, variant=var)
            out[(fmt, var)] = dY.numpy()
        ref = out[(lib.CSR, lib.VARIANT_AUTO)]
        assert rel2(out[(lib.BSR, lib.VARIANT_AUTO)], ref) <= 1e-12
        assert rel2(out[(lib.BSR, lib.VARIANT_VECTOR)], ref) <= 1e-12
        dense = np.zeros((A.m, A.n))
        rows = np.repeat(np.arange(A.m), np.diff(A.rp))
        np.add.at(dense, (rows, A.ci), A.v)
        assert rel2(out[(lib.BSR, lib.VARIANT_AUTO)], 0.9 * dense @ X - 0.4 * Y0) <= 1e-12
