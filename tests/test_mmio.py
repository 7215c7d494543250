"""MatrixMarket input and coo_to_csr against the compiled reference's output."""
import os

import numpy as np
import pytest

from sspmv_amd import lib

MTX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mtx")
NAMES = ["general_real.mtx", "symmetric_real.mtx", "pattern_general.mtx", "integer_symmetric.mtx"]


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("compact", [0, 1])
def test_read_mtx_matches_reference(golden, name, compact):
    A = lib.read_mtx(os.path.join(MTX, name), np.float64, bool(compact))
    g = lambda a: golden[f"mtx/{name}/{compact}/{a}"]
    assert (A.m, A.n, A.nnz) == (int(g("num_rows")), int(g("num_cols")), int(g("num_nnzs")))
    assert np.array_equal(A.rp, g("row_offset"))
    assert np.array_equal(A.ci, g("col_index"))
    assert np.array_equal(A.v, g("values"))


def test_read_mtx_float32_rounds_values(golden):
    A = lib.read_mtx(os.path.join(MTX, "general_real.mtx"), np.float32)
    assert A.v.dtype == np.float32
    assert np.array_equal(A.v, golden["mtx/general_real.mtx/0/values"].astype(np.float32))


def test_read_mtx_errors_are_codes(tmp_path):
    with pytest.raises(lib.LespmvError) as e:
        lib.read_mtx(str(tmp_path / "missing.mtx"))
    assert e.value.code == lib.ERR_ARG
    bad = tmp_path / "dense.mtx"
    bad.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(lib.LespmvError) as e:
        lib.read_mtx(str(bad))
    assert e.value.code == lib.ERR_UNSUPPORTED


def test_coo_to_csr_compact_matches_reference(golden):
    m, n = int(golden["coo2csr/m"]), int(golden["coo2csr/n"])
    A = lib.coo_to_csr(m, n, golden["coo2csr/ri"], golden["coo2csr/ci"], golden["coo2csr/v"], compact=True)
    assert np.array_equal(A.rp, golden["coo2csr/1/row_offset"])
    assert np.array_equal(A.ci, golden["coo2csr/1/col_index"])
    assert np.array_equal(A.v, golden["coo2csr/1/values"])


@pytest.mark.gpu
def test_coo_to_csr_gpu_matches_reference(gpu, golden):
    m, n = int(golden["coo2csr/m"]), int(golden["coo2csr/n"])
    A = gpu.coo_to_csr(m, n, golden["coo2csr/ri"], golden["coo2csr/ci"], golden["coo2csr/v"], compact=False)
    assert np.array_equal(A.rp, golden["coo2csr/0/row_offset"])
    assert np.array_equal(A.ci, golden["coo2csr/0/col_index"])
    assert np.array_equal(A.v, golden["coo2csr/0/values"])


@pytest.mark.gpu
def test_coo_to_csr_gpu_large_stable(gpu, oracle):
    r = np.random.default_rng(5)
    m, n, nnz = 50000, 40000, 600000
    ri = r.integers(0, m, nnz).astype(np.int32)
    ci = r.integers(0, n, nnz).astype(np.int32)
    v = r.uniform(-1, 1, nnz)
    A = gpu.coo_to_csr(m, n, ri, ci, v)
    res = oracle.coo_to_csr(m, n, ri, ci, v, False)
    assert np.array_equal(A.rp, res.get("row_offset", np.int32))
    assert np.array_equal(A.ci, res.get("col_index", np.int32))
    assert np.array_equal(A.v, res.get("values", np.float64))
