"""The C++ drop-in (include/lespmv/LeSpMV.h) through its benchmark driver:
every format's test routine converts on the GPU, compares against CSR with the
reference harness (alpha 0.8, beta 0.7, y0 0; maximum_relative_error) and
reports a device-timed benchmark line."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.matrices import laplace2d, random_csr

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "sspmv_amd", "bin", "benchmark_spmv")
FORMATS = ["csr", "coo", "ell", "sell", "sell_c_sigma", "sell_c_R", "dia", "bsr", "csr5"]


def write_mtx(path, A):
    rows = np.repeat(np.arange(A.m), np.diff(A.rp))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{A.m} {A.n} {A.nnz}\n")
        for r, c, v in zip(rows.tolist(), A.ci.tolist(), A.v.tolist()):
            f.write(f"{r + 1} {c + 1} {v!r}\n")


def run_driver(path, prec):
    out = subprocess.run([DRIVER, path, "--format=all", f"--precision={prec}", "--matID=7"], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return out.stdout


@pytest.mark.parametrize("prec", [64, 32])
def test_driver_all_formats(tmp_path, prec):
    cases = {"lap": laplace2d(40, 30), "rand": random_csr(500, 500, seed=3, long_rows=2, long_len=200)}
    for name, A in cases.items():
        path = str(tmp_path / f"{name}.mtx")
        write_mtx(path, A)
        text = run_driver(path, prec)
        errors = [float(e) for e in re.findall(r"\[max error\s+([0-9.eE+-]+)\]", text)]
        assert len(errors) == len(FORMATS), text
        # the harness's quiet band: include/sparse_operation.h:150-158 prints a
        # round-error note above 5 sqrt(eps) of the guarded elementwise error
        eps = np.finfo(np.float64 if prec == 64 else np.float32).eps
        assert max(errors) <= 5 * np.sqrt(eps), text
        for fmt in FORMATS:
            line = [l for l in text.splitlines() if re.match(rf"^7 {name}\.mtx {fmt} ", l)]
            assert line and "skipped" not in line[0], (fmt, text)
            ms = float(line[0].split()[4])
            assert ms > 0
        assert "POSSIBLE FAILURE" not in text


def test_driver_reads_symmetric_fixture():
    path = os.path.join(ROOT, "tests", "golden", "mtx", "symmetric_real.mtx")
    text = run_driver(path, 64)
    assert "matrix symmetric_real.mtx: 10 x 10" in text
    assert "POSSIBLE FAILURE" not in text


def test_python_mirror_matches_oracle(oracle):
    """sspmv_amd.lespmv: reference names from Python, arrays equal to the oracle's."""
    from oracle.oracle import SELL_C_SIGMA, CSR5, reference_params
    from sspmv_amd import lespmv
    A = random_csr(400, 400, seed=12, long_rows=2, long_len=180)
    csr = lespmv.CSR_Matrix(A.m, A.n, A.rp, A.ci, A.v)
    s = lespmv.csr_to_sell_c_sigma(csr, None)
    rc, res = oracle.convert(A.m, A.n, A.rp, A.ci, A.v, SELL_C_SIGMA, reference_params(SELL_C_SIGMA, np.float64))
    assert np.array_equal(s.reorder, res.get("reorder", np.int32))
    assert np.array_equal(s.col_index, res.get("col_index", np.int32))
    c5 = lespmv.csr_to_csr5(csr)
    rc, res = oracle.convert(A.m, A.n, A.rp, A.ci, A.v, CSR5, reference_params(CSR5, np.float64))
    assert np.array_equal(c5.tile_desc, res.get("tile_desc", np.uint32)) and c5._p == res.scalar("p")
    x = np.random.default_rng(0).random(A.n)
    y = np.zeros(A.m)
    lespmv.LeSpMV_sell_c_sigma(0.8, s, x, 0.7, y)
    ref = np.zeros(A.m)
    oracle.spmv_csr(A.m, A.n, A.rp, A.ci, A.v, 0.8, x, 0.7, ref)
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)
    for fn in (lespmv.test_coo_matrix_kernels, lespmv.test_dia_matrix_kernels, lespmv.test_bsr_matrix_kernels):
        ms, conv = fn(csr, 0)
        assert ms > 0
    assert lespmv.test_csr_matrix_kernels(csr, 2) > 0
