"""The C-ABI library loads and exports every declared entry point (no GPU needed),
and its host-side pieces (generators, parameters, error reporting) behave."""
import os
import re

import numpy as np
import pytest

from sspmv_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "lespmv_hip.h")).read()
    return sorted(set(re.findall(r"\b(lespmv_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = lib.load()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_default_params_reference_layout():
    p = lib.default_params(lib.SELL_C_SIGMA, np.float64, reference_layout=True)
    assert (p.chunk, p.sigma, p.alignment) == (4, 16384, 8)
    p = lib.default_params(lib.SELL_C_SIGMA, np.float32, reference_layout=True)
    assert p.alignment == 16
    p = lib.default_params(lib.CSR5, np.float64, reference_layout=True)
    assert (p.omega, p.sigma) == (8, 16)
    p = lib.default_params(lib.BSR, np.float64, reference_layout=True)
    assert (p.block_rows, p.block_cols) == (16, 8)
    p = lib.default_params(lib.SELL_C_SIGMA, np.float64, reference_layout=False)
    assert p.chunk == 64
    with pytest.raises(lib.LespmvError) as e:
        lib.default_params(42, np.float64)
    assert e.value.code == lib.ERR_ARG


def test_laplace_generator_closed_form():
    A = lib.generate(lib.GEN_LAPLACE2D, [32, 16, 4.0, -1.0], 0)
    assert A.m == 512 and A.nnz == 5 * 512 - 2 * 32 - 2 * 16
    d = np.zeros((A.m, A.n))
    rows = np.repeat(np.arange(A.m), np.diff(A.rp))
    d[rows, A.ci] = A.v
    assert np.allclose(d, d.T) and np.allclose(np.diag(d), 4.0)
    assert np.allclose(d.sum(axis=1)[np.diff(A.rp) == 5], 0.0)


def test_stencil27_nnz():
    A = lib.generate(lib.GEN_STENCIL27, [7, 26.0, -1.0], 0)
    assert A.m == 343 and A.nnz == (3 * 7 - 2) ** 3
    assert all(np.all(np.diff(A.ci[A.rp[i]:A.rp[i + 1]]) > 0) for i in range(A.m))


@pytest.mark.parametrize("family,params", [
    (lib.GEN_LOGNORMAL, [20000, 20000, 16.0, 1.0, 4096, 0.3, 64]),
    (lib.GEN_BLOCKRAND, [500, 500, 4, 4, 8]),
    (lib.GEN_BANDED, [5000, 6, 0.7]),
    (lib.GEN_POWERLAW, [5000, 2.2, 8.0, 1000]),
    (lib.GEN_BLOCKDIAG, [3000, 16]),
    (lib.GEN_UNIFORM, [3000, 4000, 1e-3]),
])
def test_generators_deterministic_and_well_formed(family, params):
    A = lib.generate(family, params, 7)
    B = lib.generate(family, params, 7)
    C = lib.generate(family, params, 8)
    assert np.array_equal(A.rp, B.rp) and np.array_equal(A.ci, B.ci) and np.array_equal(A.v, B.v)
    assert not (np.array_equal(A.ci, C.ci) and np.array_equal(A.v, C.v))
    assert A.rp[0] == 0 and A.rp[-1] == A.nnz and np.all(np.diff(A.rp) >= 0)
    assert A.nnz == 0 or (A.ci.min() >= 0 and A.ci.max() < A.n)
    for i in range(0, A.m, max(1, A.m // 200)):
        assert np.all(np.diff(A.ci[A.rp[i]:A.rp[i + 1]]) > 0)
    assert np.all(np.abs(A.v) <= 1.0)


def test_lognormal_row_statistics():
    A = lib.generate(lib.GEN_LOGNORMAL, [200000, 200000, 16.0, 1.0, 4096, 0.3, 64], 2)
    lens = np.diff(A.rp)
    assert abs(lens.mean() - 16.0) < 0.3
    assert lens.min() >= 1 and lens.max() <= 4096
    rows = np.repeat(np.arange(A.m), lens)
    near = np.abs(A.ci.astype(np.int64) - rows) <= 64
    assert 0.28 < near.mean() < 0.34


def test_fill_uniform_matches_definition():
    x = lib.fill_uniform(1000, 4)
    def splitmix(v):
        m = (1 << 64) - 1
        z = (v + 0x9E3779B97F4A7C15) & m
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        return z ^ (z >> 31)
    ref = np.array([2.0 * ((splitmix(4 ^ i) >> 11) * 2.0 ** -53) - 1.0 for i in range(1000)])
    assert np.array_equal(x, ref)


def test_errors_are_status_codes_not_exits():
    with pytest.raises(lib.LespmvError) as e:
        lib.generate(99, [1.0], 0)
    assert e.value.code == lib.ERR_ARG and "unknown generator" in str(e.value)
    L = lib.load()
    assert L.lespmv_destroy(None) == 0
    assert L.lespmv_spmv(None, 1, 0, 1.0, None, 0.0, None, None) == lib.ERR_ARG
    assert b"null matrix handle" in L.lespmv_last_error()
