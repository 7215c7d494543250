"""The oracle against fixtures written by the compiled reference.

Integer and index arrays must match bit-exactly; values are copies so they
must match exactly too; y vectors of the serial kernels (same summation order)
must match to 1e-14 relative.
"""
import numpy as np
import pytest

from oracle.oracle import (BSR, COO, CSR5, DIA, ELL, SCALARS, SELL, SELL_C_R, SELL_C_SIGMA, array_dtypes)

FORMATS = [COO, DIA, ELL, SELL, SELL_C_SIGMA, SELL_C_R, CSR5, BSR]


def cases(golden):
    out = []
    for k in golden.files:
        p = k.split("/")
        if len(p) == 5 and p[4] == "params":
            out.append((p[0], p[1], int(p[2]), int(p[3])))
    return sorted(set(out))


def matrix(golden, dn, name):
    g = lambda s: golden[f"{dn}/{name}/{s}"]
    return int(g("m")), int(g("n")), g("rp"), g("ci"), g("v"), g("x"), g("y0")


def test_fixture_has_all_formats(golden):
    fmts = {c[2] for c in cases(golden)}
    assert fmts == set(FORMATS)


def test_conversions_bit_exact(golden, oracle):
    checked = 0
    for dn, name, fmt, pi in cases(golden):
        base = f"{dn}/{name}/{fmt}/{pi}"
        m, n, rp, ci, v, x, y0 = matrix(golden, dn, name)
        prm = golden[f"{base}/params"]
        rc, res = oracle.convert(m, n, rp, ci, v, fmt, prm)
        assert rc == int(golden[f"{base}/rc"]), base
        if rc:
            continue
        for arr, adt in array_dtypes(fmt, v.dtype).items():
            key = f"{base}/{arr}"
            if key not in golden.files:
                continue
            got, exp = res.get(arr, adt), golden[key]
            if arr == "tile_desc_offset":
                mask = golden[f"{base}/tile_desc_offset_written"]
                assert got.shape == exp.shape, key
                assert np.array_equal(got[mask], exp[mask]), key
            else:
                assert got.shape == exp.shape and np.array_equal(got, exp), key
            checked += 1
        for s in SCALARS[fmt]:
            assert res.scalar(s) == int(golden[f"{base}/{s}"]), (base, s)
    assert checked > 300


def test_spmv_matches_reference(golden, oracle):
    for dn, name, fmt, pi in cases(golden):
        base = f"{dn}/{name}/{fmt}/{pi}"
        if f"{base}/y" not in golden.files:
            continue
        m, n, rp, ci, v, x, y0 = matrix(golden, dn, name)
        rc, res = oracle.convert(m, n, rp, ci, v, fmt, golden[f"{base}/params"])
        tol = 1e-13 if v.dtype == np.float64 else 2e-6
        y = np.zeros(m, v.dtype)
        oracle.spmv_fmt(res, fmt, m, n, len(ci), 0.8, x, 0.7, y)
        np.testing.assert_allclose(y, golden[f"{base}/y"], rtol=tol, atol=tol, err_msg=base)
        if f"{base}/y_beta" in golden.files:
            y = y0.copy()
            oracle.spmv_fmt(res, fmt, m, n, len(ci), 0.8, x, 0.7, y)
            np.testing.assert_allclose(y, golden[f"{base}/y_beta"], rtol=tol, atol=tol, err_msg=base + " beta")


def test_csr_kernels_match_reference(golden, oracle):
    names = {k.split("/")[1] for k in golden.files if k.startswith("float64/")}
    for dn in ("float64", "float32"):
        for name in sorted(names):
            if f"{dn}/{name}/m" not in golden.files:
                continue
            m, n, rp, ci, v, x, y0 = matrix(golden, dn, name)
            tol = 1e-13 if dn == "float64" else 2e-6
            for flag in (0, 2):
                y = np.zeros(m, v.dtype)
                oracle.spmv_csr(m, n, rp, ci, v, 0.8, x, 0.7, y, flag=flag, threads=4)
                np.testing.assert_allclose(y, golden[f"{dn}/{name}/csr/{flag}/y"], rtol=tol, atol=tol)
            y = y0.copy()
            oracle.spmv_csr(m, n, rp, ci, v, 0.8, x, 0.7, y)
            np.testing.assert_allclose(y, golden[f"{dn}/{name}/csr/0/y_beta"], rtol=tol, atol=tol)


@pytest.mark.parametrize("compact", [0, 1])
def test_coo_to_csr(golden, oracle, compact):
    m, n = int(golden["coo2csr/m"]), int(golden["coo2csr/n"])
    res = oracle.coo_to_csr(m, n, golden["coo2csr/ri"], golden["coo2csr/ci"], golden["coo2csr/v"], bool(compact))
    for arr, adt in (("row_offset", np.int32), ("col_index", np.int32), ("values", np.float64)):
        assert np.array_equal(res.get(arr, adt), golden[f"coo2csr/{compact}/{arr}"]), arr
