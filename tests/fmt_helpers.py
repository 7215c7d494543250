"""Shared helpers for the per-format GPU parity tests."""
import numpy as np

from oracle.oracle import (BSR, COO, CSR5, DIA, ELL, SCALARS, SELL, SELL_C_R, SELL_C_SIGMA, array_dtypes,
                           reference_params)


def lib_params(gpu, fmt, prm):
    """Oracle parameter list [a, b, alignment, max_diags] -> lespmv_params_t."""
    a, b, align, maxd = (int(v) for v in prm)
    if fmt == ELL:
        return gpu.make_params(ld=a)
    if fmt == SELL:
        return gpu.make_params(chunk=a, alignment=align)
    if fmt == SELL_C_SIGMA:
        return gpu.make_params(chunk=a, sigma=b, alignment=align)
    if fmt == SELL_C_R:
        return gpu.make_params(chunk=a, alignment=align)
    if fmt == DIA:
        return gpu.make_params(alignment=align, max_diags=maxd)
    if fmt == BSR:
        return gpu.make_params(block_rows=a, block_cols=b)
    if fmt == CSR5:
        return gpu.make_params(omega=a, sigma=b)
    return gpu.make_params()


def gpu_native_params(fmt):
    if fmt == SELL:
        return [[64, 0, 1, 0], [32, 0, 1, 0]]
    if fmt == SELL_C_SIGMA:
        return [[64, 4096, 1, 0], [64, 128, 1, 0], [32, 256, 1, 0]]
    if fmt == SELL_C_R:
        return [[64, 0, 1, 0], [32, 0, 1, 0]]
    if fmt == ELL:
        return [[1, 0, 0, 0]]
    if fmt == BSR:
        return [[4, 4, 0, 0], [2, 8, 0, 0], [3, 5, 0, 0], [8, 8, 0, 0]]
    if fmt == DIA:
        return [[0, 0, 1, 10240]]
    if fmt == CSR5:
        return [[64, 16, 0, 0], [64, 8, 0, 0]]
    return []


def param_sets(fmt, dt):
    return [reference_params(fmt, dt)] + gpu_native_params(fmt)


def compare_exports(M, res, fmt, dt, ctx, written_mask=None):
    """GPU export vs oracle conversion: every array bit-exact, every scalar equal."""
    for arr, adt in array_dtypes(fmt, dt).items():
        got = M.export(fmt, arr, adt)
        exp = res.get(arr, adt)
        if arr == "tile_desc_offset" and written_mask is not None:
            assert got.shape == exp.shape, (ctx, arr)
            assert np.array_equal(got[written_mask], exp[written_mask]), (ctx, arr)
            continue
        assert got.shape == exp.shape, (ctx, arr, got.shape, exp.shape)
        assert np.array_equal(got, exp), (ctx, arr, np.flatnonzero(got != exp)[:10])
    for s in SCALARS[fmt]:
        assert M.export_scalar(fmt, s) == res.scalar(s), (ctx, s)
