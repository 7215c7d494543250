"""Python mirror of the reference's SpMV interface on the MI355X library.

Same names, argument meanings and defaults as the reference's C++ API
(LeSpMV/include/sparse_format.h, sparse_conversion.h, spmv_*.h, sparse_io.h,
spmv_testroutine.h), for Python callers such as the SmartAdpter side, which in
the reference only read files. Conversions and SpMV run on the GPU through the
C ABI; struct fields are numpy arrays in the reference's layout. Errors raise
`lib.LespmvError` (the reference calls exit()).
"""
from __future__ import annotations

import time

import numpy as np

from . import lib

RowMajor, ColMajor = 0, 1

_FIELDS = {
    lib.COO: ["row_index", "col_index", "values"],
    lib.ELL: ["col_index", "values"],
    lib.SELL: ["row_width", "chunk_offset", "col_index", "values"],
    lib.SELL_C_SIGMA: ["reorder", "chunk_len", "chunk_offset", "col_index", "values"],
    lib.SELL_C_R: ["reorder", "chunk_len", "chunk_offset", "col_index", "values"],
    lib.DIA: ["diag_offsets", "diag_data"],
    lib.BSR: ["row_ptr", "block_colindex", "block_data"],
    lib.CSR5: ["row_offset", "col_index", "values", "tile_ptr", "tile_desc", "tile_desc_offset_ptr",
               "tile_desc_offset"],
}
_SCALARS = {
    lib.ELL: ["max_row_width", "ld"],
    lib.SELL: ["chunk_num", "sliceWidth", "alignment"],
    lib.SELL_C_SIGMA: ["chunkWidth_C", "sliceWidth_Sigma", "sliceNum", "chunk_num_per_slice", "chunkNum",
                       "validchunkNum", "alignment"],
    lib.SELL_C_R: ["chunkWidth_C", "validchunkNum", "alignment"],
    lib.DIA: ["complete_ndiags", "stride"],
    lib.BSR: ["mb", "nb", "nnzb", "blockDim_r", "blockDim_c"],
    lib.CSR5: ["omega", "sigma", "bit_y_offset", "bit_scansum_offset", "num_packets", "num_offsets", "p",
               "tail_partition_start"],
    lib.COO: [],
}
_INT_FIELDS = {"diag_offsets": np.int64, "chunk_offset": np.int64, "tile_ptr": np.uint32, "tile_desc": np.uint32}


class Matrix_Features:
    """Common fields (include/sparse_format.h:49-68) plus the device handle."""

    fmt = lib.CSR

    def __init__(self):
        self.num_rows = self.num_cols = self.num_nnzs = 0
        self.partition = None
        self.time = self.gflops = self.gbytes = 0.0
        self.tag = 0
        self.kernel_flag = 0
        self.gpu: lib.Matrix | None = None

    @property
    def dtype(self):
        return self.values.dtype if hasattr(self, "values") else self.gpu.dtype


class CSR_Matrix(Matrix_Features):
    fmt = lib.CSR

    def __init__(self, num_rows=0, num_cols=0, row_offset=None, col_index=None, values=None):
        super().__init__()
        self.num_rows, self.num_cols = int(num_rows), int(num_cols)
        self.row_offset = np.ascontiguousarray(row_offset, dtype=np.int32) if row_offset is not None else None
        self.col_index = np.ascontiguousarray(col_index, dtype=np.int32) if col_index is not None else None
        self.values = np.ascontiguousarray(values) if values is not None else None
        self.num_nnzs = 0 if col_index is None else int(len(col_index))
        self.kernel_flag = 1

    def device(self) -> lib.Matrix:
        if self.gpu is None:
            self.gpu = lib.Matrix(self.num_rows, self.num_cols, self.row_offset, self.col_index, self.values)
        return self.gpu


class _Converted(Matrix_Features):
    def __init__(self, fmt, csr: CSR_Matrix, params: lib.Params):
        super().__init__()
        self.fmt = fmt
        self.num_rows, self.num_cols, self.num_nnzs = csr.num_rows, csr.num_cols, csr.num_nnzs
        self.kernel_flag = csr.kernel_flag
        self.gpu = lib.Matrix(csr.num_rows, csr.num_cols, csr.row_offset, csr.col_index, csr.values)
        self.gpu.convert(fmt, params)
        vdt = csr.values.dtype
        for name in _FIELDS[fmt]:
            dt = vdt if name in ("values", "diag_data", "block_data") else _INT_FIELDS.get(name, np.int32)
            setattr(self, name, self.gpu.export(fmt, name, dt))
        for name in _SCALARS[fmt]:
            setattr(self, "_p" if name == "p" else name, self.gpu.export_scalar(fmt, name))


class COO_Matrix(_Converted):
    pass


class ELL_Matrix(_Converted):
    pass


class S_ELL_Matrix(_Converted):
    pass


class SELL_C_Sigma_Matrix(_Converted):
    pass


class SELL_C_R_Matrix(_Converted):
    pass


class DIA_Matrix(_Converted):
    pass


class BSR_Matrix(_Converted):
    pass


class CSR5_Matrix(_Converted):
    pass


def _align(dtype, alignment):
    return alignment if alignment is not None else 64 // np.dtype(dtype).itemsize  # SIMD_WIDTH/8/sizeof(V)


# ---- input (utils/sparse_io.cpp:179-204) -------------------------------------

def read_csr_matrix(mm_filename: str, dtype=np.float64, compact: bool = False) -> CSR_Matrix:
    h = lib.read_mtx(mm_filename, dtype, compact)
    return CSR_Matrix(h.m, h.n, h.rp, h.ci, h.v)


# ---- conversions (include/sparse_conversion.h) --------------------------------

def csr_to_coo(csr: CSR_Matrix) -> COO_Matrix:
    return COO_Matrix(lib.COO, csr, lib.make_params())


def csr_to_ell(csr: CSR_Matrix, ld: int = RowMajor) -> ELL_Matrix:
    return ELL_Matrix(lib.ELL, csr, lib.make_params(ld=ld))


def csr_to_sell(csr: CSR_Matrix, fp_feature=None, chunkwidth: int = 4, alignment: int | None = None) -> S_ELL_Matrix:
    return S_ELL_Matrix(lib.SELL, csr, lib.make_params(chunk=chunkwidth, alignment=_align(csr.values.dtype,
                                                                                            alignment)))


def csr_to_sell_c_sigma(csr: CSR_Matrix, fp_feature=None, slicewidth: int = 16384, chunkwidth: int = 4,
                        alignment: int | None = None) -> SELL_C_Sigma_Matrix:
    return SELL_C_Sigma_Matrix(lib.SELL_C_SIGMA, csr, lib.make_params(chunk=chunkwidth, sigma=slicewidth,
                                                                      alignment=_align(csr.values.dtype, alignment)))


def csr_to_sell_c_R(csr: CSR_Matrix, fp_feature=None, chunkwidth: int = 4,
                    alignment: int | None = None) -> SELL_C_R_Matrix:
    return SELL_C_R_Matrix(lib.SELL_C_R, csr, lib.make_params(chunk=chunkwidth,
                                                              alignment=_align(csr.values.dtype, alignment)))


def csr_to_dia(csr: CSR_Matrix, max_diags: int = 10240, fp_feature=None, alignment: int | None = None) -> DIA_Matrix:
    return DIA_Matrix(lib.DIA, csr, lib.make_params(max_diags=max_diags,
                                                    alignment=_align(csr.values.dtype, alignment)))


def csr_to_bsr(csr: CSR_Matrix, blockDimRow: int = 16, blockDimCol: int | None = None) -> BSR_Matrix:
    return BSR_Matrix(lib.BSR, csr, lib.make_params(block_rows=blockDimRow,
                                                    block_cols=_align(csr.values.dtype, blockDimCol)))


def csr_to_csr5(csr: CSR_Matrix, fp_feature=None) -> CSR5_Matrix:
    return CSR5_Matrix(lib.CSR5, csr, lib.make_params(omega=_align(csr.values.dtype, None), sigma=16))


# ---- SpMV entry points (include/spmv_*.h) --------------------------------------

def _spmv(alpha, A: Matrix_Features, x, beta, y):
    """LeSpMV_<fmt>(alpha, A, x, beta, y); x, y: numpy arrays (staged through
    the device) or lib.DeviceArray."""
    M = A.device() if isinstance(A, CSR_Matrix) else A.gpu
    if isinstance(x, lib.DeviceArray):
        M.spmv(A.fmt, alpha, x, beta, y, variant=A.kernel_flag)
        return y
    dx = lib.DeviceArray.from_numpy(np.ascontiguousarray(x, dtype=M.dtype))
    dy = lib.DeviceArray.from_numpy(np.ascontiguousarray(y, dtype=M.dtype))
    M.spmv(A.fmt, alpha, dx, beta, dy, variant=A.kernel_flag)
    y[...] = dy.numpy()
    return y


LeSpMV_csr = LeSpMV_coo = LeSpMV_ell = LeSpMV_sell = LeSpMV_sell_c_sigma = LeSpMV_sell_c_R = LeSpMV_dia = \
    LeSpMV_bsr = LeSpMV_csr5 = _spmv


# ---- harness (include/sparse_operation.h:75-193, spmv_benchmark.h:166-228) -----

def maximum_relative_error(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    eps = np.sqrt(np.finfo(a.dtype).eps)
    err = np.abs(a - b)
    nz = err != 0
    return float(np.max(err[nz] / (np.abs(a[nz]) + np.abs(b[nz]) + eps))) if nz.any() else 0.0


def compare_spmv_kernels(csr_ref: CSR_Matrix, A: Matrix_Features, seed: int = 0) -> float:
    dt = csr_ref.values.dtype
    x = np.random.default_rng(seed).random(csr_ref.num_cols).astype(dt)
    y1 = np.zeros(csr_ref.num_rows, dt)
    y2 = np.zeros(csr_ref.num_rows, dt)
    ref = CSR_Matrix(csr_ref.num_rows, csr_ref.num_cols, csr_ref.row_offset, csr_ref.col_index, csr_ref.values)
    ref.kernel_flag = 0
    LeSpMV_csr(0.8, ref, x, 0.7, y1)
    _spmv(0.8, A, x, 0.7, y2)
    return maximum_relative_error(y1, y2)


def benchmark_spmv_on_device(A: Matrix_Features, min_iter: int = 20, max_iter: int = 1000,
                             seconds: float = 20.0) -> float:
    """Device-timed benchmark with the reference's iteration policy; returns ms."""
    M = A.device() if isinstance(A, CSR_Matrix) else A.gpu
    dt = M.dtype
    dx = lib.DeviceArray.from_numpy(np.random.default_rng(1).random(A.num_cols).astype(dt))
    dy = lib.DeviceArray.empty(A.num_rows, dt)
    est = M.time_spmv(A.fmt, 1.0, dx, 0.0, dy, 1, A.kernel_flag)
    iters = max_iter if est < 20.0 else min(max_iter, max(min_iter, int(seconds * 1000 / est)))
    ms = M.time_spmv(A.fmt, 1.0, dx, 0.0, dy, iters, A.kernel_flag)
    A.time = ms
    A.gflops = 2.0 * A.num_nnzs / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    A.gbytes = M.reference_bytes(A.fmt) / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return ms


# ---- test routines (include/spmv_testroutine.h:7-50) ---------------------------

def _test(csr_ref, make, kernel_tag, name):
    t0 = time.perf_counter()
    A = make()
    convert_ms = A.gpu.convert_ms(A.fmt) if A.fmt != lib.CSR else 0.0
    A.kernel_flag = kernel_tag
    err = compare_spmv_kernels(csr_ref, A)
    print(f"\ttesting {name:<26s}[gpu]: [max error {err:9f}]")
    ms = benchmark_spmv_on_device(A)
    print(f"\tbenchmarking {name:<20s} [gpu]: {ms:8.4f} ms ( {A.gflops:5.4f} GFLOP/s {A.gbytes:5.4f} GB/s)"
          f"  (host {1e3 * (time.perf_counter() - t0):.1f} ms)")
    return ms, convert_ms, err


def test_csr_matrix_kernels(csr_ref, kernel_tag, schedule_mod=0):
    def make():
        a = CSR_Matrix(csr_ref.num_rows, csr_ref.num_cols, csr_ref.row_offset, csr_ref.col_index, csr_ref.values)
        a.device()
        return a
    return _test(csr_ref, make, kernel_tag, "csr_gpu")[0]


def test_coo_matrix_kernels(csr_ref, kernel_tag, schedule_mod=0):
    return _test(csr_ref, lambda: csr_to_coo(csr_ref), kernel_tag, "coo_gpu")[:2]


def test_ell_matrix_kernels(csr_ref, kernel_tag, ld=RowMajor, schedule_mod=0):
    return _test(csr_ref, lambda: csr_to_ell(csr_ref, ld), kernel_tag, "ell_gpu")[:2]


def test_s_ell_matrix_kernels(csr_ref, kernel_tag, schedule_mod=0):
    return _test(csr_ref, lambda: csr_to_sell(csr_ref), kernel_tag, "s_ell_gpu")[:2]


def test_sell_c_sigma_matrix_kernels(csr_ref, kernel_tag, schedule_mod=0):
    return _test(csr_ref, lambda: csr_to_sell_c_sigma(csr_ref), kernel_tag, "sell_c_sigma_gpu")[:2]


def test_sell_c_R_matrix_kernels(csr_ref, kernel_tag, schedule_mod=0):
    return _test(csr_ref, lambda: csr_to_sell_c_R(csr_ref), kernel_tag, "sell_c_R_gpu")[0]


def test_dia_matrix_kernels(csr_ref, kernel_tag, schedule_mod=0):
    return _test(csr_ref, lambda: csr_to_dia(csr_ref), kernel_tag, "dia_gpu")[:2]


def test_bsr_matrix_kernels(csr_ref, kernel_tag, schedule_mod=0):
    return _test(csr_ref, lambda: csr_to_bsr(csr_ref), kernel_tag, "bsr_gpu")[:2]


def test_csr5_matrix_kernels(csr_ref, kernel_tag, schedule_mod=0):
    return _test(csr_ref, lambda: csr_to_csr5(csr_ref), kernel_tag, "csr5_gpu")[0]
