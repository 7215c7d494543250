"""ctypes front end of the CPU oracle (and, where built, the reference checker).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg, never by the ``sspmv_amd`` package.

Format ids follow the reference's label map (LeSpMV/script/write_label.py:63-75):
COO 0, CSR 1, DIA 2, ELL 3, S-ELL 4, SELL-c-sigma 5, SELL-c-R 6, CSR5 7, BSR 8.
Conversion parameters are passed as ``[a, b, alignment, max_diags]`` where
``a`` is C (sliced formats), ld (ELL), br (BSR) or omega (CSR5) and ``b`` is
sigma (SELL-c-sigma, CSR5) or bc (BSR).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_lespmv.so")

COO, CSR, DIA, ELL, SELL, SELL_C_SIGMA, SELL_C_R, CSR5, BSR = range(9)

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def _dtype_id(v: np.ndarray) -> int:
    if v.dtype == np.float64:
        return 1
    if v.dtype == np.float32:
        return 0
    raise TypeError(f"unsupported value dtype {v.dtype}")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Result:
    """A bag of named arrays produced by one conversion."""

    def __init__(self, lib, prefix: str):
        self._lib = lib
        self._p = prefix
        self.h = getattr(lib, f"{prefix}_result_new")()

    def __del__(self):
        try:
            getattr(self._lib, f"{self._p}_result_free")(self.h)
        except Exception:
            pass

    def names(self):
        buf = C.create_string_buffer(1 << 16)
        getattr(self._lib, f"{self._p}_result_names")(self.h, buf, len(buf))
        return [n for n in buf.value.decode().split(",") if n]

    def get(self, name: str, dtype) -> np.ndarray:
        nbytes = getattr(self._lib, f"{self._p}_result_size")(self.h, name.encode())
        if nbytes < 0:
            raise KeyError(name)
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        getattr(self._lib, f"{self._p}_result_get")(self.h, name.encode(), _ptr(out))
        return out

    def scalar(self, name: str) -> int:
        return int(self.get(name, np.int64)[0])

    def error(self) -> str:
        return getattr(self._lib, f"{self._p}_result_error")(self.h).decode()


def _setup_result_api(lib, prefix):
    getattr(lib, f"{prefix}_result_new").restype = C.c_void_p
    getattr(lib, f"{prefix}_result_free").argtypes = [C.c_void_p]
    getattr(lib, f"{prefix}_result_size").argtypes = [C.c_void_p, C.c_char_p]
    getattr(lib, f"{prefix}_result_size").restype = C.c_int64
    getattr(lib, f"{prefix}_result_get").argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
    getattr(lib, f"{prefix}_result_names").argtypes = [C.c_void_p, C.c_char_p, C.c_int64]
    getattr(lib, f"{prefix}_result_error").argtypes = [C.c_void_p]
    getattr(lib, f"{prefix}_result_error").restype = C.c_char_p


# Array names (and dtypes) each format's conversion exports, shared by the
# oracle, the reference driver and the GPU library's export calls.
def array_dtypes(fmt: int, vdt) -> dict:
    vdt = np.dtype(vdt)
    if fmt == COO:
        return {"row_index": np.int32, "col_index": np.int32, "values": vdt}
    if fmt == ELL:
        return {"col_index": np.int32, "values": vdt}
    if fmt == SELL:
        return {"row_width": np.int32, "chunk_offset": np.int64, "col_index": np.int32, "values": vdt}
    if fmt in (SELL_C_SIGMA, SELL_C_R):
        return {"reorder": np.int32, "chunk_len": np.int32, "chunk_offset": np.int64,
                "col_index": np.int32, "values": vdt}
    if fmt == DIA:
        return {"diag_offsets": np.int64, "diag_data": vdt}
    if fmt == BSR:
        return {"row_ptr": np.int32, "block_colindex": np.int32, "block_data": vdt}
    if fmt == CSR5:
        return {"row_offset": np.int32, "col_index": np.int32, "values": vdt, "tile_ptr": np.uint32,
                "tile_desc": np.uint32, "tile_desc_offset_ptr": np.int32, "tile_desc_offset": np.int32}
    raise ValueError(fmt)


SCALARS = {
    ELL: ["max_row_width", "ld"],
    SELL: ["chunk_num", "sliceWidth", "alignment"],
    SELL_C_SIGMA: ["chunkWidth_C", "sliceWidth_Sigma", "sliceNum", "chunk_num_per_slice", "chunkNum",
                   "validchunkNum", "alignment"],
    SELL_C_R: ["chunkWidth_C", "validchunkNum", "alignment"],
    DIA: ["complete_ndiags", "stride"],
    BSR: ["mb", "nb", "nnzb", "blockDim_r", "blockDim_c"],
    CSR5: ["omega", "sigma", "bit_y_offset", "bit_scansum_offset", "num_packets", "num_offsets", "p",
           "tail_partition_start"],
    COO: [],
}


def reference_params(fmt: int, vdt) -> list:
    """The parameters the reference's test routines use (utils/test_spmv_*.cpp)."""
    align = 64 // np.dtype(vdt).itemsize  # SIMD_WIDTH/8/sizeof(V), plat_config.h:17
    if fmt == ELL:
        return [0, 0, align, 0]
    if fmt == SELL:
        return [4, 0, align, 0]  # CHUNK_SIZE, general_config.h:16
    if fmt == SELL_C_SIGMA:
        return [4, 16384, align, 0]  # SELL_SIGMA, general_config.h:15
    if fmt == SELL_C_R:
        return [4, 0, align, 0]
    if fmt == DIA:
        return [0, 0, align, 10240]  # MAX_DIAG_NUM, general_config.h:9
    if fmt == BSR:
        return [16, align, 0, 0]  # BSR_BlockDimRow, test_spmv_bsr.cpp:25-26
    if fmt == CSR5:
        return [align, 16, 0, 0]  # omega, CSR5_SIGMA (general_config.h:24)
    return [0, 0, 0, 0]


IMAGE_DTYPES = {"tile_nnz": np.int32, "tile_max": np.int32, "tile_ave": np.float64, "tile_std": np.float64,
                "tile_uniq_rows": np.int32, "tile_uniq_cols": np.int32, "rb_nnz": np.int32, "rb_max": np.int32,
                "rb_ave": np.float64, "rb_std": np.float64, "cb_nnz": np.int32, "cb_max": np.int32,
                "cb_ave": np.float64, "cb_std": np.float64, "row_nnz": np.int32, "col_nnz": np.int32}


def feature_dicts(r: Result):
    scal, img = {}, {}
    for name in r.names():
        if name.startswith("f:"):
            scal[name[2:]] = float(r.get(name, np.float64)[0])
        elif name.startswith("img:"):
            img[name[4:]] = r.get(name, IMAGE_DTYPES[name[4:]])
    return scal, img


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
        lib = C.CDLL(path)
        _setup_result_api(lib, "oracle")
        lib.oracle_convert.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p,
                                       C.c_void_p, np.ctypeslib.ndpointer(np.int64)]
        lib.oracle_coo_to_csr.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p,
                                          C.c_void_p, C.c_int]
        lib.oracle_spmv_csr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p, C.c_void_p,
                                        C.c_double, C.c_void_p, C.c_double, C.c_void_p, C.c_int]
        lib.oracle_spmv_fmt.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                        C.c_void_p, C.c_double, C.c_void_p]
        lib.oracle_time_csr_lb.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        lib.oracle_time_csr_lb.restype = C.c_double
        lib.oracle_max_rel_error.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
        lib.oracle_max_rel_error.restype = C.c_double
        lib.oracle_features.argtypes = [C.c_void_p, C.c_char_p]
        self.lib = lib

    def num_threads(self) -> int:
        return self.lib.oracle_num_threads()

    def convert(self, m, n, rp, ci, v, fmt, params) -> tuple[int, Result]:
        r = Result(self.lib, "oracle")
        rc = self.lib.oracle_convert(r.h, _dtype_id(v), fmt, m, n, len(ci), rp, ci, _ptr(v),
                                     np.asarray(params, dtype=np.int64))
        return rc, r

    def coo_to_csr(self, m, n, ri, ci, v, compact=False) -> Result:
        r = Result(self.lib, "oracle")
        self.lib.oracle_coo_to_csr(r.h, _dtype_id(v), m, n, len(ci), ri, ci, _ptr(v), int(compact))
        return r

    def spmv_csr(self, m, n, rp, ci, v, alpha, x, beta, y, flag=0, threads=1):
        self.lib.oracle_spmv_csr(_dtype_id(v), flag, m, n, len(ci), rp, ci, _ptr(v), alpha, _ptr(x), beta,
                                 _ptr(y), threads)
        return y

    def spmv_fmt(self, res: Result, fmt, m, n, nnz, alpha, x, beta, y):
        rc = self.lib.oracle_spmv_fmt(res.h, _dtype_id(x), fmt, m, n, nnz, alpha, _ptr(x), beta, _ptr(y))
        if rc:
            raise RuntimeError(f"oracle spmv failed for format {fmt}: {rc}")
        return y

    def time_csr_lb(self, m, n, rp, ci, v, x, y, threads, iters) -> float:
        return self.lib.oracle_time_csr_lb(_dtype_id(v), m, n, len(ci), rp, ci, _ptr(v), _ptr(x), _ptr(y),
                                           threads, iters)

    def features(self, path: str):
        """The MTX features of a MatrixMarket file: (scalars, images)."""
        r = Result(self.lib, "oracle")
        if self.lib.oracle_features(r.h, path.encode()):
            raise RuntimeError(f"oracle features failed: {r.error()}")
        return feature_dicts(r)

    def max_rel_error(self, a, b) -> float:
        return self.lib.oracle_max_rel_error(_dtype_id(a), _ptr(a), _ptr(b), len(a))


class Reference:
    """The compiled reference (oracle/_ref), present only in the build container."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise RuntimeError(f"reference checker not built: {path}")
        lib = C.CDLL(path)
        _setup_result_api(lib, "ref")
        lib.ref_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p, C.c_void_p,
                                np.ctypeslib.ndpointer(np.int64), C.c_int, C.c_double, C.c_void_p, C.c_double,
                                C.c_void_p]
        lib.ref_coo_to_csr.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p, C.c_void_p,
                                       C.c_int]
        lib.ref_set_threads.argtypes = [C.c_int]
        lib.ref_read_mtx.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int]
        lib.ref_features.argtypes = [C.c_void_p, C.c_char_p]
        self.lib = lib

    def set_threads(self, t: int):
        self.lib.ref_set_threads(t)

    def run(self, m, n, rp, ci, v, fmt, params, flag=0, alpha=0.8, x=None, beta=0.7, y=None):
        r = Result(self.lib, "ref")
        rc = self.lib.ref_run(r.h, _dtype_id(v), fmt, m, n, len(ci), rp, ci, _ptr(v),
                              np.asarray(params, dtype=np.int64), flag, alpha, _ptr(x), beta, _ptr(y))
        return rc, r

    def read_mtx(self, path: str, dtype=np.float64, compact=False) -> Result:
        r = Result(self.lib, "ref")
        self.lib.ref_read_mtx(r.h, 1 if np.dtype(dtype) == np.float64 else 0, path.encode(), int(compact))
        return r

    def features(self, path: str):
        """Get_features' values for a file (MTX<int, double>): (scalars, images)."""
        r = Result(self.lib, "ref")
        if self.lib.ref_features(r.h, path.encode()):
            raise RuntimeError("reference features failed")
        return feature_dicts(r)

    def coo_to_csr(self, m, n, ri, ci, v, compact=False) -> Result:
        r = Result(self.lib, "ref")
        self.lib.ref_coo_to_csr(r.h, _dtype_id(v), m, n, len(ci), ri, ci, _ptr(v), int(compact))
        return r
