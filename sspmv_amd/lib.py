"""ctypes binding of the C ABI in include/lespmv_hip.h (libsspmv_hip.so).

The library is the product: every SpMV here runs as a gfx950 HIP kernel. If the
shared library is missing this module raises at import of `load()`; there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsspmv_hip.so")

# format ids (LeSpMV/script/write_label.py:63-75)
COO, CSR, DIA, ELL, SELL, SELL_C_SIGMA, SELL_C_R, CSR5, BSR = range(9)
FORMAT_NAMES = ["COO", "CSR", "DIA", "ELL", "S-ELL", "SELL-c-sigma", "SELL-c-R", "CSR5", "BSR"]
F32, F64 = 0, 1
VARIANT_AUTO, VARIANT_VECTOR, VARIANT_MERGE, VARIANT_ATOMIC, VARIANT_PANEL, VARIANT_PANEL16 = 0, 1, 2, 3, 4, 5
VARIANT_PANEL4 = 6
VARIANT_ROW_TILES = 7
GEN_LAPLACE2D, GEN_STENCIL27, GEN_LOGNORMAL, GEN_BLOCKRAND, GEN_BANDED, GEN_POWERLAW, GEN_BLOCKDIAG, GEN_UNIFORM = \
    range(8)

ERR_ARG, ERR_OOM, ERR_UNSUPPORTED, ERR_MAX_DIAGS, ERR_DEVICE, ERR_NOT_CONVERTED = range(1, 7)


class LespmvError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[lespmv error {code}] {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("chunk", C.c_int32), ("sigma", C.c_int32), ("alignment", C.c_int32), ("ld", C.c_int32),
                ("max_diags", C.c_int32), ("block_rows", C.c_int32), ("block_cols", C.c_int32),
                ("omega", C.c_int32), ("max_fill", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class HostCsrC(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64), ("dtype", C.c_int32),
                ("mm_flags", C.c_int32), ("row_offset", C.POINTER(C.c_int32)), ("col_index", C.POINTER(C.c_int32)),
                ("values", C.c_void_p)]


# lespmv_features_t, field for field (include/lespmv_hip.h)
FEATURE_INT_FIELDS = [
    "num_rows", "num_cols", "num_nnzs", "nnz_mtx", "nnz_lower", "nnz_upper", "nnz_diagonal", "complete_ndiags",
    "close_nnz", "min_nnz_each_row", "max_nnz_each_row", "min_nnz_each_col", "max_nnz_each_col", "nz_rows",
    "nz_cols", "diag_dominant_rows", "t_num_blocks", "t_num_RB", "t_num_CB", "t_mod_RB", "t_mod_CB",
    "t_min_nnz_all_tiles", "t_max_nnz_all_tiles", "t_min_nnz_each_RB", "t_max_nnz_each_RB", "t_min_nnz_each_CB",
    "t_max_nnz_each_CB", "uniq_rows_sum", "uniq_cols_sum", "symm_pattern_pairs", "symm_value_pairs",
    "is_symmetric"]
FEATURE_DOUBLE_FIELDS = [
    "pattern_symm", "value_symm", "nnz_ratio", "diag_close_ratio", "distance_per_row", "nz_row_ratio",
    "ave_nnz_each_row", "var_nnz_each_row", "standard_dev_row", "Gini_row", "P_ratio_row", "nz_col_ratio",
    "ave_nnz_each_col", "var_nnz_each_col", "standard_dev_col", "Gini_col", "P_ratio_col",
    "diagonal_dominant_ratio", "max_value_offdiag", "max_value_diagonal", "row_variability", "col_variability",
    "t_ave_nnz_all_tiles", "t_ave_nnz_RB", "t_ave_nnz_CB", "t_var_nnz_all_tiles", "t_var_nnz_RB", "t_var_nnz_CB",
    "t_standard_dev_all_tiles", "t_standard_dev_RB", "t_standard_dev_CB", "t_Gini_all_tiles", "t_Gini_RB",
    "t_Gini_CB", "t_P_ratio_all_tiles", "t_P_ratio_RB", "t_P_ratio_CB", "t_nz_ratio_tiles", "t_nz_ratio_RB",
    "t_nz_ratio_CB", "uniqR", "uniqC", "potReuseR", "potReuseC", "GrX_uniqR", "GrX_uniqC", "GrX_potReuseR",
    "GrX_potReuseC", "device_ms"]
FEAT_SYMMETRIC_FILE = 1
FEAT_TILES = 256
NUM_FEATURES = 40


class TuneResultC(C.Structure):
    _fields_ = [("fmt", C.c_int32), ("variant", C.c_int32), ("params", Params), ("status", C.c_int32),
                ("ms", C.c_double), ("convert_ms", C.c_double)]

    def as_dict(self):
        return {"fmt": self.fmt, "variant": self.variant, "params": self.params.as_dict(), "status": self.status,
                "ms": self.ms, "convert_ms": self.convert_ms}


class ServerStatsC(C.Structure):
    _fields_ = [("resident", C.c_int32), ("queued", C.c_int32), ("bytes_used", C.c_int64),
                ("bytes_budget", C.c_int64), ("batches", C.c_int64), ("items", C.c_int64), ("completed", C.c_int64)]


class FeaturesC(C.Structure):
    _fields_ = [(n, C.c_int64) for n in FEATURE_INT_FIELDS] + [(n, C.c_double) for n in FEATURE_DOUBLE_FIELDS]


# lespmv_feature_images_t: (name, element type, length)
FEATURE_IMAGES = [("tile_nnz", np.int32, 65536), ("tile_max", np.int32, 65536), ("tile_ave", np.float64, 65536),
                  ("tile_std", np.float64, 65536), ("tile_uniq_rows", np.int32, 65536),
                  ("tile_uniq_cols", np.int32, 65536), ("rb_nnz", np.int32, 256), ("rb_max", np.int32, 256),
                  ("rb_ave", np.float64, 256), ("rb_std", np.float64, 256), ("cb_nnz", np.int32, 256),
                  ("cb_max", np.int32, 256), ("cb_ave", np.float64, 256), ("cb_std", np.float64, 256)]


class FeatureImagesC(C.Structure):
    _fields_ = [(n, C.c_void_p) for n, _, _ in FEATURE_IMAGES]


_lib = None


def load() -> C.CDLL:
    """Load libsspmv_hip.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    sig = {
        "lespmv_last_error": ([], C.c_char_p),
        "lespmv_init": ([C.c_int], C.c_int),
        "lespmv_device_info": ([C.POINTER(i64), C.POINTER(i32), C.POINTER(i32)], C.c_int),
        "lespmv_default_params": ([C.c_int, C.c_int, C.c_int, C.POINTER(Params)], C.c_int),
        "lespmv_create_csr": ([C.POINTER(vp), C.c_int, i64, i64, i64, vp, vp, vp, C.c_int], C.c_int),
        "lespmv_destroy": ([vp], C.c_int),
        "lespmv_matrix_info": ([vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64), C.POINTER(i32)], C.c_int),
        "lespmv_convert": ([vp, C.c_int, C.POINTER(Params)], C.c_int),
        "lespmv_release_format": ([vp, C.c_int], C.c_int),
        "lespmv_has_format": ([vp, C.c_int], C.c_int),
        "lespmv_convert_ms": ([vp, C.c_int, C.POINTER(dbl)], C.c_int),
        "lespmv_prepare": ([vp, C.c_int, C.c_int], C.c_int),
        "lespmv_csr_plan": ([vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(dbl)], C.c_int),
        "lespmv_csr_spmv_noalloc": ([C.c_int, i64, i64, i64, vp, vp, vp, dbl, vp, dbl, vp, vp], C.c_int),
        "lespmv_csr_noalloc_max_items": ([], i64),
        "lespmv_spmv": ([vp, C.c_int, C.c_int, dbl, vp, dbl, vp, vp], C.c_int),
        "lespmv_spmm": ([vp, C.c_int, C.c_int, i32, dbl, vp, i64, dbl, vp, i64, vp], C.c_int),
        "lespmv_time_spmv": ([vp, C.c_int, C.c_int, dbl, vp, dbl, vp, i32, C.POINTER(dbl)], C.c_int),
        "lespmv_time_spmm": ([vp, C.c_int, C.c_int, i32, dbl, vp, i64, dbl, vp, i64, i32, C.POINTER(dbl)],
                             C.c_int),
        "lespmv_time_spmv_rotating": ([C.POINTER(vp), i32, C.c_int, C.c_int, dbl, C.POINTER(vp), dbl,
                                       C.POINTER(vp), i32, C.POINTER(dbl)], C.c_int),
        "lespmv_effective_bytes": ([vp, C.c_int, i32, C.c_int, C.POINTER(i64)], C.c_int),
        "lespmv_reference_bytes": ([vp, C.c_int, C.POINTER(i64)], C.c_int),
        "lespmv_format_bytes": ([vp, C.c_int, C.POINTER(i64)], C.c_int),
        "lespmv_export_size": ([vp, C.c_int, C.c_char_p, C.POINTER(i64)], C.c_int),
        "lespmv_export": ([vp, C.c_int, C.c_char_p, vp, i64], C.c_int),
        "lespmv_malloc": ([C.POINTER(vp), i64], C.c_int),
        "lespmv_free": ([vp], C.c_int),
        "lespmv_memcpy_h2d": ([vp, vp, i64], C.c_int),
        "lespmv_memcpy_d2h": ([vp, vp, i64], C.c_int),
        "lespmv_memset": ([vp, C.c_int, i64], C.c_int),
        "lespmv_synchronize": ([], C.c_int),
        "lespmv_time_copy": ([vp, vp, i64, i32, C.POINTER(dbl)], C.c_int),
        "lespmv_stream_create": ([C.POINTER(vp)], C.c_int),
        "lespmv_stream_destroy": ([vp], C.c_int),
        "lespmv_stream_synchronize": ([vp], C.c_int),
        "lespmv_generate": ([C.c_int, C.POINTER(dbl), i32, C.c_uint64, C.c_int, C.POINTER(HostCsrC)], C.c_int),
        "lespmv_host_csr_free": ([C.POINTER(HostCsrC)], C.c_int),
        "lespmv_fill_uniform": ([vp, C.c_int, i64, C.c_uint64, dbl, dbl], C.c_int),
        "lespmv_read_mtx": ([C.c_char_p, C.c_int, C.c_int, C.POINTER(HostCsrC)], C.c_int),
        "lespmv_coo_to_csr": ([C.c_int, i64, i64, i64, vp, vp, vp, C.c_int, C.POINTER(HostCsrC)], C.c_int),
        "lespmv_features": ([vp, i32, C.POINTER(FeaturesC), C.POINTER(FeatureImagesC)], C.c_int),
        "lespmv_tune_candidates": ([i32, i32, C.POINTER(TuneResultC), i32, C.POINTER(i32)], C.c_int),
        "lespmv_tune": ([vp, C.POINTER(TuneResultC), i32, i32, i32], C.c_int),
        "lespmv_autotune": ([vp, i32, i32, i32, C.POINTER(TuneResultC)], C.c_int),
        "lespmv_block_counts": ([vp, vp, i32, vp], C.c_int),
        "lespmv_mlp_create": ([C.POINTER(vp), i32, i32, i32, i32, vp, C.c_uint64], C.c_int),
        "lespmv_mlp_destroy": ([vp], C.c_int),
        "lespmv_mlp_shape": ([vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64)], C.c_int),
        "lespmv_mlp_get_params": ([vp, vp], C.c_int),
        "lespmv_mlp_set_params": ([vp, vp], C.c_int),
        "lespmv_mlp_set_normalizer": ([vp, vp, vp], C.c_int),
        "lespmv_mlp_loss_grad": ([vp, vp, vp, i32, C.POINTER(dbl), vp], C.c_int),
        "lespmv_mlp_train": ([vp, vp, vp, i32, i32, i32, C.c_float, C.c_float, C.c_uint64, vp], C.c_int),
        "lespmv_mlp_forward": ([vp, vp, i32, vp], C.c_int),
        "lespmv_mlp_save": ([vp, C.c_char_p], C.c_int),
        "lespmv_mlp_load": ([C.POINTER(vp), C.c_char_p], C.c_int),
        "lespmv_selector_layout": ([vp, vp, vp], C.c_int),
        "lespmv_selector_inputs": ([vp, i32, vp, vp, C.POINTER(FeaturesC)], C.c_int),
        "lespmv_selector_decode": ([vp, vp, C.POINTER(i32), C.POINTER(i32)], C.c_int),
        "lespmv_select": ([vp, vp, i32, C.POINTER(TuneResultC), C.POINTER(i32)], C.c_int),
        "lespmv_create_csr64": ([C.POINTER(vp), C.c_int, i64, i64, i64, vp, vp, vp, C.c_int], C.c_int),
        "lespmv_server_create": ([C.POINTER(vp), i32, i64, i32, i32], C.c_int),
        "lespmv_server_destroy": ([vp], C.c_int),
        "lespmv_server_add_csr": ([vp, i64, i64, i64, vp, vp, vp, i32, C.POINTER(i32)], C.c_int),
        "lespmv_server_remove": ([vp, i32], C.c_int),
        "lespmv_server_matrix": ([vp, i32, C.POINTER(vp), C.POINTER(i32)], C.c_int),
        "lespmv_server_lanes": ([vp, i32, i32, C.POINTER(i32)], C.c_int),
        "lespmv_server_use": ([vp, i32, i32, vp, i32], C.c_int),
        "lespmv_server_format": ([vp, i32, i32, C.POINTER(i32)], C.c_int),
        "lespmv_server_submit": ([vp, i32, i32, dbl, vp, i64, dbl, vp, i64, C.POINTER(i64)], C.c_int),
        "lespmv_server_poll": ([vp, i32, vp, vp, i32, C.POINTER(i32)], C.c_int),
        "lespmv_server_drain": ([vp], C.c_int),
        "lespmv_server_stats": ([vp, C.POINTER(ServerStatsC)], C.c_int),
        "lespmv_features_vector": ([C.POINTER(FeaturesC), C.POINTER(dbl)], C.c_int),
        "lespmv_feature_name": ([i32], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc: int):
    if rc != 0:
        raise LespmvError(rc, load().lespmv_last_error().decode(errors="replace"))


def dtype_id(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.float64:
        return F64
    if dt == np.float32:
        return F32
    raise TypeError(f"unsupported dtype {dt}")


def np_dtype(did: int):
    return np.float64 if did == F64 else np.float32


def default_params(fmt: int, dtype, reference_layout: bool = False) -> Params:
    p = Params()
    check(load().lespmv_default_params(fmt, dtype_id(dtype), int(reference_layout), C.byref(p)))
    return p


def make_params(**kw) -> Params:
    p = Params()
    for k, v in kw.items():
        setattr(p, k, int(v))
    return p


class DeviceArray:
    """A device buffer owned by the library's allocator."""

    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        check(load().lespmv_malloc(C.byref(self.ptr), self.nbytes))

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.dtype, d.shape = a.dtype, a.shape
        if a.nbytes:
            check(load().lespmv_memcpy_h2d(d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return d

    @classmethod
    def empty(cls, shape, dtype) -> "DeviceArray":
        dt = np.dtype(dtype)
        n = int(np.prod(shape))
        d = cls(n * dt.itemsize)
        d.dtype, d.shape = dt, tuple(np.atleast_1d(shape))
        return d

    def fill_bytes(self, value: int = 0):
        check(load().lespmv_memset(self.ptr, value, self.nbytes))

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        check(load().lespmv_memcpy_h2d(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if out.nbytes:
            check(load().lespmv_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def __del__(self):
        try:
            if self.ptr:
                load().lespmv_free(self.ptr)
                self.ptr = C.c_void_p()
        except Exception:
            pass


@dataclass
class HostCsr:
    m: int
    n: int
    rp: np.ndarray
    ci: np.ndarray
    v: np.ndarray
    symmetric: bool = False  # read from a symmetric MatrixMarket file

    @property
    def nnz(self) -> int:
        return int(self.ci.shape[0])


def _take_host_csr(out: HostCsrC) -> HostCsr:
    lib = load()
    try:
        m, nnz = out.m, out.nnz
        rp = np.ctypeslib.as_array(out.row_offset, shape=(m + 1,)).copy()
        ci = np.ctypeslib.as_array(out.col_index, shape=(max(nnz, 1),))[:nnz].copy()
        vt = C.c_double if out.dtype == F64 else C.c_float
        v = np.ctypeslib.as_array(C.cast(out.values, C.POINTER(vt)), shape=(max(nnz, 1),))[:nnz].copy()
    finally:
        lib.lespmv_host_csr_free(C.byref(out))
    return HostCsr(int(out.m), int(out.n), rp, ci, v, bool(out.mm_flags & 1))


def generate(family: int, params, seed: int, dtype=np.float64) -> HostCsr:
    """Synthetic CSR from the library's counter-based generators (host)."""
    arr = (C.c_double * max(len(params), 1))(*params)
    out = HostCsrC()
    check(load().lespmv_generate(family, arr, len(params), seed, dtype_id(dtype), C.byref(out)))
    return _take_host_csr(out)


def read_mtx(path: str, dtype=np.float64, compact: bool = False) -> HostCsr:
    """read_csr_matrix semantics (symmetric expansion, pattern -> 1.0)."""
    out = HostCsrC()
    check(load().lespmv_read_mtx(path.encode(), dtype_id(dtype), int(compact), C.byref(out)))
    return _take_host_csr(out)


def coo_to_csr(m: int, n: int, ri, ci, v, compact: bool = False) -> HostCsr:
    """coo_to_csr; without compaction it runs as a GPU radix sort."""
    ri = np.ascontiguousarray(ri, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    v = np.ascontiguousarray(v)
    out = HostCsrC()
    check(load().lespmv_coo_to_csr(dtype_id(v.dtype), m, n, len(ci), ri.ctypes.data_as(C.c_void_p),
                                   ci.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p), int(compact),
                                   C.byref(out)))
    return _take_host_csr(out)


def fill_uniform(n: int, seed: int, dtype=np.float64, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """x[i] = lo + (hi-lo)*u(splitmix64(seed ^ i))."""
    out = np.empty(n, dtype=dtype)
    check(load().lespmv_fill_uniform(out.ctypes.data_as(C.c_void_p), dtype_id(dtype), n, seed, lo, hi))
    return out


class Matrix:
    """A device-resident sparse matrix handle (CSR plus any converted formats)."""

    def __init__(self, m: int, n: int, rp: np.ndarray, ci: np.ndarray, v: np.ndarray):
        lib = load()
        self.h = C.c_void_p()
        rp = np.ascontiguousarray(rp, dtype=np.int32)
        ci = np.ascontiguousarray(ci, dtype=np.int32)
        v = np.ascontiguousarray(v)
        self.dtype = v.dtype
        self.m, self.n, self.nnz = int(m), int(n), int(ci.shape[0])
        check(lib.lespmv_create_csr(C.byref(self.h), dtype_id(v.dtype), self.m, self.n, self.nnz,
                                    rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                                    v.ctypes.data_as(C.c_void_p), 0))

    @classmethod
    def from_host(cls, A) -> "Matrix":
        return cls(A.m, A.n, A.rp, A.ci, A.v)

    @classmethod
    def csr64(cls, m: int, n: int, rp: np.ndarray, ci: np.ndarray, v: np.ndarray) -> "Matrix":
        """A handle with 64-bit row offsets (lespmv_create_csr64): CSR SpMV / SpMM only."""
        M = cls.__new__(cls)
        M.h = C.c_void_p()
        rp = np.ascontiguousarray(rp, dtype=np.int64)
        ci = np.ascontiguousarray(ci, dtype=np.int32)
        v = np.ascontiguousarray(v)
        M.dtype = v.dtype
        M.m, M.n, M.nnz = int(m), int(n), int(ci.shape[0])
        check(load().lespmv_create_csr64(C.byref(M.h), dtype_id(v.dtype), M.m, M.n, M.nnz,
                                         rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                                         v.ctypes.data_as(C.c_void_p), 0))
        return M

    @classmethod
    def borrowed(cls, h, m: int, n: int, nnz: int, dtype, owner=None) -> "Matrix":
        """A handle owned elsewhere (a server's resident matrix); not destroyed here."""
        M = cls.__new__(cls)
        M.h, M.dtype, M.m, M.n, M.nnz = h, np.dtype(dtype), int(m), int(n), int(nnz)
        M._owner = owner  # keeps the owner alive while the handle is in use
        return M

    def __del__(self):
        try:
            if self.h and getattr(self, "_owner", None) is None:
                load().lespmv_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass

    def convert(self, fmt: int, params: Params | None = None):
        check(load().lespmv_convert(self.h, fmt, C.byref(params) if params is not None else None))
        return self

    def has(self, fmt: int) -> bool:
        return bool(load().lespmv_has_format(self.h, fmt))

    def release(self, fmt: int):
        check(load().lespmv_release_format(self.h, fmt))

    def prepare(self, fmt: int, variant: int = 0) -> float:
        """Builds what `variant` needs ahead of the first call (CSR: the merge-path
        partition, the column-panel layout); returns its device ms."""
        check(load().lespmv_prepare(self.h, fmt, variant))
        return self.convert_ms(fmt)

    def csr_plan(self):
        """(variant AUTO runs for CSR, vector lanes per row, far x lines per entry)."""
        v, lanes, far = C.c_int32(), C.c_int32(), C.c_double()
        check(load().lespmv_csr_plan(self.h, C.byref(v), C.byref(lanes), C.byref(far)))
        return v.value, lanes.value, far.value

    def convert_ms(self, fmt: int) -> float:
        ms = C.c_double()
        check(load().lespmv_convert_ms(self.h, fmt, C.byref(ms)))
        return ms.value

    def spmv(self, fmt: int, alpha: float, x: DeviceArray, beta: float, y: DeviceArray, variant: int = 0,
             stream=None):
        check(load().lespmv_spmv(self.h, fmt, variant, alpha, x.ptr, beta, y.ptr, stream))

    def spmm(self, fmt: int, k: int, alpha: float, X: DeviceArray, ldx: int, beta: float, Y: DeviceArray, ldy: int,
             variant: int = 0, stream=None):
        check(load().lespmv_spmm(self.h, fmt, variant, k, alpha, X.ptr, ldx, beta, Y.ptr, ldy, stream))

    def time_spmv(self, fmt, alpha, x, beta, y, iters, variant=0) -> float:
        ms = C.c_double()
        check(load().lespmv_time_spmv(self.h, fmt, variant, alpha, x.ptr, beta, y.ptr, iters, C.byref(ms)))
        return ms.value

    def time_spmm(self, fmt, k, alpha, X, ldx, beta, Y, ldy, iters, variant=0) -> float:
        ms = C.c_double()
        check(load().lespmv_time_spmm(self.h, fmt, variant, k, alpha, X.ptr, ldx, beta, Y.ptr, ldy, iters,
                                      C.byref(ms)))
        return ms.value

    def effective_bytes(self, fmt: int, k: int = 1, beta_nonzero: bool = False) -> int:
        b = C.c_int64()
        check(load().lespmv_effective_bytes(self.h, fmt, k, int(beta_nonzero), C.byref(b)))
        return b.value

    def reference_bytes(self, fmt: int) -> int:
        b = C.c_int64()
        check(load().lespmv_reference_bytes(self.h, fmt, C.byref(b)))
        return b.value

    def format_bytes(self, fmt: int) -> int:
        b = C.c_int64()
        check(load().lespmv_format_bytes(self.h, fmt, C.byref(b)))
        return b.value

    def export(self, fmt: int, name: str, dtype) -> np.ndarray:
        lib = load()
        nb = C.c_int64()
        check(lib.lespmv_export_size(self.h, fmt, name.encode(), C.byref(nb)))
        out = np.empty(nb.value // np.dtype(dtype).itemsize, dtype=dtype)
        check(lib.lespmv_export(self.h, fmt, name.encode(), out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def features(self, symmetric: bool = False, images: bool = False):
        return features(self, symmetric, images)

    def tune(self, candidates=None, iters: int = 10, k: int = 1):
        """Time every candidate (converted, timed and released in turn)."""
        if candidates is None:
            candidates = tune_candidates(ALL_FORMATS_MASK, self.dtype)
        check(load().lespmv_tune(self.h, candidates, len(candidates), iters, k))
        return candidates

    def autotune(self, fmt: int, iters: int = 10, k: int = 1) -> TuneResultC:
        """Best parameters of one format; the winner stays converted."""
        best = TuneResultC()
        check(load().lespmv_autotune(self.h, fmt, iters, k, C.byref(best)))
        return best

    def export_scalar(self, fmt: int, name: str) -> int:
        return int(self.export(fmt, name, np.int64)[0])


ALL_FORMATS_MASK = (1 << 9) - 1
TUNE_PANELS = 1 << 9  # adds the CSR column-panel candidates after all others (LESPMV_TUNE_PANELS)


def tune_candidates(fmt_mask: int = ALL_FORMATS_MASK, dtype=np.float64):
    """The library's standard candidate grid (a ctypes array of TuneResultC)."""
    lib = load()
    n = C.c_int32()
    check(lib.lespmv_tune_candidates(fmt_mask, dtype_id(dtype), None, 0, C.byref(n)))
    arr = (TuneResultC * n.value)()
    check(lib.lespmv_tune_candidates(fmt_mask, dtype_id(dtype), arr, n.value, C.byref(n)))
    return arr


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.c_void_p)


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.c_void_p)


class Mlp:
    """The GPU-trained MLP (lespmv_mlp_*): in -> h1 -> h2 -> heads."""

    def __init__(self, in_dim=None, h1=128, h2=128, head_sizes=(), seed=0, path=None):
        lib = load()
        self.h = C.c_void_p()
        if path is not None:
            check(lib.lespmv_mlp_load(C.byref(self.h), path.encode()))
        else:
            hs, hp = _i32(head_sizes)
            check(lib.lespmv_mlp_create(C.byref(self.h), in_dim, h1, h2, len(hs), hp, seed))
        i, o, n = C.c_int32(), C.c_int32(), C.c_int64()
        check(lib.lespmv_mlp_shape(self.h, C.byref(i), C.byref(o), C.byref(n)))
        self.in_dim, self.out_dim, self.nparams = i.value, o.value, n.value

    def __del__(self):
        try:
            if self.h:
                load().lespmv_mlp_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparams, dtype=np.float32)
        check(load().lespmv_mlp_get_params(self.h, p.ctypes.data_as(C.c_void_p)))
        return p

    def set_params(self, p):
        p, pp = _f32(p)
        check(load().lespmv_mlp_set_params(self.h, pp))

    def set_normalizer(self, mean, std):
        m, mp = _f32(mean)
        s, sp = _f32(std)
        check(load().lespmv_mlp_set_normalizer(self.h, mp, sp))

    def loss_grad(self, X, labels):
        X, xp = _f32(X)
        L, lp = _i32(labels)
        loss = C.c_double()
        g = np.zeros(self.nparams, dtype=np.float32)
        check(load().lespmv_mlp_loss_grad(self.h, xp, lp, X.shape[0], C.byref(loss), g.ctypes.data_as(C.c_void_p)))
        return loss.value, g

    def train(self, X, labels, epochs, batch=256, lr=1e-3, weight_decay=0.0, seed=0):
        X, xp = _f32(X)
        L, lp = _i32(labels)
        hist = np.zeros(epochs, dtype=np.float32)
        check(load().lespmv_mlp_train(self.h, xp, lp, X.shape[0], epochs, batch, lr, weight_decay, seed,
                                      hist.ctypes.data_as(C.c_void_p)))
        return hist

    def forward(self, X) -> np.ndarray:
        X, xp = _f32(X)
        out = np.zeros((X.shape[0], self.out_dim), dtype=np.float32)
        check(load().lespmv_mlp_forward(self.h, xp, X.shape[0], out.ctypes.data_as(C.c_void_p)))
        return out

    def save(self, path: str):
        check(load().lespmv_mlp_save(self.h, path.encode()))


def selector_layout():
    cf = np.zeros(10, dtype=np.int32)
    hc = np.zeros(9, dtype=np.int32)
    hs = np.zeros(7, dtype=np.int32)
    check(load().lespmv_selector_layout(cf.ctypes.data_as(C.c_void_p), hc.ctypes.data_as(C.c_void_p),
                                        hs.ctypes.data_as(C.c_void_p)))
    return cf, hc, hs


SELECTOR_OUTPUTS = 35  # LESPMV_SELECTOR_OUTPUTS


def selector_inputs(M: "Matrix", symmetric: bool = False):
    """(128 float inputs, 35 allowed flags, features dict) of a handle."""
    x = np.zeros(128, dtype=np.float32)
    allowed = np.zeros(SELECTOR_OUTPUTS, dtype=np.uint8)
    f = FeaturesC()
    check(load().lespmv_selector_inputs(M.h, FEAT_SYMMETRIC_FILE if symmetric else 0, x.ctypes.data_as(C.c_void_p),
                                        allowed.ctypes.data_as(C.c_void_p), C.byref(f)))
    return x, allowed, {n: getattr(f, n) for n, _ in FeaturesC._fields_}


def selector_decode(probs, allowed):
    p, pp = _f32(probs)
    a = np.ascontiguousarray(allowed, dtype=np.uint8)
    cls, cand = C.c_int32(), C.c_int32()
    check(load().lespmv_selector_decode(pp, a.ctypes.data_as(C.c_void_p), C.byref(cls), C.byref(cand)))
    return cls.value, cand.value


def select(M: "Matrix", mlp: Mlp, symmetric: bool = False):
    """(chosen TuneResultC, format class) for a handle."""
    choice = TuneResultC()
    cls = C.c_int32()
    check(load().lespmv_select(M.h, mlp.h, FEAT_SYMMETRIC_FILE if symmetric else 0, C.byref(choice), C.byref(cls)))
    return choice, cls.value


def block_counts(M: "Matrix", dims) -> np.ndarray:
    d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32).reshape(-1))
    out = np.zeros(len(d) // 2, dtype=np.int64)
    check(load().lespmv_block_counts(M.h, d.ctypes.data_as(C.c_void_p), len(out), out.ctypes.data_as(C.c_void_p)))
    return out


class Server:
    """Continuous-batching SpMV / SpMM server (lespmv_server_*)."""

    def __init__(self, dtype=np.float64, budget_bytes: int = 1 << 34, max_batch: int = 128, slots: int = 2):
        self.h = C.c_void_p()
        self.dtype = np.dtype(dtype)
        check(load().lespmv_server_create(C.byref(self.h), dtype_id(dtype), int(budget_bytes), max_batch, slots))

    def __del__(self):
        try:
            if self.h:
                load().lespmv_server_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass

    def add(self, A, lanes: int = 0) -> int:
        rp = np.ascontiguousarray(A.rp, dtype=np.int32)
        ci = np.ascontiguousarray(A.ci, dtype=np.int32)
        v = np.ascontiguousarray(A.v, dtype=self.dtype)
        mid = C.c_int32()
        check(load().lespmv_server_add_csr(self.h, A.m, A.n, len(ci), rp.ctypes.data_as(C.c_void_p),
                                           ci.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p), lanes,
                                           C.byref(mid)))
        return mid.value

    def remove(self, mid: int):
        check(load().lespmv_server_remove(self.h, mid))

    def lanes(self, mid: int, k: int = 1) -> int:
        """Lanes per row the server uses for this matrix's requests with k right-hand sides."""
        lanes = C.c_int32()
        check(load().lespmv_server_lanes(self.h, mid, k, C.byref(lanes)))
        return lanes.value

    def use(self, mid: int, fmt: int, params: "Params | None" = None, k_from: int = 0):
        """Serve the matrix from a layout: a sliced format for k = 1, BSR 4x4 for k >= k_from, or CSR only."""
        check(load().lespmv_server_use(self.h, mid, fmt, C.byref(params) if params is not None else None, k_from))

    def format(self, mid: int, k: int = 1) -> int:
        """The format this matrix's requests with k right-hand sides run on."""
        f = C.c_int32()
        check(load().lespmv_server_format(self.h, mid, k, C.byref(f)))
        return f.value

    def matrix(self, mid: int, A) -> "Matrix":
        """The server's handle of resident matrix `mid` (host matrix A gives its shape), borrowed."""
        h = C.c_void_p()
        check(load().lespmv_server_matrix(self.h, mid, C.byref(h), None))
        return Matrix.borrowed(h, A.m, A.n, len(A.ci), self.dtype, owner=self)

    def submit(self, mid: int, k: int, alpha: float, X: "DeviceArray", ldx: int, beta: float, Y: "DeviceArray",
               ldy: int) -> int:
        rid = C.c_int64()
        check(load().lespmv_server_submit(self.h, mid, k, alpha, X.ptr, ldx, beta, Y.ptr, ldy, C.byref(rid)))
        return rid.value

    def poll(self, wait: bool = False, cap: int = 4096):
        ids = np.zeros(cap, dtype=np.int64)
        lat = np.zeros(cap, dtype=np.float64)
        n = C.c_int32()
        check(load().lespmv_server_poll(self.h, int(wait), ids.ctypes.data_as(C.c_void_p),
                                        lat.ctypes.data_as(C.c_void_p), cap, C.byref(n)))
        return ids[:n.value], lat[:n.value]

    def drain(self):
        check(load().lespmv_server_drain(self.h))

    def stats(self) -> dict:
        s = ServerStatsC()
        check(load().lespmv_server_stats(self.h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in ServerStatsC._fields_}


def feature_names() -> list:
    """The 40 line names of a .features file, in file order."""
    lib = load()
    return [lib.lespmv_feature_name(i).decode() for i in range(NUM_FEATURES)]


def features(M: "Matrix", symmetric: bool = False, images: bool = False):
    """GPU features of a handle: (dict of lespmv_features_t fields, 40-vector,
    dict of images or None)."""
    lib = load()
    f = FeaturesC()
    img = None
    imgc = None
    if images:
        img = {n: np.zeros(k, dtype=t) for n, t, k in FEATURE_IMAGES}
        imgc = FeatureImagesC(*[img[n].ctypes.data_as(C.c_void_p) for n, _, _ in FEATURE_IMAGES])
    check(lib.lespmv_features(M.h, FEAT_SYMMETRIC_FILE if symmetric else 0, C.byref(f),
                              C.byref(imgc) if imgc is not None else None))
    vec = np.zeros(NUM_FEATURES)
    check(lib.lespmv_features_vector(C.byref(f), vec.ctypes.data_as(C.POINTER(C.c_double))))
    vals = {n: getattr(f, n) for n, _ in FeaturesC._fields_}
    return vals, vec, img


def time_spmv_rotating(mats, fmt: int, alpha: float, xs, beta: float, ys, iters: int, variant: int = 0) -> float:
    """Device ms per SpMV over `iters` rounds of one call per matrix in turn
    (each matrix with its own x and y), on the first matrix's stream."""
    n = len(mats)
    H = (C.c_void_p * n)(*[M.h.value for M in mats])
    X = (C.c_void_p * n)(*[x.ptr.value for x in xs])
    Y = (C.c_void_p * n)(*[y.ptr.value for y in ys])
    ms = C.c_double()
    check(load().lespmv_time_spmv_rotating(H, n, fmt, variant, alpha, X, beta, Y, iters, C.byref(ms)))
    return ms.value


def time_copy(dst: DeviceArray, src: DeviceArray, nbytes: int, iters: int) -> float:
    ms = C.c_double()
    check(load().lespmv_time_copy(dst.ptr, src.ptr, nbytes, iters, C.byref(ms)))
    return ms.value


def synchronize():
    check(load().lespmv_synchronize())


def device_info() -> dict:
    hbm, cu, arch = C.c_int64(), C.c_int32(), C.c_int32()
    check(load().lespmv_device_info(C.byref(hbm), C.byref(cu), C.byref(arch)))
    return {"hbm_bytes": hbm.value, "num_cu": cu.value, "gfx": arch.value}
