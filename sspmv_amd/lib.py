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
VARIANT_AUTO, VARIANT_VECTOR, VARIANT_MERGE, VARIANT_ATOMIC = 0, 1, 2, 3
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
        "lespmv_spmv": ([vp, C.c_int, C.c_int, dbl, vp, dbl, vp, vp], C.c_int),
        "lespmv_spmm": ([vp, C.c_int, C.c_int, i32, dbl, vp, i64, dbl, vp, i64, vp], C.c_int),
        "lespmv_time_spmv": ([vp, C.c_int, C.c_int, dbl, vp, dbl, vp, i32, C.POINTER(dbl)], C.c_int),
        "lespmv_time_spmm": ([vp, C.c_int, C.c_int, i32, dbl, vp, i64, dbl, vp, i64, i32, C.POINTER(dbl)],
                             C.c_int),
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

    def __del__(self):
        try:
            if self.h:
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

    def export_scalar(self, fmt: int, name: str) -> int:
        return int(self.export(fmt, name, np.int64)[0])


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
