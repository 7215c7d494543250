"""Matrix features and their files, as Get_features / MTX::FeaturesWrite make
them (src/sparse_features.cpp:1134-1410) and SmartAdpter reads them
(SmartAdpter/utils/load_MMdata.py:18-83, 123-152).

The values come from the GPU (``lib.features``); this module only lays them out.
Files of one matrix go to ``<root>/<name>/``:
  <name>.features       "[matID]: name" then 40 "label value" lines
  <name>.nnz/.max/.ave/.std      65,536 per-tile lines (row-major tiles)
  <name>.RB{nnz,max,ave,std}     256 per-row-block lines
  <name>.CB{nnz,max,ave,std}     256 per-column-block lines
The reference writes under a fixed root (/data/lsl/MModel-Data); here the
root is an argument. Numbers are printed as C++ iostreams print them by
default: integers as integers, doubles as %g with six significant digits.
"""
from __future__ import annotations

import os

import numpy as np

from . import lib

# .features lines whose value is an integer (IndexType) in the reference
INTEGER_LINES = {"num_rows", "num_cols", "num_nnz", "max_nnz_each_row", "max_nnz_each_col", "max_nnz_each_tiles",
                 "max_nnz_each_RB", "max_nnz_each_CB"}

IMAGE_FILES = [("ave", "tile_ave"), ("max", "tile_max"), ("nnz", "tile_nnz"), ("std", "tile_std"),
               ("RBmax", "rb_max"), ("RBnnz", "rb_nnz"), ("RBave", "rb_ave"), ("RBstd", "rb_std"),
               ("CBmax", "cb_max"), ("CBnnz", "cb_nnz"), ("CBave", "cb_ave"), ("CBstd", "cb_std")]


def fmt_value(v, integer: bool) -> str:
    if integer:
        return str(int(v))
    return "%g" % float(v)


def features_text(mat_id: int, name: str, vec: np.ndarray) -> str:
    labels = lib.feature_names()
    lines = [f"[{mat_id}]: {name}"]
    lines += [f"{lab} {fmt_value(v, lab in INTEGER_LINES)}" for lab, v in zip(labels, vec)]
    return "\n".join(lines) + "\n"


def write_features(root: str, mat_id: int, name: str, vec: np.ndarray, images: dict | None = None) -> str:
    """Write <root>/<name>/<name>.features (and the image files when given);
    returns the directory."""
    d = os.path.join(root, name)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, name + ".features"), "w") as f:
        f.write(features_text(mat_id, name, vec))
    if images is not None:
        for suffix, key in IMAGE_FILES:
            arr = images[key]
            integer = arr.dtype.kind in "iu"
            with open(os.path.join(d, f"{name}.{suffix}"), "w") as f:
                f.write("".join(fmt_value(v, integer) + "\n" for v in arr.tolist()))
    return d


def read_features(path: str) -> tuple[str, np.ndarray]:
    """The header line and the 40 values of a .features file (the loader's view,
    load_MMdata.py:123-152)."""
    with open(path) as f:
        header = f.readline().rstrip("\n")
        vals = [float(line.split()[1]) for line in f if line.strip()]
    return header, np.array(vals, dtype=np.float64)


def extract(path: str, mat_id: int = 0, dtype=np.float64):
    """Get_features on a MatrixMarket file: read it, build the device CSR, and
    compute every feature on the GPU. Returns (fields, 40-vector, images)."""
    A = lib.read_mtx(path, dtype)
    M = lib.Matrix.from_host(A)
    return lib.features(M, symmetric=A.symmetric, images=True)
