"""Small deterministic CSR test matrices (numpy only, no GPU, no product code).

These cover the shapes the reference's conversions and kernels branch on:
empty rows, rows longer than a CSR5 tile, rectangular matrices, unsorted
columns inside a row, duplicate entries (for coo_to_csr compact).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class HostCsr:
    m: int
    n: int
    rp: np.ndarray
    ci: np.ndarray
    v: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.ci.shape[0])

    def astype(self, dt) -> "HostCsr":
        return HostCsr(self.m, self.n, self.rp, self.ci, self.v.astype(dt))

    def dense(self) -> np.ndarray:
        a = np.zeros((self.m, self.n), dtype=np.float64)
        rows = np.repeat(np.arange(self.m), np.diff(self.rp))
        np.add.at(a, (rows, self.ci), self.v.astype(np.float64))
        return a


def from_rows(m: int, n: int, rows: list, rng, dtype=np.float64) -> HostCsr:
    rp = np.zeros(m + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate([np.asarray(r, dtype=np.int32) for r in rows]) if rp[-1] else np.zeros(0, np.int32)
    v = rng.uniform(-1.0, 1.0, size=int(rp[-1])).astype(dtype)
    return HostCsr(m, n, rp, ci.astype(np.int32), v)


def laplace2d(nx: int, ny: int, dtype=np.float64) -> HostCsr:
    """5-point Laplacian, natural order, diag 4, off-diagonals -1."""
    rows, vals = [], []
    for j in range(ny):
        for i in range(nx):
            r = j * nx + i
            cols = []
            if j > 0:
                cols.append((r - nx, -1.0))
            if i > 0:
                cols.append((r - 1, -1.0))
            cols.append((r, 4.0))
            if i < nx - 1:
                cols.append((r + 1, -1.0))
            if j < ny - 1:
                cols.append((r + nx, -1.0))
            rows.append([c for c, _ in cols])
            vals.extend(v for _, v in cols)
    m = nx * ny
    rp = np.zeros(m + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    return HostCsr(m, m, rp, np.concatenate([np.array(r, np.int32) for r in rows]), np.array(vals, dtype=dtype))


def random_csr(m: int, n: int, seed: int, mean: float = 6.0, empty_frac: float = 0.1, long_rows: int = 0,
               long_len: int = 300, sort_cols: bool = True, dtype=np.float64) -> HostCsr:
    """Random pattern: Poisson row lengths, some empty rows, a few long rows."""
    rng = np.random.default_rng(seed)
    lens = rng.poisson(mean, size=m)
    lens[rng.random(m) < empty_frac] = 0
    if long_rows:
        lens[rng.choice(m, size=long_rows, replace=False)] = long_len
    lens = np.minimum(lens, n)
    rows = []
    for r in range(m):
        c = rng.choice(n, size=int(lens[r]), replace=False)
        rows.append(np.sort(c) if sort_cols else c)
    return from_rows(m, n, rows, rng, dtype)


def banded(m: int, offsets, seed: int, dtype=np.float64) -> HostCsr:
    rng = np.random.default_rng(seed)
    rows = []
    for r in range(m):
        rows.append([r + o for o in offsets if 0 <= r + o < m])
    return from_rows(m, m, rows, rng, dtype)


def block_random(mb: int, nb: int, br: int, bc: int, blocks_per_row: int, seed: int, fill: float = 1.0,
                 dtype=np.float64) -> HostCsr:
    rng = np.random.default_rng(seed)
    rows = [[] for _ in range(mb * br)]
    for b in range(mb):
        bcols = np.sort(rng.choice(nb, size=min(blocks_per_row, nb), replace=False))
        for r in range(br):
            cols = []
            for bcl in bcols:
                for c in range(bc):
                    if rng.random() < fill:
                        cols.append(bcl * bc + c)
            rows[b * br + r] = cols
    return from_rows(mb * br, nb * bc, rows, rng, dtype)


def suite(dtype=np.float64) -> dict:
    """The named matrices used by fixtures and parity tests."""
    return {
        "lap8x8": laplace2d(8, 8, dtype),
        "lap33x17": laplace2d(33, 17, dtype),
        "rand200": random_csr(200, 200, seed=1, dtype=dtype),
        "rand_long": random_csr(300, 300, seed=2, long_rows=3, long_len=150, dtype=dtype),
        "rand_empty": random_csr(257, 257, seed=3, mean=2.0, empty_frac=0.5, dtype=dtype),
        "rect_wide": random_csr(60, 150, seed=4, dtype=dtype),
        "rect_tall": random_csr(150, 60, seed=5, dtype=dtype),
        "unsorted": random_csr(120, 120, seed=6, sort_cols=False, dtype=dtype),
        "banded": banded(500, [-40, -3, -1, 0, 1, 2, 17], seed=7, dtype=dtype),
        "blocks4": block_random(30, 30, 4, 4, 5, seed=8, fill=0.8, dtype=dtype),
        "single_row": random_csr(1, 700, seed=9, mean=300.0, empty_frac=0.0, dtype=dtype),
    }
