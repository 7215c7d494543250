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


# ---- MatrixMarket files for the feature extractor ---------------------------------


def mtx_text(m: int, n: int, r, c, v=None, field: str = "real", symmetry: str = "general") -> str:
    """A MatrixMarket coordinate file listing (r, c[, v]) 0-based, in the given order."""
    lines = [f"%%MatrixMarket matrix coordinate {field} {symmetry}", "% generated test matrix",
             f"{m} {n} {len(r)}"]
    r = np.asarray(r).tolist()
    c = np.asarray(c).tolist()
    if field == "pattern":
        lines += [f"{a + 1} {b + 1}" for a, b in zip(r, c)]
    elif field == "integer":
        lines += [f"{a + 1} {b + 1} {int(x)}" for a, b, x in zip(r, c, np.asarray(v).tolist())]
    else:
        lines += [f"{a + 1} {b + 1} {float(x)!r}" for a, b, x in zip(r, c, np.asarray(v).tolist())]
    return "\n".join(lines) + "\n"


def _random_entries(rng, m, n, mean, empty_frac=0.1, long_rows=0, long_len=0, dups=0):
    lens = rng.poisson(mean, size=m)
    lens[rng.random(m) < empty_frac] = 0
    for i in rng.choice(m, size=long_rows, replace=False) if long_rows else []:
        lens[i] = min(long_len, n)
    r, c = [], []
    for i, k in enumerate(lens):
        cols = rng.choice(n, size=min(int(k), n), replace=False)
        r += [i] * len(cols)
        c += cols.tolist()
    r, c = np.array(r, dtype=np.int64), np.array(c, dtype=np.int64)
    if dups and len(r):
        pick = rng.choice(len(r), size=dups)
        r, c = np.concatenate([r, r[pick]]), np.concatenate([c, c[pick]])
    order = rng.permutation(len(r))  # file order is not row order
    return r[order], c[order]


def _values(rng, k):
    """Magnitudes over six decades, both signs, a few exact zeros."""
    v = rng.uniform(-1.0, 1.0, size=k) * 10.0 ** rng.uniform(-3, 3, size=k)
    v[rng.random(k) < 0.02] = 0.0
    return v


def _lower_entries(rng, n, mean):
    r, c = [], []
    for i in range(n):
        k = min(rng.poisson(mean), i)
        cols = rng.choice(i, size=k, replace=False) if k else []
        r += [i] * len(cols)
        c += list(cols)
        if rng.random() < 0.8:
            r.append(i)
            c.append(i)
    order = rng.permutation(len(r))
    return np.array(r)[order], np.array(c)[order]


def _symvals(rng, n):
    """A square general matrix whose mirrored pairs exercise the "%lf" value
    comparison: equal values, differences below and above 1e-6, rounding ties
    (k/128), signed zeros, and duplicates whose first value counts."""
    r, c, v = [], [], []

    def pair(a, b, x, y):
        r.extend([a, b]); c.extend([b, a]); v.extend([x, y])

    cells = rng.choice(n * n, size=900, replace=False)
    specials = [(0.5, 0.5), (0.0078125, 0.00781249), (0.0234375, 0.0234384), (0.0234375, 0.0234374),
                (1e-9, -1e-9), (0.0, -0.0), (1.25, 1.2500004), (1.25, 1.2500006), (-3.0, -3.0),
                (123456.1234565, 123456.123456), (2.0 ** 53, 2.0 ** 53 + 2), (1e-320, 0.0)]
    for idx, cell in enumerate(cells):
        a, b = divmod(int(cell), n)
        if a == b:
            r.append(a); c.append(b); v.append(float(rng.uniform(-2, 2)))
            continue
        kind = idx % 4
        x = float(rng.uniform(-1, 1))
        if kind == 0:
            pair(a, b, *specials[idx % len(specials)])
        elif kind == 1:
            pair(a, b, x, x)
        elif kind == 2:
            pair(a, b, x, x + 3e-6)
        else:
            r.append(a); c.append(b); v.append(x)  # no mirror
    # duplicates: the second value differs, the first is the one compared
    for k in range(0, 40, 2):
        r.append(r[k]); c.append(c[k]); v.append(v[k] + 1.0)
    return np.array(r), np.array(c), np.array(v)


def feature_cases() -> dict:
    """name -> MatrixMarket text for the feature extractor tests (deterministic)."""
    rng = np.random.default_rng(20240123)
    out = {}
    r, c = _random_entries(rng, 700, 650, 8, dups=30)
    out["gen_real_700x650"] = mtx_text(700, 650, r, c, _values(rng, len(r)))
    r, c = _lower_entries(rng, 600, 4)
    out["sym_real_600"] = mtx_text(600, 600, r, c, _values(rng, len(r)), symmetry="symmetric")
    r, c = _random_entries(rng, 520, 777, 5, empty_frac=0.3)
    out["pat_gen_520x777"] = mtx_text(520, 777, r, c, field="pattern")
    r, c = _lower_entries(rng, 530, 3)
    out["pat_sym_530"] = mtx_text(530, 530, r, c, field="pattern", symmetry="symmetric")
    r, c = _random_entries(rng, 300, 1000, 12)
    out["int_gen_300x1000"] = mtx_text(300, 1000, r, c, rng.integers(-5, 6, size=len(r)), field="integer")
    r, c = _random_entries(rng, 1000, 300, 3, empty_frac=0.2)
    out["real_tall_1000x300"] = mtx_text(1000, 300, r, c, _values(rng, len(r)))
    r, c = _random_entries(rng, 50, 40, 4)
    out["small_50x40"] = mtx_text(50, 40, r, c, _values(rng, len(r)))
    m = 3000
    rr, cc = [], []
    for off in (-50, -1, 0, 1, 50, 1200):
        i = np.arange(max(0, -off), min(m, m - off))
        rr.append(i); cc.append(i + off)
    r, c = np.concatenate(rr), np.concatenate(cc)
    out["band_3000"] = mtx_text(m, m, r, c, rng.uniform(-1, 1, size=len(r)))
    r, c, v = _symvals(rng, 800)
    out["symvals_800"] = mtx_text(800, 800, r, c, v)
    nx = 30
    r, c, v = [], [], []
    for y in range(nx):
        for x in range(nx):
            i = y * nx + x
            for j, w in ((i - nx, -1.0), (i - 1, -1.0), (i, 4.0), (i + 1, -1.0), (i + nx, -1.0)):
                if 0 <= j < nx * nx and (abs(j - i) != 1 or j // nx == y):
                    r.append(i); c.append(j); v.append(w)
    out["lap_30x30"] = mtx_text(nx * nx, nx * nx, r, c, v)
    r, c = _random_entries(rng, 2000, 2100, 6, long_rows=6, long_len=700)
    out["gen_long_2000x2100"] = mtx_text(2000, 2100, r, c, _values(rng, len(r)))
    return out
