"""Learned format selector: timing database, GPU training, regret.

No reference counterpart in code: the reference's selector (SmartAdpter's
MM-Adapter, SmartAdpter/models/SpMV_Adapter.py) is a Keras model trained on
CPU timing spreadsheets. Here every step runs through the library:

* ``collect``: synthetic matrices from the library's generators (banded,
  power-law rows, block-diagonal, uniform random, plus lognormal, 4x4-block
  random, 2-D and 3-D stencils), GPU features -> 128 selector inputs, and an
  exhaustive GPU timing of the 27-candidate grid (``lespmv_tune``) plus the
  library's default CSR kernel. Exhaustive timing is the arbiter and the fallback.
* ``labels``: per matrix the class of the fastest candidate, and for each class
  with parameters the fastest candidate of that class.
* ``train``: the MLP (128 -> 128 -> 128 -> 32, six softmax heads) trained by the
  library's own HIP forward / backward / Adam kernels.
* ``evaluate``: the decoded pick's measured time over the best time (regret),
  against the library default (CSR, automatic kernel) and a fixed rule.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import lib

FAMILIES = ["banded", "powerlaw", "blockdiag", "uniform", "lognormal", "blockrand", "laplace2d", "stencil27"]
CLASS_NAMES = ["CSR-vector", "CSR-merge", "COO", "ELL", "SELL", "DIA", "BSR", "CSR5", "CSR-panel"]
N_CLASSES = len(CLASS_NAMES)
# candidates of the standard grid (all formats plus the two CSR column-panel
# variants, grid indices 27 and 28); times[:, N_GRID] is the default CSR kernel
N_GRID = 29
GRID_MASK = lib.ALL_FORMATS_MASK | lib.TUNE_PANELS


def sample_matrix(rng: np.random.Generator, max_nnz: float, n_range=(1e4, 1e6)):
    """(family, generator id, params) of one synthetic matrix."""
    fam = FAMILIES[int(rng.integers(len(FAMILIES)))]
    n = int(10 ** rng.uniform(math.log10(n_range[0]), math.log10(n_range[1])))
    if fam == "banded":
        hb = int(10 ** rng.uniform(0, 2.5))
        fill = float(rng.uniform(0.05, 1.0))
        fill = min(fill, max_nnz / (n * (2 * hb + 1)))
        return fam, lib.GEN_BANDED, [n, hb, max(fill, 1e-3)]
    if fam == "powerlaw":
        a = float(rng.uniform(1.5, 3.0))
        mean = float(10 ** rng.uniform(0.3, 1.5))
        mean = min(mean, max_nnz / n)
        return fam, lib.GEN_POWERLAW, [n, a, mean, 4096]
    if fam == "blockdiag":
        blk = int(rng.integers(2, 65))
        blk = max(2, min(blk, int(max_nnz / n)))
        return fam, lib.GEN_BLOCKDIAG, [n, blk]
    if fam == "uniform":
        d = 10 ** rng.uniform(-6, -3)
        d = min(d, max_nnz / (float(n) * n))
        return fam, lib.GEN_UNIFORM, [n, n, d]
    if fam == "lognormal":
        mean = float(10 ** rng.uniform(0.5, 1.8))
        mean = min(mean, max_nnz / n)
        return fam, lib.GEN_LOGNORMAL, [n, n, mean, float(rng.uniform(0.3, 1.5)), 4096, float(rng.uniform(0, 1)),
                                        64]
    if fam == "blockrand":
        b = int(rng.choice([2, 4, 8]))
        bpr = int(rng.integers(2, 17))
        bpr = max(1, min(bpr, int(max_nnz / (n * b))))
        mb = max(n // b, 1)
        return fam, lib.GEN_BLOCKRAND, [mb, mb, b, b, bpr]
    if fam == "laplace2d":
        nx = max(int(math.sqrt(n)), 8)
        return fam, lib.GEN_LAPLACE2D, [nx, nx, 4.0, -1.0]
    k = max(int(round(n ** (1 / 3))), 4)
    k = min(k, int((max_nnz / 27) ** (1 / 3)))
    return "stencil27", lib.GEN_STENCIL27, [k, 26.0, -1.0]


def collect(count: int, out: str, seed: int = 1, iters: int = 10, max_nnz: float = 6e6, n_range=(1e4, 1e6),
            log=print):
    """Time the grid on `count` generated matrices; writes an .npz timing database."""
    rng = np.random.default_rng(seed)
    grid = lib.tune_candidates(GRID_MASK)
    rows = {k: [] for k in ("X", "allowed", "times", "convert_ms", "family", "m", "nnz", "seed")}
    t0 = time.time()
    for i in range(count):
        fam, gen, params = sample_matrix(rng, max_nnz, n_range)
        mseed = seed * 1_000_003 + i
        A = lib.generate(gen, params, mseed)
        M = lib.Matrix.from_host(A)
        x, allowed, _ = lib.selector_inputs(M)
        cands = (lib.TuneResultC * (N_GRID + 1))()
        for j in range(N_GRID):
            cands[j] = grid[j]
        cands[N_GRID].fmt = lib.CSR
        cands[N_GRID].variant = lib.VARIANT_AUTO
        M.tune(cands, iters=iters)
        times = np.array([c.ms if c.status == 0 else np.inf for c in cands])
        rows["X"].append(x)
        rows["allowed"].append(allowed)
        rows["times"].append(times)
        rows["convert_ms"].append(np.array([c.convert_ms for c in cands]))
        rows["family"].append(FAMILIES.index(fam))
        rows["m"].append(A.m)
        rows["nnz"].append(A.nnz)
        rows["seed"].append(mseed)
        del M, A
        if (i + 1) % 50 == 0 or i + 1 == count:
            log(f"collect {i + 1}/{count} {time.time() - t0:.0f}s", flush=True)
            save_db(out, rows)
    return save_db(out, rows)


def save_db(path, rows):
    db = {k: np.asarray(v) for k, v in rows.items()}
    np.savez_compressed(path, **db)
    return db


def load_db(path):
    with np.load(path, allow_pickle=False) as z:
        db = {k: z[k] for k in z.files}
    # inputs written before non-finite values were mapped to 0 in
    # lespmv_selector_inputs (empty matrices): the same mapping
    db["X"] = np.where(np.isfinite(db["X"]), db["X"], 0.0).astype(np.float32)
    return widen_db(db)


def labels(times: np.ndarray, tie: float = 0.0) -> np.ndarray:
    """(n, 7) labels: class of the fastest grid candidate, then per parameter
    head the fastest candidate of that class (-1 when none of it applies).
    With tie > 0, candidates within a factor (1 + tie) of the fastest count as
    equally fast and the label goes to the one whose class wins most often
    overall, so near-equal timings do not scatter the labels."""
    cf, hc, hs = lib.selector_layout()
    g = times[:, :N_GRID]
    cls_of = np.zeros(N_GRID, dtype=np.int32)
    for c in range(N_CLASSES):
        cls_of[cf[c]:cf[c + 1]] = c
    L = np.full((len(g), len(hs)), -1, dtype=np.int32)
    best = np.argmin(g, axis=1)
    if tie > 0:
        prior = np.bincount(cls_of[best], minlength=N_CLASSES).astype(np.float64)
        near = g <= (1.0 + tie) * g[np.arange(len(g)), best][:, None]
        score = np.where(near, prior[cls_of][None, :] - g / g.max(), -np.inf)
        best = np.argmax(score, axis=1)
    L[:, 0] = cls_of[best]
    for c in range(N_CLASSES):
        h = hc[c]
        if h < 0:
            continue
        seg = g[:, cf[c]:cf[c + 1]]
        ok = np.isfinite(seg).any(axis=1)
        L[ok, h] = np.argmin(seg[ok], axis=1)
    return L


# The reference's nine labels (LeSpMV/script/write_label.py:63-75).
REFERENCE_LABELS = ["COO", "CSR", "DIA", "ELL", "S-ELL", "S-ELL-sigma", "S-ELL-R", "CSR5", "BSR"]


def reference_format_times(times: np.ndarray) -> np.ndarray:
    """(n, 9) best time per reference format label (inf where none applies):
    every grid candidate belongs to one label (every CSR kernel to CSR)."""
    grid = lib.tune_candidates(GRID_MASK)
    fmt_label = {lib.COO: 0, lib.CSR: 1, lib.DIA: 2, lib.ELL: 3, lib.SELL: 4, lib.SELL_C_SIGMA: 5,
                 lib.SELL_C_R: 6, lib.CSR5: 7, lib.BSR: 8}
    out = np.full((len(times), 9), np.inf)
    for j in range(N_GRID):
        lab = fmt_label[grid[j].fmt]
        out[:, lab] = np.minimum(out[:, lab], times[:, j])
    return out


def probability_vectors(times: np.ndarray, nnz: np.ndarray) -> np.ndarray:
    """The reference's probability label: each format's best GFLOP/s divided
    by the sum over the nine formats (0 for formats that do not apply)."""
    t = reference_format_times(times)
    gflops = np.where(np.isfinite(t), 2.0 * np.asarray(nnz, dtype=np.float64)[:, None] / (t * 1e6), 0.0)
    s = gflops.sum(axis=1, keepdims=True)
    return gflops / np.where(s > 0, s, 1.0)


def write_label_files(root: str, name: str, times_row: np.ndarray, nnz: int) -> str:
    """<root>/<name>/<name>.format_label ("<FORMAT> <id>") and .prob_label
    ("<best FORMAT> p_COO ... p_BSR"), as write_label.py lays them out
    (LeSpMV/script/write_label.py:14-61)."""
    p = probability_vectors(times_row[None, :], np.array([nnz]))[0]
    best = int(np.argmax(p))
    d = os.path.join(root, name)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, name + ".format_label"), "w") as f:
        f.write(f"{REFERENCE_LABELS[best]} {best}\n")
    with open(os.path.join(d, name + ".prob_label"), "w") as f:
        f.write(" ".join([REFERENCE_LABELS[best]] + [repr(float(v)) for v in p]) + "\n")
    return d


def standardiser(X):
    mean = X.mean(axis=0)
    std = X.std(axis=0)
    std[std < 1e-3] = 1.0
    return mean.astype(np.float32), std.astype(np.float32)


def train(X, L, epochs=400, batch=256, lr=2e-3, weight_decay=1e-4, seed=7, hidden=128):
    """The selector network trained on the GPU; returns (mlp, loss history)."""
    _, _, hs = lib.selector_layout()
    mlp = lib.Mlp(X.shape[1], hidden, hidden, hs.tolist(), seed=seed)
    mean, std = standardiser(X)
    mlp.set_normalizer(mean, std)
    hist = mlp.train(X, L, epochs, batch=batch, lr=lr, weight_decay=weight_decay, seed=seed)
    return mlp, hist


def picks(mlp, X, allowed):
    probs = mlp.forward(X)
    return np.array([lib.selector_decode(p, a)[1] for p, a in zip(probs, allowed)], dtype=np.int64)


def pick_time(times, cand):
    """Measured time of a pick; a pick that does not apply falls back to the
    library default (CSR, automatic kernel)."""
    t = times[np.arange(len(times)), cand]
    return np.where(np.isfinite(t), t, times[:, N_GRID])


def rule_pick(db):
    """A fixed rule without the network: DIA when the features allow it, else
    the default CSR kernel."""
    allowed = db["allowed"]
    cand = np.full(len(allowed), N_GRID, dtype=np.int64)
    cf, _, _ = lib.selector_layout()
    cand[allowed[:, 5] == 1] = cf[5]  # DIA, thread per row
    return cand


def best_fixed(train_times, test_times):
    """The single grid candidate with the lowest mean regret on the training
    matrices (the strongest selector-free static choice), on the test set."""
    best = train_times[:, :N_GRID].min(axis=1, keepdims=True)
    fallback = train_times[:, N_GRID:N_GRID + 1]
    t = np.where(np.isfinite(train_times[:, :N_GRID]), train_times[:, :N_GRID], fallback)
    j = int(np.argmin(np.mean(t / best - 1.0, axis=0)))
    r = evaluate(test_times, np.full(len(test_times), j))
    r["candidate"] = j
    return r


def evaluate(times, cand):
    best = times[:, :N_GRID].min(axis=1)
    t = pick_time(times, cand)
    regret = t / best - 1.0
    return {"within_10pct": float(np.mean(regret <= 0.10)), "mean_regret": float(np.mean(regret)),
            "median_regret": float(np.median(regret)), "p90_regret": float(np.quantile(regret, 0.9)),
            "geomean_speedup_vs_default": float(np.exp(np.mean(np.log(times[:, N_GRID] / t))))}


def split(n, frac=0.8, seed=3):
    idx = np.random.default_rng(seed).permutation(n)
    k = int(round(frac * n))
    return idx[:k], idx[k:]


def run_training(db_path, model_out, report_out=None, epochs=400, seed=7, tie=0.0, weight_decay=1e-4, hidden=128,
                 log=print):
    db = load_db(db_path)
    X, times = db["X"].astype(np.float32), db["times"]
    L = labels(times, tie)
    tr, te = split(len(X))
    mlp, hist = train(X[tr], L[tr], epochs=epochs, seed=seed, weight_decay=weight_decay, hidden=hidden)
    mlp.save(model_out)
    cand = picks(mlp, X[te], db["allowed"][te])
    res = {
        "matrices": int(len(X)), "train": int(len(tr)), "test": int(len(te)), "epochs": epochs, "tie": tie,
        "weight_decay": weight_decay, "hidden": hidden,
        "final_loss": float(hist[-1]), "first_loss": float(hist[0]),
        "selector": evaluate(times[te], cand),
        "default_csr": evaluate(times[te], np.full(len(te), N_GRID)),
        "rule_dia_else_csr": evaluate(times[te], rule_pick({k: v[te] for k, v in db.items()})),
        "best_fixed_candidate": best_fixed(times[tr], times[te]),
        "train_selector": evaluate(times[tr], picks(mlp, X[tr], db["allowed"][tr])),
        "class_histogram": np.bincount(L[:, 0], minlength=N_CLASSES).tolist(),
        "class_names": CLASS_NAMES,
    }
    if report_out:
        with open(report_out, "w") as f:
            json.dump(res, f, indent=1)
    log(json.dumps(res["selector"]), flush=True)
    return mlp, res


def run_evaluation(db_path, model_path, report_out=None, log=print):
    """A trained model on a timing database it was not trained on (for
    instance one collected with a different seed, or after kernel changes):
    the same regret figures as the training report, all matrices held out."""
    db = load_db(db_path)
    X, times = db["X"].astype(np.float32), db["times"]
    mlp = lib.Mlp(path=model_path)
    res = {
        "matrices": int(len(X)), "model": os.path.basename(model_path),
        "selector": evaluate(times, picks(mlp, X, db["allowed"])),
        "default_csr": evaluate(times, np.full(len(X), N_GRID)),
        "rule_dia_else_csr": evaluate(times, rule_pick(db)),
    }
    if report_out:
        with open(report_out, "w") as f:
            json.dump(res, f, indent=1)
    log(json.dumps(res), flush=True)
    return res


def widen_db(db):
    """A database timed on the 27-candidate grid (times: 27 + the default CSR
    kernel; allowed: 32 flags) laid out for the grid with the two panel
    candidates: their columns (inf until timed) go before the default, and the
    allowed flags gain the panel class and its two parameters."""
    t = db["times"]
    if t.shape[1] == N_GRID + 1:
        return db
    assert t.shape[1] == 28, t.shape
    n = len(t)
    out = dict(db)
    inf2 = np.full((n, 2), np.inf)
    out["times"] = np.concatenate([t[:, :27], inf2, t[:, 27:]], axis=1)
    out["convert_ms"] = np.concatenate([db["convert_ms"][:, :27], np.zeros((n, 2)), db["convert_ms"][:, 27:]],
                                       axis=1)
    a = db["allowed"]
    ones = np.ones((n, 1), np.uint8)
    out["allowed"] = np.concatenate([a[:, :8], ones, a[:, 8:], ones, ones], axis=1)
    return out


def retime(db_path, out, columns, iters=10, max_nnz=6e6, n_max=2e6, rows=None, log=print):
    """Times the grid candidates `columns` (grid indices; N_GRID is the default
    CSR kernel) again on the matrices of a timing database (all rows, or the
    rows [start, stop) given, which alone are written to `out`): after a
    kernel change, or for the two column-panel candidates (27, 28) in a
    database collected before they joined the grid. The matrices are made
    again from their stored seeds (collection seed * 1,000,003 + index: the
    sampler is replayed from that collection's start, and the regenerated
    sizes and family must equal the stored ones), so the other columns stay
    as they were timed."""
    columns = [int(c) for c in columns]
    assert all(0 <= c <= N_GRID for c in columns), columns
    db = widen_db(load_db(db_path))
    if rows is not None:
        db = {k: v[rows[0]:rows[1]] for k, v in db.items()}
    grid = lib.tune_candidates(GRID_MASK)
    seeds, idx = np.divmod(db["seed"], 1_000_003)
    rngs, pos = {}, {}
    t0 = time.time()
    for r in range(len(seeds)):
        s, i = int(seeds[r]), int(idx[r])
        if s not in rngs:
            rngs[s], pos[s] = np.random.default_rng(s), 0
        assert pos[s] <= i, f"rows of seed {s} out of order"
        while pos[s] < i:  # matrices of this collection before the first row taken
            sample_matrix(rngs[s], max_nnz, (1e4, n_max))
            pos[s] += 1
        fam, gen, params = sample_matrix(rngs[s], max_nnz, (1e4, n_max))
        pos[s] += 1
        A = lib.generate(gen, params, int(db["seed"][r]))
        assert (A.m, A.nnz, FAMILIES.index(fam)) == (db["m"][r], db["nnz"][r], db["family"][r]), r
        M = lib.Matrix.from_host(A)
        cands = (lib.TuneResultC * len(columns))()
        for j, c in enumerate(columns):
            if c < N_GRID:
                cands[j] = grid[c]
            else:
                cands[j].fmt, cands[j].variant = lib.CSR, lib.VARIANT_AUTO
        M.tune(cands, iters=iters)
        for j, c in enumerate(columns):
            db["times"][r, c] = cands[j].ms if cands[j].status == 0 else np.inf
            db["convert_ms"][r, c] = cands[j].convert_ms
        del M, A
        if (r + 1) % 100 == 0 or r + 1 == len(seeds):
            log(f"retime {r + 1}/{len(seeds)} {time.time() - t0:.0f}s", flush=True)
    np.savez_compressed(out, **db)
    return db


class Selector:
    """A trained selector (weights file written by ``run_training``)."""

    def __init__(self, path: str):
        self.mlp = lib.Mlp(path=path)

    def select(self, M, symmetric: bool = False):
        return lib.select(M, self.mlp, symmetric)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m sspmv_amd.selector")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--count", type=int, default=2000)
    c.add_argument("--out", required=True)
    c.add_argument("--seed", type=int, default=1)
    c.add_argument("--iters", type=int, default=10)
    c.add_argument("--max-nnz", type=float, default=6e6)
    c.add_argument("--n-max", type=float, default=1e6)
    t = sub.add_parser("train")
    t.add_argument("--db", required=True, nargs="+")
    t.add_argument("--model", required=True)
    t.add_argument("--report")
    t.add_argument("--db-out", help="where to write the merged database")
    t.add_argument("--epochs", type=int, default=400)
    t.add_argument("--tie", type=float, default=0.0)
    t.add_argument("--weight-decay", type=float, default=1e-4)
    t.add_argument("--hidden", type=int, default=128)
    for name, hlp in (("retime", "time grid candidates again on a database's matrices"),
                      ("retime-panels", "time the CSR column-panel candidates into an older database")):
        rp = sub.add_parser(name, help=hlp)
        rp.add_argument("--db", required=True)
        rp.add_argument("--out", required=True)
        rp.add_argument("--iters", type=int, default=10)
        rp.add_argument("--n-max", type=float, default=2e6)
        rp.add_argument("--rows", help="START:STOP, the rows to time (written alone to --out)")
        if name == "retime":
            rp.add_argument("--columns", required=True,
                            help=f"comma list of grid indices ({N_GRID} = the default CSR kernel)")
    e = sub.add_parser("evaluate")
    e.add_argument("--db", required=True)
    e.add_argument("--model", required=True)
    e.add_argument("--report")
    a = ap.parse_args(argv)
    lib.check(lib.load().lespmv_init(0))
    if a.cmd == "collect":
        collect(a.count, a.out, a.seed, a.iters, a.max_nnz, (1e4, a.n_max))
    elif a.cmd in ("retime", "retime-panels"):
        rows = tuple(int(v) for v in a.rows.split(":")) if a.rows else None
        cols = a.columns.split(",") if a.cmd == "retime" else (N_GRID - 2, N_GRID - 1)
        retime(a.db, a.out, cols, a.iters, n_max=a.n_max, rows=rows)
    elif a.cmd == "evaluate":
        run_evaluation(a.db, a.model, a.report)
    else:
        path = a.db[0]
        if len(a.db) > 1 or a.db_out:  # merge several databases
            parts = [load_db(p) for p in a.db]
            merged = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
            path = a.db_out or a.model + ".db.npz"
            np.savez_compressed(path, **merged)
        run_training(path, a.model, a.report, a.epochs, tie=a.tie, weight_decay=a.weight_decay, hidden=a.hidden)


if __name__ == "__main__":
    sys.exit(main())
