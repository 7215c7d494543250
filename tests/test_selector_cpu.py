"""Selector bookkeeping on the CPU: labels from timings, regret, sampling."""
import numpy as np

from sspmv_amd import lib, selector


def test_labels_pick_fastest_and_per_class_best():
    cf, hc, hs = lib.selector_layout()
    t = np.full((3, selector.N_GRID + 1), np.inf)
    t[0, cf[4] + 5] = 1.0   # SELL (C=64, S-ELL) fastest
    t[0, cf[0] + 2] = 2.0   # CSR vector, 8 lanes
    t[0, cf[6] + 1] = 3.0   # BSR 4x4
    t[1, cf[5] + 1] = 0.5   # DIA with LDS windows
    t[1, cf[5]] = 0.7
    t[2, cf[8] + 1] = 0.4   # CSR with 4 column panels
    t[2, cf[8]] = 0.6       # 8 panels
    t[2, cf[1]] = 0.9       # merge path
    t[:, selector.N_GRID] = 5.0
    L = selector.labels(t)
    assert L[0].tolist() == [4, 2, 5, -1, 1, -1, -1]
    assert L[1].tolist() == [5, -1, -1, 1, -1, -1, -1]
    assert L[2].tolist() == [8, -1, -1, -1, -1, -1, 1]
    assert hs.tolist() == [9, 6, 8, 2, 5, 3, 2] and cf[-1] == selector.N_GRID == 29
    grid = lib.tune_candidates(selector.GRID_MASK)
    assert [(grid[j].fmt, grid[j].variant) for j in (27, 28)] == [(lib.CSR, lib.VARIANT_PANEL),
                                                                  (lib.CSR, lib.VARIANT_PANEL4)]
    assert len(lib.tune_candidates()) == 27  # without the panel bit the grid is unchanged


def test_older_database_widened():
    """A database timed on the 27-candidate grid gains the two panel columns
    (not yet timed: inf) before the default kernel's, and three allowed flags."""
    n = 3
    old = {"times": np.arange(n * 28, dtype=np.float64).reshape(n, 28), "convert_ms": np.ones((n, 28)),
           "allowed": np.zeros((n, 32), np.uint8)}
    new = selector.widen_db(old)
    assert new["times"].shape == (n, 30) and np.isinf(new["times"][:, 27:29]).all()
    np.testing.assert_array_equal(new["times"][:, 29], old["times"][:, 27])
    np.testing.assert_array_equal(new["times"][:, :27], old["times"][:, :27])
    assert new["allowed"].shape == (n, 35) and (new["allowed"][:, [8, 33, 34]] == 1).all()
    assert selector.widen_db(new) is new


def test_regret_and_fallback():
    t = np.array([[1.0] + [2.0] * 28 + [4.0], [np.inf, 3.0] + [6.0] * 27 + [3.0]])
    r = selector.evaluate(t, np.array([0, 0]))  # second pick does not apply: default CSR (3.0) is used
    assert r["within_10pct"] == 1.0 and r["mean_regret"] == 0.0
    r = selector.evaluate(t, np.array([1, 2]))
    assert np.isclose(r["mean_regret"], (1.0 + 1.0) / 2)


def test_sampling_covers_families_and_respects_nnz_cap():
    rng = np.random.default_rng(0)
    fams = set()
    for _ in range(400):
        fam, gen, params = selector.sample_matrix(rng, 6e6)
        fams.add(fam)
        if fam == "uniform":
            assert params[0] * params[1] * params[2] <= 6e6 * 1.0001
        if fam == "blockdiag":
            assert params[0] * params[1] <= 6e6 * 1.0001 or params[1] == 2
    assert fams == set(selector.FAMILIES)


def test_reference_labels_and_probability_vector(tmp_path):
    """Grid timings folded into the reference's nine labels; the probability
    label is each format's best GFLOP/s over their sum."""
    cf, _, _ = lib.selector_layout()
    t = np.full((1, selector.N_GRID + 1), np.inf)
    t[0, cf[0] + 1] = 2.0      # CSR vector
    t[0, cf[1]] = 1.0          # CSR merge: CSR's best is 1.0
    t[0, cf[5]] = 4.0          # DIA
    grid = lib.tune_candidates()
    sell = [j for j in range(27) if grid[j].fmt == lib.SELL_C_SIGMA][0]
    t[0, sell] = 2.0
    ft = selector.reference_format_times(t)[0]
    assert ft[1] == 1.0 and ft[2] == 4.0 and ft[5] == 2.0 and np.isinf(ft[0])
    p = selector.probability_vectors(t, np.array([1000]))[0]
    g = np.array([0, 1 / 1.0, 1 / 4.0, 0, 0, 1 / 2.0, 0, 0, 0])
    np.testing.assert_allclose(p, g / g.sum())
    d = selector.write_label_files(str(tmp_path), "m", t[0], 1000)
    assert open(f"{d}/m.format_label").read() == "CSR 1\n"
    toks = open(f"{d}/m.prob_label").read().split()
    assert toks[0] == "CSR" and len(toks) == 10
