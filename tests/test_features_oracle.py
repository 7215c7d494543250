"""The CPU restatement of the feature extractor (oracle/features.cpp) against
Get_features' own values (tests/golden/features_golden.npz, written from the
compiled reference), and both against brute force on the entry list.

Bit-exact: every count, every ratio of counts, p-ratios and Gini coefficients,
the log10 variabilities and the distance feature. The sums of squared
deviations (var_*, standard_dev_*, t_var_*, t_standard_dev_*, the std images)
may differ in the last bit: the reference build fuses d*d + acc into an FMA.
"""
import os

import numpy as np
import pytest

from tests.matrices import feature_cases, mtx_text

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "features_golden.npz")
CASES = feature_cases()
ROUNDING = ("var_", "standard_dev_", "t_var_", "t_standard_dev_")


@pytest.fixture(scope="module")
def fgolden():
    with np.load(FIXTURE, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_case(g, name):
    names = [str(s) for s in g[f"{name}/scalar_names"]]
    scal = dict(zip(names, g[f"{name}/scalar_values"].tolist()))
    img = {k.split("/")[-1]: v for k, v in g.items() if k.startswith(f"{name}/img/")}
    return scal, img


def same(a, b):
    return a == b or (np.isnan(a) and np.isnan(b))


def compare_scalars(got, ref, rounding_rtol):
    bad = []
    for k, b in ref.items():
        a = got[k]
        if k.startswith(ROUNDING):
            if not (same(a, b) or abs(a - b) <= rounding_rtol * max(abs(b), 1e-300)):
                bad.append((k, a, b))
        elif not same(a, b):
            bad.append((k, a, b))
    return bad


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference(tmp_path, oracle, fgolden, name):
    path = str(tmp_path / f"{name}.mtx")
    with open(path, "w") as f:
        f.write(CASES[name])
    scal, img = oracle.features(path)
    rscal, rimg = golden_case(fgolden, name)
    assert compare_scalars(scal, rscal, 1e-13) == []
    for k, ref in rimg.items():
        got = img[k]
        assert got.shape == ref.shape, k
        if k.endswith("_std"):
            np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-15, err_msg=k)
        else:
            np.testing.assert_array_equal(got, ref, err_msg=k)


def entries(text):
    """(m, n, rows, cols, vals) of the expanded entry list of a MatrixMarket text."""
    lines = text.splitlines()
    sym = "symmetric" in lines[0]
    pattern = "pattern" in lines[0]
    body = [l for l in lines[1:] if not l.startswith("%")]
    m, n, _ = map(int, body[0].split())
    tok = [l.split() for l in body[1:]]
    r = np.array([int(t[0]) - 1 for t in tok])
    c = np.array([int(t[1]) - 1 for t in tok])
    v = np.ones(len(tok)) if pattern else np.array([float(t[2]) for t in tok])
    if sym:
        off = r != c
        r, c, v = np.concatenate([r, c[off]]), np.concatenate([c, r[off]]), np.concatenate([v, v[off]])
    return m, n, r, c, v


@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_histograms_brute_force(fgolden, name):
    """Row/column/tile/block counts and distinct rows/columns per tile, from the
    expanded entries with numpy's near-equal split of rows and columns into 256
    blocks (the first m % 256 blocks one longer)."""
    m, n, r, c, _ = entries(CASES[name])
    scal, img = golden_case(fgolden, name)
    rb = np.zeros(m, dtype=np.int64)
    for b, idx in enumerate(np.array_split(np.arange(m), 256)):
        rb[idx] = b
    cb = np.zeros(n, dtype=np.int64)
    for b, idx in enumerate(np.array_split(np.arange(n), 256)):
        cb[idx] = b
    np.testing.assert_array_equal(img["row_nnz"], np.bincount(r, minlength=m))
    np.testing.assert_array_equal(img["col_nnz"], np.bincount(c, minlength=n))
    tiles = np.bincount(rb[r] * 256 + cb[c], minlength=65536)
    np.testing.assert_array_equal(img["tile_nnz"], tiles)
    np.testing.assert_array_equal(img["rb_nnz"], np.bincount(rb[r], minlength=256))
    np.testing.assert_array_equal(img["cb_nnz"], np.bincount(cb[c], minlength=256))
    urows = np.unique(r * 256 + cb[c])  # (row, column block) pairs
    np.testing.assert_array_equal(img["tile_uniq_rows"], np.bincount(rb[urows // 256] * 256 + urows % 256,
                                                                     minlength=65536))
    ucols = np.unique(c * 256 + rb[r])
    np.testing.assert_array_equal(img["tile_uniq_cols"], np.bincount((ucols % 256) * 256 + cb[ucols // 256],
                                                                     minlength=65536))
    # per-tile maximum row count
    cnt = np.bincount(r * 256 + cb[c], minlength=m * 256).reshape(m, 256)
    tmax = np.zeros((256, 256), dtype=np.int64)
    np.maximum.at(tmax, rb, cnt)
    np.testing.assert_array_equal(img["tile_max"], tmax.ravel())
    assert scal["num_nnzs"] == len(r)
    assert scal["nnz_lower"] == np.sum(r > c) and scal["nnz_upper"] == np.sum(r < c)
    assert scal["nnz_diagonal"] == np.sum(r == c)
    assert scal["complete_ndiags"] == len(np.unique(c - r))
    # the -1.0 start of the nz-ratio sums: (non-empty count - 1) / blocks
    assert scal["t_nz_ratio_tiles"] == (np.count_nonzero(tiles) - 1) / 65536.0
    assert scal["uniqR"] == len(urows) / len(r)
    assert scal["potReuseC"] == len(ucols) / n


def test_reference_single_nonzero_quirk(tmp_path, oracle):
    """One nonzero: every nz-ratio over tiles/blocks is 0 because its sum starts
    at -1 (the shipped feature table shows nz_tiles_ratio = 0 for such a
    matrix), and the Gini coefficient of 65,536 tiles is 1 - 1/65536."""
    path = str(tmp_path / "one.mtx")
    with open(path, "w") as f:
        f.write(mtx_text(1000, 1000, [5], [7], [2.5]))
    s, _ = oracle.features(path)
    assert s["t_nz_ratio_tiles"] == 0.0 and s["t_nz_ratio_RB"] == 0.0 and s["t_nz_ratio_CB"] == 0.0
    assert s["t_Gini_all_tiles"] == pytest.approx(1.0 - 1.0 / 65536, abs=1e-15)
    assert s["P_ratio_row"] == 1.0 / 1000  # one row holds everything
    assert s["num_nnzs"] == 1 and s["complete_ndiags"] == 1


def test_gini_pratio_closed_forms(tmp_path, oracle):
    """Equal row counts: Gini 0, p-ratio 1/2 (the first k with k/m + k/m >= 1).
    A 2x2 block diagonal of sizes 1 and 3 rows: Gini of [1, 3] over two rows."""
    m = 512
    path = str(tmp_path / "eq.mtx")
    r = np.repeat(np.arange(m), 3)
    c = (r * 7 + np.tile([0, 1, 2], m)) % m
    with open(path, "w") as f:
        f.write(mtx_text(m, m, r, c, np.ones(len(r))))
    s, _ = oracle.features(path)
    assert s["Gini_row"] == 0.0 and s["P_ratio_row"] == 0.5
    path = str(tmp_path / "two.mtx")
    with open(path, "w") as f:
        f.write(mtx_text(2, 3, [0, 1, 1, 1], [0, 0, 1, 2], [1.0, 1.0, 1.0, 1.0]))
    s, _ = oracle.features(path)
    # sorted counts [1, 3]: cumulative [1, 4]; trapezoid 1/2 + 5/2 = 3; total 4*2/2 = 4
    assert s["Gini_row"] == (4.0 - 3.0) / 4.0
    # k = 1: 1/2 + 3/4 >= 1
    assert s["P_ratio_row"] == 0.5


def test_feature_names_match_reference_order():
    """The 40 .features labels in the order FeaturesWrite prints them
    (src/sparse_features.cpp:1347-1400)."""
    from sspmv_amd import lib
    names = lib.feature_names()
    assert len(names) == 40
    assert names[:6] == ["num_rows", "num_cols", "num_nnz", "nnz_ratio", "nnz_on_maindiag_ratio",
                         "nnz_row_distance_ratio"]
    assert names[-4:] == ["uniq_R", "uniq_C", "potReuseR", "potReuseC"]
    assert names[18:24] == ["nz_tiles_ratio", "max_nnz_each_tiles", "ave_nnz_each_tiles", "standard_dev_tiles",
                            "P-ratio_tiles", "Gini_coeff_tiles"]


def test_features_file_layout(tmp_path):
    from sspmv_amd import features
    vec = np.arange(40, dtype=np.float64) + 0.123456789
    vec[0], vec[1], vec[2] = 1000, 900, 12345
    d = features.write_features(str(tmp_path), 7, "mat", vec)
    text = open(os.path.join(d, "mat.features")).read().splitlines()
    assert text[0] == "[7]: mat" and len(text) == 41
    assert text[1] == "num_rows 1000" and text[3] == "num_nnz 12345"
    assert text[4] == "nnz_ratio 3.12346"
    header, vals = features.read_features(os.path.join(d, "mat.features"))
    assert header == "[7]: mat" and vals.shape == (40,)


def test_image_number_format():
    """Image files print numbers as std::ostream does by default (the reference
    writes `file << value << std::endl`, src/sparse_features.cpp:1116-1132):
    integers as integers, doubles like %g with six significant digits."""
    from sspmv_amd import features
    cases = {0.0: "0", 2.0: "2", 0.1: "0.1", 1.0 / 3.0: "0.333333", 2.0 / 3.0: "0.666667", 1234567.0: "1.23457e+06",
             1e-5: "1e-05", 123456.0: "123456", 0.0001234565: "0.000123457", 16.5: "16.5"}
    for v, text in cases.items():
        assert features.fmt_value(v, False) == text, (v, features.fmt_value(v, False))
    assert features.fmt_value(np.int32(65536), True) == "65536"
