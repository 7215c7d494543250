"""GPU feature extraction (lespmv_features) against Get_features' own values
(tests/golden/features_golden.npz) and the CPU restatement.

Counts, ratios of counts, p-ratios and Gini coefficients are compared
bit-exactly. Variances and standard deviations come from exact integer
moments on the GPU and from a running sum of squared differences in the
reference: relative 1e-12. The distance feature is a sum of per-row
quotients, reduced in a different order on the GPU: relative 1e-12. The
log10 variabilities use the device log10: relative 1e-13.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

from tests.matrices import feature_cases, mtx_text
from tests.test_features_oracle import ROUNDING, golden_case, same

pytestmark = pytest.mark.gpu
CASES = feature_cases()
TOL = {"distance_per_row": 1e-12, "row_variability": 1e-13, "col_variability": 1e-13}


@pytest.fixture(scope="module")
def fgolden():
    from tests.test_features_oracle import FIXTURE
    with np.load(FIXTURE, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def gpu_features(lib, path, dtype=np.float64):
    A = lib.read_mtx(path, dtype)
    M = lib.Matrix.from_host(A)
    return lib.features(M, symmetric=A.symmetric, images=True)


def check_against(vals, img, rscal, rimg, rounding_rtol=1e-12):
    bad = []
    for k, b in rscal.items():
        if k not in vals:
            continue
        a = float(vals[k])
        tol = rounding_rtol if k.startswith(ROUNDING) else TOL.get(k, 0.0)
        if not (same(a, b) or abs(a - b) <= tol * max(abs(b), 1e-300)):
            bad.append((k, a, b))
    assert bad == []
    for k, ref in rimg.items():
        if k not in img:
            continue
        got = img[k]
        if k.endswith("_std"):
            np.testing.assert_allclose(got, ref, rtol=rounding_rtol, atol=1e-14, err_msg=k)
        else:
            np.testing.assert_array_equal(got, ref, err_msg=k)


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_features_match_reference(tmp_path, gpu, fgolden, name):
    lib = gpu
    path = str(tmp_path / f"{name}.mtx")
    with open(path, "w") as f:
        f.write(CASES[name])
    vals, vec, img = gpu_features(lib, path)
    rscal, rimg = golden_case(fgolden, name)
    check_against(vals, img, rscal, rimg)
    # the 40-vector in .features order
    names = lib.feature_names()
    assert vec[0] == rscal["num_rows"] and vec[2] == rscal["num_nnzs"]
    assert vec[names.index("nz_tiles_ratio")] == rscal["t_nz_ratio_tiles"]
    assert vec[names.index("Gini_coeff_CB")] == rscal["t_Gini_CB"]
    assert vec[names.index("potReuseC")] == rscal["potReuseC"]
    assert vec[names.index("nnz_row_distance_ratio")] == pytest.approx(
        rscal["distance_per_row"] / rscal["num_cols"], rel=1e-12)
    assert vals["device_ms"] > 0


def test_gpu_features_larger_vs_oracle(tmp_path, gpu, oracle):
    """A 40k-row matrix with power-law rows and a 60k-row banded matrix: every
    tile row holds many rows, and rows longer than a wavefront.

    The banded matrix's column counts are all 7 but near its edges, so the sum
    of squared deviations cancels: the reference's running sum of (c - mean)^2
    is off by about 1e-12 relative there, while the GPU forms it from exact
    integer moments. The deviations are checked against exact rational
    arithmetic instead, and against the running sum within 1e-9."""
    lib = gpu
    rng = np.random.default_rng(11)
    m, n = 40000, 36000
    lens = np.minimum((rng.pareto(1.5, size=m) * 3).astype(np.int64), 3000)
    r = np.repeat(np.arange(m), lens)
    c = rng.integers(0, n, size=len(r))
    v = rng.uniform(-1, 1, size=len(r)) * 10.0 ** rng.uniform(-2, 2, size=len(r))
    order = rng.permutation(len(r))
    cases = {"power": mtx_text(m, n, r[order], c[order], v[order])}
    mb = 60000
    rr, cc = [], []
    for off in (-300, -2, -1, 0, 1, 2, 300):
        i = np.arange(max(0, -off), min(mb, mb - off))
        rr.append(i)
        cc.append(i + off)
    rr, cc = np.concatenate(rr), np.concatenate(cc)
    cases["band"] = mtx_text(mb, mb, rr, cc, rng.uniform(-1, 1, size=len(rr)))
    for name, text in cases.items():
        path = str(tmp_path / f"{name}.mtx")
        with open(path, "w") as f:
            f.write(text)
        vals, _, img = gpu_features(lib, path)
        oscal, oimg = oracle.features(path)
        check_against(vals, img, oscal, oimg, rounding_rtol=1e-9)
        for key, counts in (("row", oimg["row_nnz"]), ("col", oimg["col_nnz"])):
            n = len(counts)
            S = int(counts.sum())
            Q = sum(int(c) * int(c) for c in counts)
            exact = float(Fraction(n * Q - S * S, n * n))
            assert vals[f"var_nnz_each_{key}"] == pytest.approx(exact, rel=1e-15)


def test_gpu_features_fp32_handle_counts(tmp_path, gpu, fgolden):
    """An FP32 handle gives the same counts and count ratios (values only enter
    the value features)."""
    lib = gpu
    name = "gen_real_700x650"
    path = str(tmp_path / f"{name}.mtx")
    with open(path, "w") as f:
        f.write(CASES[name])
    vals, _, img = gpu_features(lib, path, np.float32)
    rscal, rimg = golden_case(fgolden, name)
    for k in ("num_nnzs", "complete_ndiags", "nnz_lower", "t_max_nnz_all_tiles", "uniqR", "uniqC", "Gini_row",
              "P_ratio_col", "t_Gini_all_tiles", "nz_row_ratio"):
        assert float(vals[k]) == rscal[k], k
    np.testing.assert_array_equal(img["tile_uniq_cols"], rimg["tile_uniq_cols"])


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_features_write_files(tmp_path, gpu, fgolden, name):
    """FeaturesWrite's files (src/sparse_features.cpp:1134-1410) from the GPU
    values: the .features file, and every image file (.nnz/.max/.ave/.std over
    the 256 x 256 tiles, .RB* and .CB* over the 256 blocks) read back line by
    line against the compiled reference's image arrays. Integer images must
    print the same integers; double images (ave, std) must print what
    std::ostream's default format (%g, six significant digits) prints for the
    reference's value, or, where the last digit sits on a rounding boundary,
    agree within that digit."""
    from sspmv_amd import features
    path = str(tmp_path / f"{name}.mtx")
    with open(path, "w") as f:
        f.write(CASES[name])
    vals, vec, img = features.extract(path, 3)
    d = features.write_features(str(tmp_path / "out"), 3, name, vec, img)
    header, back = features.read_features(os.path.join(d, name + ".features"))
    assert header == f"[3]: {name}" and back.shape == (40,)
    np.testing.assert_allclose(back, vec, rtol=1e-5)
    _, rimg = golden_case(fgolden, name)
    boundary = 0
    for suffix, key in features.IMAGE_FILES:
        lines = open(os.path.join(d, f"{name}.{suffix}")).read().splitlines()
        ref = rimg[key]
        assert len(lines) == (65536 if key.startswith("tile_") else 256) == len(ref), (suffix, len(lines))
        if ref.dtype.kind in "iu":
            assert lines == [str(int(v)) for v in ref], suffix
            continue
        want = ["%g" % v for v in ref.tolist()]
        for i, (got, w) in enumerate(zip(lines, want)):
            if got != w:
                boundary += 1
                assert abs(float(got) - ref[i]) <= 1e-5 * abs(ref[i]), (suffix, i, got, w)
    assert boundary <= 4, boundary


def test_get_features_driver(tmp_path, fgolden):
    """The C++ drop-in (class MTX in include/lespmv/LeSpMV.h) through the
    get_features driver: printed values and the written .features file."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    name = "sym_real_600"
    path = str(tmp_path / f"{name}.mtx")
    with open(path, "w") as f:
        f.write(CASES[name])
    out = subprocess.run([os.path.join(root, "sspmv_amd", "bin", "get_features"), path, "--matID=5",
                          f"--write={tmp_path / 'feat'}"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rscal, _ = golden_case(fgolden, name)
    lines = out.stdout.splitlines()
    assert f"real_nnz   = {int(rscal['num_nnzs'])}" in lines
    assert f"num_of_diags  = {int(rscal['complete_ndiags'])}" in lines
    assert any(l.startswith(f"5 {name} ") for l in lines)
    from sspmv_amd import features
    header, vals = features.read_features(str(tmp_path / "feat" / name / f"{name}.features"))
    assert header == f"[5]: {name}"
    assert vals[2] == rscal["num_nnzs"] and vals[7] == rscal["max_nnz_each_row"]
    assert vals[11] == pytest.approx(rscal["Gini_row"], rel=1e-5)
