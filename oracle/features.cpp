// CPU restatement of LeSpMV's matrix feature extractor (the MTX class).
// TEST INFRASTRUCTURE ONLY: used by tests/ to check the GPU feature kernels;
// nothing in the product links this file.
//
// Follows, with IndexType = int and ValueType = double (Get_features
// --precision=64):
//   MTX::MtxLoad                      src/sparse_features.cpp:121-604
//   P_ratioAndGini, trapezoidalRule    src/sparse_features.cpp:611-640, :79-96
//   MTX::CalculateFeatures            src/sparse_features.cpp:642-776
//   MTX::CalculateTilesFeatures       src/sparse_features.cpp:782-853
//   MTX::CalculateTilesExtraFeatures  src/sparse_features.cpp:941-968 (BSR_ANA off)
//   initial member values             include/sparse_features.h:150-328
// and, for the CSR that CalculateTilesExtraFeatures re-reads, read_coo_matrix /
// read_csr_matrix (utils/sparse_io.cpp:26-204) with coo_to_csr non-compact.
//
// Every feature is stored as a one-element double array "f:<member name>"
// (trailing underscore dropped); the per-tile and per-block images under
// "img:<name>".
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "named.hpp"

namespace {

struct MtxFile {
  int m = 0, n = 0, nnz = 0;
  bool symmetric = false, pattern = false;
  std::vector<int> r, c;  // 0-based, file order
  std::vector<double> v;
};

std::string lower(std::string s) {
  for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

// MatrixMarket coordinate reader (banner, comments, size line, entries).
bool read_mtx(const char* path, MtxFile& f, std::string& err) {
  FILE* fp = std::fopen(path, "r");
  if (!fp) { err = "cannot open"; return false; }
  char line[4096];
  if (!std::fgets(line, sizeof line, fp)) { std::fclose(fp); err = "no banner"; return false; }
  char t[5][64] = {};
  if (std::sscanf(line, "%63s %63s %63s %63s %63s", t[0], t[1], t[2], t[3], t[4]) < 4) {
    std::fclose(fp); err = "bad banner"; return false;
  }
  const std::string field = lower(t[3]), sym = lower(t[4]);
  f.pattern = field == "pattern";
  f.symmetric = sym == "symmetric" || sym == "hermitian";
  do {
    if (!std::fgets(line, sizeof line, fp)) { std::fclose(fp); err = "no size line"; return false; }
  } while (line[0] == '%');
  if (std::sscanf(line, "%d %d %d", &f.m, &f.n, &f.nnz) != 3) { std::fclose(fp); err = "bad size"; return false; }
  f.r.resize(f.nnz); f.c.resize(f.nnz); f.v.assign(f.nnz, 1.0);
  for (int i = 0; i < f.nnz; i++) {
    int got = f.pattern ? std::fscanf(fp, "%d %d", &f.r[i], &f.c[i])
                        : std::fscanf(fp, "%d %d %lf", &f.r[i], &f.c[i], &f.v[i]);
    if (got != (f.pattern ? 2 : 3)) { std::fclose(fp); err = "truncated"; return false; }
    f.r[i]--; f.c[i]--;
  }
  std::fclose(fp);
  return true;
}

// p-ratio and Gini of a count vector (sort ascending; walk down from the top
// until p + (1-p) >= 1; Gini by the trapezoid over the cumulative sum).
void pratio_gini(const std::vector<int>& vec, int num_nnzs, double& p_ratio, double& gini) {
  std::vector<int> s = vec;
  std::sort(s.begin(), s.end());
  const int len = (int)s.size();
  int quit = len, step = 0, count = 0;
  double p = 0.0, q = 0.0;
  do {
    --quit; ++count;
    step += s[quit];
    p = (double)count / len;
    q = (double)step / num_nnzs;
  } while (p + q < 1.0);
  p_ratio = p;
  std::vector<int> cum(len);
  std::partial_sum(s.begin(), s.end(), cum.begin());
  double area = (cum[0] + 0) / 2.0;
  const double dx = (double)len / cum.size();
  for (size_t i = 0; i + 1 < cum.size(); ++i) area += (double)(cum[i] + cum[i + 1]) / 2.0 * dx;
  const double total = (double)num_nnzs * len / 2.0;
  gini = (total - area) / total;
}

struct Features {
  NamedArrays& out;
  explicit Features(NamedArrays& o) : out(o) {}
  void f(const std::string& name, double v) { out.put("f:" + name, &v, 1); }
};

int features(NamedArrays& res, const char* path) {
  MtxFile F;
  std::string err;
  if (!read_mtx(path, F, err)) { res.error = err; return 1; }
  const int M = F.m, N = F.n;
  const int T = 256;  // MAT_TILE_SIZE, include/general_config.h:8
  const int qR = M / T, modR = M % T, qC = N / T, modC = N % T;
  const int thrR = modR * (qR + 1), thrC = modC * (qC + 1);
  auto rb_of = [&](int r) { return r < thrR ? r / (qR + 1) : modR + (r - thrR) / qR; };
  auto cb_of = [&](int c) { return c < thrC ? c / (qC + 1) : modC + (c - thrC) / qC; };

  std::vector<int> row_cnt(M, 0), col_cnt(N, 0), tiles(T * T, 0), rbn(T, 0), cbn(T, 0), uniq_rb(T * T, 0),
      uniq_cb(T * T, 0);
  std::vector<int> rows_cnt((size_t)M * T, 0), cols_cnt((size_t)N * T, 0);
  std::vector<char> rows_seen((size_t)M * T, 0), cols_seen((size_t)N * T, 0);
  std::vector<double> diag_dom(M, 0.0), max_row(M, 0.0), min_row(M, 0.0), max_col(N, 0.0), min_col(N, 0.0);
  std::unordered_map<int, int> diag_offset;
  std::unordered_map<std::string, std::string> first_value;
  std::vector<std::string> pair_keys;
  long long nnz_lower = 0, nnz_upper = 0, nnz_diagonal = 0;
  double max_offdiag = -std::numeric_limits<double>::max(), max_diagonal = -std::numeric_limits<double>::max();

  auto tile_hit = [&](int row, int col, int tr, int tc) {
    const int tid = tr * T + tc;
    tiles[tid]++; rbn[tr]++; cbn[tc]++;
    rows_cnt[(size_t)row * T + tc]++;
    cols_cnt[(size_t)col * T + tr]++;
    if (!rows_seen[(size_t)row * T + tc]) { uniq_rb[tid]++; rows_seen[(size_t)row * T + tc] = 1; }
    if (!cols_seen[(size_t)col * T + tr]) { uniq_cb[tid]++; cols_seen[(size_t)col * T + tr] = 1; }
  };
  auto log_minmax = [](double& mx, double& mn, double a) {
    const double l = std::log10(a);
    if (mx < l) mx = l;
    if (a > 0.0 && mn > l) mn = l;
  };

  for (int i = 0; i < F.nnz; i++) {
    const int r = F.r[i], c = F.c[i];
    const double val = F.v[i], a = std::fabs(val);
    const int off = c - r;
    row_cnt[r]++; col_cnt[c]++; diag_offset[off]++;
    const int tr = rb_of(r), tc = cb_of(c);
    tile_hit(r, c, tr, tc);
    if (F.symmetric) {
      if (r == c) {
        nnz_diagonal++;
        diag_dom[r] += a;
        max_diagonal = F.pattern ? 1.0 : std::max(max_diagonal, a);
      } else {
        row_cnt[c]++; col_cnt[r]++; diag_offset[-off]++;
        // mirrored entry: the reference swaps the block indices it already
        // has (identical to recomputing them for square matrices)
        const int tid = tc * T + tr;
        tiles[tid]++; rbn[tc]++; cbn[tr]++;
        rows_cnt[(size_t)c * T + tr]++;
        cols_cnt[(size_t)r * T + tc]++;
        if (!rows_seen[(size_t)c * T + tr]) { uniq_rb[tid]++; rows_seen[(size_t)c * T + tr] = 1; }
        if (!cols_seen[(size_t)r * T + tc]) { uniq_cb[tid]++; cols_seen[(size_t)r * T + tc] = 1; }
        nnz_lower++; nnz_upper++;
        diag_dom[r] -= F.pattern ? 1.0 : a;
        diag_dom[c] -= F.pattern ? 1.0 : a;
        max_offdiag = F.pattern ? 1.0 : std::max(max_offdiag, a);
      }
      if (!F.pattern) {  // the pattern branch never updates the log ranges
        log_minmax(max_row[r], min_row[r], a);
        log_minmax(max_row[c], min_row[c], a);
        log_minmax(max_col[c], min_col[c], a);
        log_minmax(max_col[r], min_col[r], a);
      }
    } else {
      if (r == c) {
        nnz_diagonal++;
        max_diagonal = std::max(max_diagonal, a);
        diag_dom[r] += a;
      } else {
        if (r > c) nnz_lower++; else nnz_upper++;
        max_offdiag = std::max(max_offdiag, a);
        diag_dom[r] -= a;
      }
      char buf[100];
      std::snprintf(buf, sizeof buf, "%lf", val);
      const std::string key = std::to_string(r) + "_" + std::to_string(c);
      first_value.emplace(key, buf);  // keeps the first value of a duplicate
      pair_keys.push_back(key);
      if (!F.pattern) {
        log_minmax(max_row[r], min_row[r], a);
        log_minmax(max_col[c], min_col[c], a);
      }
    }
  }
  const long long num_nnzs = F.symmetric ? nnz_lower * 2 + nnz_diagonal : F.nnz;

  // per-tile and per-block images (the post pass of MtxLoad)
  std::vector<int> tile_max(T * T, 0), rb_max(T, 0), cb_max(T, 0);
  std::vector<double> tile_ave(T * T, 0.0), tile_std(T * T, 0.0), rb_ave(T, 0.0), rb_std(T, 0.0), cb_ave(T, 0.0),
      cb_std(T, 0.0);
  for (int i = 0; i < T * T; i++) {
    if (i < modR * T) tile_ave[i] = (double)tiles[i] / (qR + 1);
    else if (qR) tile_ave[i] = (double)tiles[i] / qR;
  }
  for (int r = 0; r < M; r++) {
    const int tr = rb_of(r);
    for (int tc = 0; tc < T; tc++) {
      const int tid = tr * T + tc;
      const int cnt = rows_cnt[(size_t)r * T + tc];
      tile_max[tid] = std::max(tile_max[tid], cnt);
      const double d = cnt - tile_ave[tid];
      tile_std[tid] += d * d;
    }
  }
  for (int i = 0; i < T * T; i++) {
    if (i < modR * T) tile_std[i] = tile_std[i] / (qR + 1);
    else if (qR) tile_std[i] = tile_std[i] / qR;
    tile_std[i] = std::sqrt(tile_std[i]);
  }
  for (int i = 0; i < T; i++) {
    if (i < modR) rb_ave[i] = (double)rbn[i] / (qR + 1);
    else if (qR) rb_ave[i] = (double)rbn[i] / qR;
    if (i < modC) cb_ave[i] = (double)cbn[i] / (qC + 1);
    else if (qC) cb_ave[i] = (double)cbn[i] / qC;
  }
  for (int i = 0; i < T; i++) {
    const int r0 = i < modR ? i * (qR + 1) : thrR + (i - modR) * qR;
    const int r1 = i < modR ? r0 + qR + 1 : r0 + qR;
    for (int r = r0; r < r1; r++) {
      int s = 0;
      for (int tc = 0; tc < T; tc++) s += rows_cnt[(size_t)r * T + tc];
      rb_max[i] = std::max(rb_max[i], s);
      const double d = s - rb_ave[i];
      rb_std[i] += d * d;
    }
    const int c0 = i < modC ? i * (qC + 1) : thrC + (i - modC) * qC;
    const int c1 = i < modC ? c0 + qC + 1 : c0 + qC;
    for (int c = c0; c < c1; c++) {
      int s = 0;
      for (int tr = 0; tr < T; tr++) s += cols_cnt[(size_t)c * T + tr];
      cb_max[i] = std::max(cb_max[i], s);
      const double d = s - cb_ave[i];
      cb_std[i] += d * d;
    }
  }
  for (int i = 0; i < T; i++) {
    if (i < modR) rb_std[i] = rb_std[i] / (qR + 1);
    else if (qR) rb_std[i] = rb_std[i] / qR;
    rb_std[i] = std::sqrt(rb_std[i]);
    if (i < modC) cb_std[i] = cb_std[i] / (qC + 1);
    else if (qC) cb_std[i] = cb_std[i] / qC;
    cb_std[i] = std::sqrt(cb_std[i]);
  }

  // CalculateFeatures
  Features out(res);
  long long complete_ndiags = 0, close_nnz = 0;
  const int close_thr = (M + N - 1) / 20;
  for (auto& kv : diag_offset) {
    if (std::abs(kv.first) <= close_thr) close_nnz += kv.second;
    complete_ndiags++;
  }
  const double nn = (double)num_nnzs;
  double row_div = -1.0, col_div = -1.0;
  int min_row_nnz = 100000000, max_row_nnz = 0, min_col_nnz = 100000000, max_col_nnz = 0;
  long long nz_rows = 0, nz_cols = 0, dominance = 0;
  const double ave_row = nn / M, ave_col = nn / N;
  double var_row = 0.0, var_col = 0.0;
  for (int i = 0; i < M; i++) {
    min_row_nnz = std::min(min_row_nnz, row_cnt[i]);
    max_row_nnz = std::max(max_row_nnz, row_cnt[i]);
    if (row_cnt[i]) nz_rows++;
    if (min_row[i] != 0.0) row_div = std::max(row_div, max_row[i] - min_row[i]);
    const double d = row_cnt[i] - ave_row;
    var_row += d * d;
    if (diag_dom[i] > 0) dominance++;
  }
  var_row /= M;
  double p_row, g_row, p_col, g_col;
  pratio_gini(row_cnt, (int)num_nnzs, p_row, g_row);
  for (int j = 0; j < N; j++) {
    min_col_nnz = std::min(min_col_nnz, col_cnt[j]);
    max_col_nnz = std::max(max_col_nnz, col_cnt[j]);
    if (col_cnt[j]) nz_cols++;
    if (min_col[j] != 0.0) col_div = std::max(col_div, max_col[j] - min_col[j]);
    const double d = col_cnt[j] - ave_col;
    var_col += d * d;
  }
  var_col /= N;
  pratio_gini(col_cnt, (int)num_nnzs, p_col, g_col);

  double pattern_symm = 0.0, value_symm = 0.0;
  if (F.symmetric) {
    pattern_symm = value_symm = 1.0;
  } else if (M == N) {
    long long sp = 0, sv = 0;
    for (int i = 0; i < F.nnz; i++) {
      const int r = F.r[i], c = F.c[i];
      if (r == c) continue;
      auto it = first_value.find(std::to_string(c) + "_" + std::to_string(r));
      if (it != first_value.end()) {
        sp++;
        if (it->second == first_value[pair_keys[i]]) sv++;
      }
    }
    pattern_symm = (double)sp / (F.nnz - nnz_diagonal);
    value_symm = (double)sv / (F.nnz - nnz_diagonal);
  }
  const bool is_symmetric = pattern_symm == 1.0 && value_symm == 1.0;

  // CalculateTilesFeatures (the -1.0 initial values of the t_var / t_nz_ratio
  // members are part of the result)
  const double t_ave_tiles = nn / ((double)T * T), t_ave_rb = nn / T, t_ave_cb = nn / T;
  double nz_rb = -1.0, nz_cb = -1.0, var_rb = -1.0, var_cb = -1.0, nz_t = -1.0, var_t = -1.0;
  int min_rb = 100000000, max_rb = 0, min_cb = 100000000, max_cb = 0, min_t = 100000000, max_t = 0;
  for (int i = 0; i < T; i++) {
    if (rbn[i]) ++nz_rb;
    if (cbn[i]) ++nz_cb;
    const double dr = rbn[i] - t_ave_rb, dc = cbn[i] - t_ave_cb;
    var_rb += dr * dr;
    var_cb += dc * dc;
    min_rb = std::min(min_rb, rbn[i]); max_rb = std::max(max_rb, rbn[i]);
    min_cb = std::min(min_cb, cbn[i]); max_cb = std::max(max_cb, cbn[i]);
  }
  nz_rb /= (double)T; var_rb /= (double)T; nz_cb /= (double)T; var_cb /= (double)T;
  for (int i = 0; i < T * T; i++) {
    if (tiles[i]) ++nz_t;
    const double d = tiles[i] - t_ave_tiles;
    var_t += d * d;
    min_t = std::min(min_t, tiles[i]); max_t = std::max(max_t, tiles[i]);
  }
  nz_t /= ((double)T * T); var_t /= ((double)T * T);
  double p_t, g_t, p_rb, g_rb, p_cb, g_cb;
  pratio_gini(tiles, (int)num_nnzs, p_t, g_t);
  pratio_gini(rbn, (int)num_nnzs, p_rb, g_rb);
  pratio_gini(cbn, (int)num_nnzs, p_cb, g_cb);
  const int sum_urb = std::accumulate(uniq_rb.begin(), uniq_rb.end(), 0);
  const int sum_ucb = std::accumulate(uniq_cb.begin(), uniq_cb.end(), 0);

  // CalculateTilesExtraFeatures: mean over rows of (sum of column gaps in the
  // CSR read back from the file) / row length
  std::vector<int> er, ec;
  for (int i = 0; i < F.nnz; i++) {
    er.push_back(F.r[i]); ec.push_back(F.c[i]);
    if (F.symmetric && F.r[i] != F.c[i]) { er.push_back(F.c[i]); ec.push_back(F.r[i]); }
  }
  std::vector<int> rp(M + 1, 0), ci(er.size());
  for (int r : er) rp[r + 1]++;
  for (int i = 0; i < M; i++) rp[i + 1] += rp[i];
  {
    std::vector<int> fill(rp.begin(), rp.end() - 1);
    for (size_t k = 0; k < er.size(); k++) ci[fill[er[k]]++] = ec[k];
  }
  double distance = 0.0;
  {
    std::vector<int> dis(M, 0);
    for (int i = 0; i < M; i++)
      for (int j = rp[i]; j + 1 < rp[i + 1]; j++) dis[i] += ci[j + 1] - ci[j];
    for (int i = 0; i < M; i++)
      if (rp[i + 1] - rp[i]) distance += (double)dis[i] / (rp[i + 1] - rp[i]);
    distance = distance / M;
  }

  out.f("num_rows", M); out.f("num_cols", N); out.f("num_nnzs", (double)num_nnzs); out.f("nnz_mtx", F.nnz);
  out.f("nnz_lower", (double)nnz_lower); out.f("nnz_upper", (double)nnz_upper);
  out.f("nnz_diagonal", (double)nnz_diagonal); out.f("complete_ndiags", (double)complete_ndiags);
  out.f("close_nnz", (double)close_nnz);
  out.f("min_nnz_each_row", min_row_nnz); out.f("max_nnz_each_row", max_row_nnz);
  out.f("min_nnz_each_col", min_col_nnz); out.f("max_nnz_each_col", max_col_nnz);
  out.f("is_symmetric", is_symmetric ? 1.0 : 0.0);
  out.f("pattern_symm", pattern_symm); out.f("value_symm", value_symm);
  out.f("nnz_ratio", nn / ((double)M * N)); out.f("diag_close_ratio", (double)close_nnz / nn);
  out.f("distance_per_row", distance);
  out.f("nz_row_ratio", (double)nz_rows / M); out.f("ave_nnz_each_row", ave_row);
  out.f("var_nnz_each_row", var_row); out.f("standard_dev_row", std::sqrt(var_row));
  out.f("Gini_row", g_row); out.f("P_ratio_row", p_row);
  out.f("nz_col_ratio", (double)nz_cols / N); out.f("ave_nnz_each_col", ave_col);
  out.f("var_nnz_each_col", var_col); out.f("standard_dev_col", std::sqrt(var_col));
  out.f("Gini_col", g_col); out.f("P_ratio_col", p_col);
  out.f("diagonal_dominant_ratio", (double)dominance / std::min(M, N));
  out.f("diag_dominant_rows", (double)dominance);
  out.f("max_value_offdiag", max_offdiag); out.f("max_value_diagonal", max_diagonal);
  out.f("row_variability", row_div); out.f("col_variability", col_div);
  out.f("t_num_blocks", T); out.f("t_num_RB", qR); out.f("t_num_CB", qC); out.f("t_mod_RB", modR);
  out.f("t_mod_CB", modC);
  out.f("t_ave_nnz_all_tiles", t_ave_tiles); out.f("t_ave_nnz_RB", t_ave_rb); out.f("t_ave_nnz_CB", t_ave_cb);
  out.f("t_var_nnz_all_tiles", var_t); out.f("t_var_nnz_RB", var_rb); out.f("t_var_nnz_CB", var_cb);
  out.f("t_standard_dev_all_tiles", std::sqrt(std::max(var_t, 0.0)));
  out.f("t_standard_dev_RB", std::sqrt(std::max(var_rb, 0.0)));
  out.f("t_standard_dev_CB", std::sqrt(std::max(var_cb, 0.0)));
  out.f("t_min_nnz_all_tiles", min_t); out.f("t_max_nnz_all_tiles", max_t);
  out.f("t_min_nnz_each_RB", min_rb); out.f("t_max_nnz_each_RB", max_rb);
  out.f("t_min_nnz_each_CB", min_cb); out.f("t_max_nnz_each_CB", max_cb);
  out.f("t_Gini_all_tiles", g_t); out.f("t_Gini_RB", g_rb); out.f("t_Gini_CB", g_cb);
  out.f("t_P_ratio_all_tiles", p_t); out.f("t_P_ratio_RB", p_rb); out.f("t_P_ratio_CB", p_cb);
  out.f("t_nz_ratio_tiles", nz_t); out.f("t_nz_ratio_RB", nz_rb); out.f("t_nz_ratio_CB", nz_cb);
  out.f("uniqR", (double)sum_urb / num_nnzs); out.f("uniqC", (double)sum_ucb / num_nnzs);
  out.f("potReuseR", (double)sum_urb / M); out.f("potReuseC", (double)sum_ucb / N);
  out.f("GrX_uniqR", -1.0); out.f("GrX_uniqC", -1.0); out.f("GrX_potReuseR", -1.0); out.f("GrX_potReuseC", -1.0);

  res.put("img:tile_nnz", tiles); res.put("img:tile_max", tile_max); res.put("img:tile_ave", tile_ave);
  res.put("img:tile_std", tile_std); res.put("img:tile_uniq_rows", uniq_rb); res.put("img:tile_uniq_cols", uniq_cb);
  res.put("img:rb_nnz", rbn); res.put("img:rb_max", rb_max); res.put("img:rb_ave", rb_ave); res.put("img:rb_std", rb_std);
  res.put("img:cb_nnz", cbn); res.put("img:cb_max", cb_max); res.put("img:cb_ave", cb_ave); res.put("img:cb_std", cb_std);
  res.put("img:row_nnz", row_cnt); res.put("img:col_nnz", col_cnt);
  return 0;
}

}  // namespace

extern "C" int oracle_features(void* res, const char* path) {
  return features(*static_cast<NamedArrays*>(res), path);
}
