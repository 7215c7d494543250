// Exhaustive timing and the autotuner.
//
// No reference counterpart: the reference times one format per driver run
// (test/benchmark_spmv_<fmt>.cpp) and its selector learns from spreadsheets of
// those runs. Here one call times a list of (format, kernel variant,
// conversion parameters) candidates on the device, which is both the label
// source for the learned selector and the tuner that picks SELL C / sigma, the
// BSR block size, CSR lanes per row (rows per wave = 64 / lanes), CSR5 omega /
// sigma and the DIA / COO kernel variant.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "common.hpp"

namespace lsp {
extern thread_local std::string g_last_error;
}  // namespace lsp

using namespace lsp;

namespace {

lespmv_tune_result_t candidate(int fmt, int variant, lespmv_params_t p) {
  lespmv_tune_result_t r{};
  r.fmt = fmt;
  r.variant = variant;
  r.params = p;
  r.status = LESPMV_OK;
  r.ms = std::numeric_limits<double>::infinity();
  r.convert_ms = 0.0;
  return r;
}

// The candidate grid (fill limits keep padded layouts from exploding).
std::vector<lespmv_tune_result_t> grid(int32_t fmt_mask, int dtype) {
  std::vector<lespmv_tune_result_t> c;
  auto want = [&](int f) { return (fmt_mask >> f) & 1; };
  lespmv_params_t base{};
  base.alignment = 1;
  base.max_diags = 10240;  // MAX_DIAG_NUM, general_config.h:9
  base.max_fill = 4;
  (void)dtype;
  if (want(LESPMV_CSR)) {
    for (int lg = 1; lg <= 6; lg++) c.push_back(candidate(LESPMV_CSR, 16 + lg, base));  // 2..64 lanes per row
    c.push_back(candidate(LESPMV_CSR, LESPMV_VARIANT_MERGE, base));
  }
  if (want(LESPMV_COO)) c.push_back(candidate(LESPMV_COO, LESPMV_VARIANT_AUTO, base));
  if (want(LESPMV_ELL)) c.push_back(candidate(LESPMV_ELL, LESPMV_VARIANT_AUTO, base));
  for (int C : {32, 64}) {
    lespmv_params_t p = base;
    p.chunk = C;
    if (want(LESPMV_SELL)) c.push_back(candidate(LESPMV_SELL, LESPMV_VARIANT_AUTO, p));
    if (want(LESPMV_SELL_C_SIGMA))
      for (int sg : {256, 4096}) {
        p.sigma = sg;
        c.push_back(candidate(LESPMV_SELL_C_SIGMA, LESPMV_VARIANT_AUTO, p));
      }
    if (want(LESPMV_SELL_C_R)) c.push_back(candidate(LESPMV_SELL_C_R, LESPMV_VARIANT_AUTO, p));
  }
  if (want(LESPMV_DIA)) {
    c.push_back(candidate(LESPMV_DIA, LESPMV_VARIANT_VECTOR, base));
    c.push_back(candidate(LESPMV_DIA, LESPMV_VARIANT_MERGE, base));
  }
  if (want(LESPMV_BSR))
    for (auto bs : {std::pair<int, int>{2, 2}, {4, 4}, {8, 4}, {8, 8}, {16, 4}}) {
      lespmv_params_t p = base;
      p.block_rows = bs.first;
      p.block_cols = bs.second;
      c.push_back(candidate(LESPMV_BSR, LESPMV_VARIANT_AUTO, p));
    }
  if (want(LESPMV_CSR5))
    for (auto ws : {std::pair<int, int>{64, 16}, {32, 16}, {64, 8}}) {
      lespmv_params_t p = base;
      p.omega = ws.first;
      p.sigma = ws.second;
      c.push_back(candidate(LESPMV_CSR5, LESPMV_VARIANT_AUTO, p));
    }
  // the CSR column panels (8 and 4 panels), last so that the other candidates
  // keep their places in the grid
  if ((fmt_mask & LESPMV_TUNE_PANELS) != 0) {
    c.push_back(candidate(LESPMV_CSR, LESPMV_VARIANT_PANEL, base));
    c.push_back(candidate(LESPMV_CSR, LESPMV_VARIANT_PANEL4, base));
  }
  return c;
}

bool is_panel(const lespmv_tune_result_t& r) {
  return r.fmt == LESPMV_CSR && (r.variant == LESPMV_VARIANT_PANEL || r.variant == LESPMV_VARIANT_PANEL16 ||
                                 r.variant == LESPMV_VARIANT_PANEL4);
}

// Converts (unless CSR), times `iters` calls after two warm-up calls, and
// releases the format again unless `keep`.
void time_one(lespmv_matrix_t h, lespmv_tune_result_t& r, int32_t iters, int32_t k, void* x, void* y, int64_t n,
              int64_t m, bool keep) {
  r.ms = std::numeric_limits<double>::infinity();
  r.convert_ms = 0.0;
  if (r.fmt != LESPMV_CSR) {
    r.status = lespmv_convert(h, r.fmt, &r.params);
    if (r.status != LESPMV_OK) return;
    lespmv_convert_ms(h, r.fmt, &r.convert_ms);
  } else if (is_panel(r)) {  // the panel layout, built and timed like a conversion
    r.status = lespmv_prepare(h, LESPMV_CSR, r.variant);
    if (r.status != LESPMV_OK) return;
    lespmv_convert_ms(h, LESPMV_CSR, &r.convert_ms);
  }
  double ms = 0.0;
  int rc;
  if (k <= 1) {
    rc = lespmv_time_spmv(h, r.fmt, r.variant, 1.0, x, 0.0, y, 2, &ms);
    if (rc == LESPMV_OK) rc = lespmv_time_spmv(h, r.fmt, r.variant, 1.0, x, 0.0, y, iters, &ms);
  } else {
    rc = lespmv_time_spmm(h, r.fmt, r.variant, k, 1.0, x, k, 0.0, y, k, 2, &ms);
    if (rc == LESPMV_OK) rc = lespmv_time_spmm(h, r.fmt, r.variant, k, 1.0, x, k, 0.0, y, k, iters, &ms);
  }
  (void)n;
  (void)m;
  r.status = rc;
  if (rc == LESPMV_OK) r.ms = ms;
  if (!keep && (r.fmt != LESPMV_CSR || is_panel(r))) lespmv_release_format(h, r.fmt);
}

}  // namespace

extern "C" int lespmv_tune_candidates(int32_t fmt_mask, int32_t dtype, lespmv_tune_result_t* out, int32_t cap,
                                      int32_t* count) {
  return guarded_call([&] {
    LSP_REQUIRE(count, LESPMV_ERR_ARG, "null count");
    const auto c = grid(fmt_mask, dtype);
    *count = (int32_t)c.size();
    if (out) {
      LSP_REQUIRE(cap >= (int32_t)c.size(), LESPMV_ERR_ARG, "candidate buffer too small");
      std::copy(c.begin(), c.end(), out);
    }
  });
}

extern "C" int lespmv_tune(lespmv_matrix_t h, lespmv_tune_result_t* cand, int32_t n, int32_t iters, int32_t k) {
  return guarded_call([&] {
    LSP_REQUIRE(h && cand && n >= 0 && iters > 0 && k >= 1 && k <= 32, LESPMV_ERR_ARG, "bad tune arguments");
    int64_t m = 0, ncols = 0, nnz = 0;
    int32_t dtype = 0;
    LSP_REQUIRE(lespmv_matrix_info(h, &m, &ncols, &nnz, &dtype) == LESPMV_OK, LESPMV_ERR_ARG, "bad handle");
    const int64_t vs = dtype == LESPMV_F64 ? 8 : 4;
    DBuf x, y;
    x.alloc(vs * std::max<int64_t>(ncols, 1) * k);
    y.alloc(vs * std::max<int64_t>(m, 1) * k);
    // x = 1 everywhere (bit pattern of 1.0 in the value type)
    std::vector<unsigned char> ones(vs * std::max<int64_t>(ncols, 1) * k);
    for (int64_t i = 0; i < (int64_t)ones.size() / vs; i++) {
      if (vs == 8) {
        const double one = 1.0;
        std::memcpy(&ones[i * 8], &one, 8);
      } else {
        const float one = 1.0f;
        std::memcpy(&ones[i * 4], &one, 4);
      }
    }
    LSP_HIP(hipMemcpy(x.p, ones.data(), ones.size(), hipMemcpyHostToDevice));
    for (int32_t i = 0; i < n; i++) time_one(h, cand[i], iters, k, x.p, y.p, ncols, m, false);
    g_last_error.clear();
  });
}

extern "C" int lespmv_autotune(lespmv_matrix_t h, int32_t fmt, int32_t iters, int32_t k, lespmv_tune_result_t* best) {
  return guarded_call([&] {
    LSP_REQUIRE(h && best && fmt >= 0 && fmt < LESPMV_NUM_FORMATS, LESPMV_ERR_ARG, "bad autotune arguments");
    int64_t m = 0, n = 0, nnz = 0;
    int32_t dtype = 0;
    LSP_REQUIRE(lespmv_matrix_info(h, &m, &n, &nnz, &dtype) == LESPMV_OK, LESPMV_ERR_ARG, "bad handle");
    auto c = grid(1 << fmt, dtype);
    LSP_REQUIRE(!c.empty(), LESPMV_ERR_UNSUPPORTED, "no candidates for this format");
    const int rc = lespmv_tune(h, c.data(), (int32_t)c.size(), iters, k);
    LSP_REQUIRE(rc == LESPMV_OK, rc, "tuning failed");
    const auto it = std::min_element(c.begin(), c.end(), [](const auto& a, const auto& b) { return a.ms < b.ms; });
    LSP_REQUIRE(std::isfinite(it->ms), LESPMV_ERR_UNSUPPORTED, "no candidate of this format applies to the matrix");
    *best = *it;
    if (fmt != LESPMV_CSR) {  // leave the winner converted on the handle
      const int rc2 = lespmv_convert(h, fmt, &best->params);
      LSP_REQUIRE(rc2 == LESPMV_OK, rc2, "re-converting the tuned format failed");
    }
  });
}
