// The learned format selector: GPU features -> 128 inputs -> MLP (mlp.hip)
// -> (format class, parameter) -> a candidate of the standard tuning grid
// (tune.cpp). No reference counterpart; the reference's selector reads the
// 40 .features values plus tile images into a Keras model
// (SmartAdpter/models/SpMV_Adapter.py:29-116) and predicts one of the nine
// format labels (LeSpMV/script/write_label.py:63-75).
//
// Heads: 0 = format class (CSR-vector, CSR-merge, COO, ELL, SELL family, DIA,
// BSR, CSR5, CSR column panels), 1 = CSR lanes per row (2..64), 2 = SELL
// (C, sigma): C in {32, 64} x {S-ELL, sigma 256, sigma 4096, SELL-C-R},
// 3 = DIA variant, 4 = BSR block (2x2, 4x4, 8x4, 8x8, 16x4), 5 = CSR5
// (omega, sigma), 6 = panels (8, 4). A class's parameter head indexes its run
// of candidates in the grid (all formats plus LESPMV_TUNE_PANELS).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace lsp {
extern thread_local std::string g_last_error;
}  // namespace lsp

using namespace lsp;

namespace {

constexpr int kClasses = 9, kHeads = 7, kOut = LESPMV_SELECTOR_OUTPUTS;
// first grid index of each class (the grid order of lespmv_tune_candidates with
// all formats and LESPMV_TUNE_PANELS)
constexpr int kClassFirst[kClasses + 1] = {0, 6, 7, 8, 9, 17, 19, 24, 27, 29};
constexpr int kHeadOfClass[kClasses] = {1, -1, -1, -1, 2, 3, 4, 5, 6};
constexpr int kHeadSizes[kHeads] = {9, 6, 8, 2, 5, 3, 2};
constexpr int kHeadOff[kHeads] = {0, 9, 15, 23, 25, 30, 33};
static_assert(kHeadOff[kHeads - 1] + kHeadSizes[kHeads - 1] == kOut, "head layout");
constexpr int kBlockDims[3][2] = {{2, 2}, {4, 4}, {8, 8}};
constexpr int kBsrKnown[3] = {0, 1, 3};  // positions of 2x2, 4x4, 8x8 in the BSR head
constexpr double kMaxFill = 4.0;          // the grid's params.max_fill
constexpr int64_t kMaxDiags = 10240;      // MAX_DIAG_NUM, general_config.h:9

float lg(double x) { return (float)std::log2(1.0 + std::max(x, 0.0)); }

void inputs_from(const lespmv_features_t& f, const int32_t* rb, const int32_t* cb, const int64_t* nnzb, float* o,
                 uint8_t* allowed) {
  double v[LESPMV_NUM_FEATURES];
  lespmv_features_vector(&f, v);
  const double nnz = (double)std::max<int64_t>(f.num_nnzs, 1), M = (double)f.num_rows, N = (double)f.num_cols;
  // the 40 .features values: sizes and spreads on a log scale, ratios as they are
  static const bool logscale[LESPMV_NUM_FEATURES] = {
      1, 1, 1, 0, 0, 0,  // rows, cols, nnz, nnz_ratio (below), diag ratio, distance ratio
      0, 1, 1, 1, 0, 0,  // row: nz ratio, max, ave, std, p, gini
      0, 1, 1, 1, 0, 0,  // col
      0, 1, 1, 1, 0, 0,  // tiles
      0, 1, 1, 1, 0, 0,  // RB
      0, 1, 1, 1, 0, 0,  // CB
      0, 0, 1, 1};       // uniqR, uniqC, potReuseR, potReuseC
  for (int i = 0; i < LESPMV_NUM_FEATURES; i++) o[i] = logscale[i] ? lg(v[i]) : (float)v[i];
  o[3] = (float)std::log10(f.nnz_ratio + 1e-12);
  float* e = o + 40;
  e[0] = lg((double)f.complete_ndiags);
  e[1] = lg((double)f.complete_ndiags * M / nnz);
  e[2] = lg((double)f.max_nnz_each_row * M / nnz);
  e[3] = (float)(f.nnz_lower / nnz);
  e[4] = (float)(f.nnz_upper / nnz);
  e[5] = (float)(f.nnz_diagonal / std::max(1.0, std::min(M, N)));
  e[6] = lg((double)f.min_nnz_each_row);
  e[7] = lg((double)f.min_nnz_each_col);
  e[8] = (float)f.diagonal_dominant_ratio;
  e[9] = (float)f.pattern_symm;
  e[10] = (float)f.value_symm;
  e[11] = (float)f.is_symmetric;
  e[12] = (float)std::clamp(f.row_variability, -1.0, 30.0);
  e[13] = (float)std::clamp(f.col_variability, -1.0, 30.0);
  e[14] = lg((double)f.max_nnz_each_row / std::max(f.ave_nnz_each_row, 1e-9));
  e[15] = lg((double)f.max_nnz_each_col / std::max(f.ave_nnz_each_col, 1e-9));
  e[16] = (float)(1.0 - f.nz_row_ratio);
  e[17] = (float)std::log2(M / N);
  for (int b = 0; b < 3; b++) e[18 + b] = lg((double)nnzb[b] * kBlockDims[b][0] * kBlockDims[b][1] / nnz);
  e[21] = (float)(nnz / std::max<double>((double)nnzb[1] * 16.0, 1.0));
  e[22] = lg(f.ave_nnz_each_row / 64.0);
  e[23] = (float)(f.standard_dev_row / std::max(f.ave_nnz_each_row, 1e-9));
  // row-block and column-block nonzero shares, 256 blocks folded into 32 bins
  for (int b = 0; b < 32; b++) {
    double r = 0, c = 0;
    for (int k = 0; k < 8; k++) {
      r += rb[b * 8 + k];
      c += cb[b * 8 + k];
    }
    o[64 + b] = (float)(r / nnz);
    o[96 + b] = (float)(c / nnz);
  }
  // an empty matrix leaves 0/0 ratios (NaN, as in the reference's features):
  // the network sees 0 there
  for (int i = 0; i < LESPMV_SELECTOR_INPUTS; i++)
    if (!std::isfinite(o[i])) o[i] = 0.f;
  // what can be ruled out from the features (the grid's fill and diagonal limits)
  std::fill(allowed, allowed + kOut, (uint8_t)1);
  const double limit = kMaxFill * nnz + M;
  if ((double)f.max_nnz_each_row * M > limit) allowed[3] = 0;  // ELL
  if (f.complete_ndiags > kMaxDiags || (double)f.complete_ndiags * M > limit) allowed[5] = 0;  // DIA
  if (f.num_nnzs == 0) allowed[7] = 0;  // CSR5
  for (int b = 0; b < 3; b++)
    if ((double)nnzb[b] * kBlockDims[b][0] * kBlockDims[b][1] > limit) allowed[kHeadOff[4] + kBsrKnown[b]] = 0;
}

void compute_inputs(lespmv_matrix_t h, int32_t flags, float* out, uint8_t* allowed, lespmv_features_t* fo) {
  lespmv_features_t f{};
  std::vector<int32_t> rb(LESPMV_FEAT_TILES), cb(LESPMV_FEAT_TILES);
  lespmv_feature_images_t img{};
  img.rb_nnz = rb.data();
  img.cb_nnz = cb.data();
  int rc = lespmv_features(h, flags, &f, &img);
  LSP_REQUIRE(rc == LESPMV_OK, rc, "feature extraction failed");
  int64_t nnzb[3];
  const int32_t dims[6] = {2, 2, 4, 4, 8, 8};
  rc = lespmv_block_counts(h, dims, 3, nnzb);
  LSP_REQUIRE(rc == LESPMV_OK, rc, "block counts failed");
  inputs_from(f, rb.data(), cb.data(), nnzb, out, allowed);
  if (fo) *fo = f;
}

int masked_argmax(const float* p, const uint8_t* allowed, int n) {
  int best = -1;
  for (int i = 0; i < n; i++)
    if (allowed[i] && (best < 0 || p[i] > p[best])) best = i;
  return best;
}

}  // namespace

extern "C" int lespmv_selector_layout(int32_t* class_first, int32_t* head_of_class, int32_t* head_sizes) {
  return guarded_call([&] {
    if (class_first) std::copy(kClassFirst, kClassFirst + kClasses + 1, class_first);
    if (head_of_class) std::copy(kHeadOfClass, kHeadOfClass + kClasses, head_of_class);
    if (head_sizes) std::copy(kHeadSizes, kHeadSizes + kHeads, head_sizes);
  });
}

extern "C" int lespmv_selector_inputs(lespmv_matrix_t h, int32_t flags, float* out, uint8_t* allowed,
                                      lespmv_features_t* features) {
  return guarded_call([&] {
    LSP_REQUIRE(h && out && allowed, LESPMV_ERR_ARG, "null argument");
    compute_inputs(h, flags, out, allowed, features);
  });
}

extern "C" int lespmv_selector_decode(const float* probs, const uint8_t* allowed, int32_t* cls, int32_t* candidate) {
  return guarded_call([&] {
    LSP_REQUIRE(probs && allowed && cls && candidate, LESPMV_ERR_ARG, "null argument");
    // a class whose parameter head is fully ruled out is ruled out too
    uint8_t cmask[kClasses];
    for (int c = 0; c < kClasses; c++) {
      cmask[c] = allowed[c];
      const int hh = kHeadOfClass[c];
      if (hh >= 0) {
        bool any = false;
        for (int j = 0; j < kHeadSizes[hh]; j++) any |= allowed[kHeadOff[hh] + j] != 0;
        if (!any) cmask[c] = 0;
      }
    }
    int c = masked_argmax(probs, cmask, kClasses);
    if (c < 0) c = 1;  // nothing allowed: CSR merge-path
    int j = 0;
    const int hh = kHeadOfClass[c];
    if (hh >= 0) j = std::max(0, masked_argmax(probs + kHeadOff[hh], allowed + kHeadOff[hh], kHeadSizes[hh]));
    *cls = c;
    *candidate = kClassFirst[c] + j;
  });
}

extern "C" int lespmv_select(lespmv_matrix_t h, lespmv_mlp_t mlp, int32_t flags, lespmv_tune_result_t* choice,
                             int32_t* cls) {
  return guarded_call([&] {
    LSP_REQUIRE(h && mlp && choice, LESPMV_ERR_ARG, "null argument");
    int32_t in_dim = 0, out_dim = 0;
    int rc = lespmv_mlp_shape(mlp, &in_dim, &out_dim, nullptr);
    LSP_REQUIRE(rc == LESPMV_OK && in_dim == LESPMV_SELECTOR_INPUTS && out_dim == kOut, LESPMV_ERR_ARG,
                "the network does not have the selector's shape");
    float in[LESPMV_SELECTOR_INPUTS], probs[kOut];
    uint8_t allowed[kOut];
    compute_inputs(h, flags, in, allowed, nullptr);
    rc = lespmv_mlp_forward(mlp, in, 1, probs);
    LSP_REQUIRE(rc == LESPMV_OK, rc, "selector forward pass failed");
    int32_t c = 0, cand = 0;
    rc = lespmv_selector_decode(probs, allowed, &c, &cand);
    LSP_REQUIRE(rc == LESPMV_OK, rc, "decode failed");
    int64_t m = 0, n = 0, nnz = 0;
    int32_t dtype = 0;
    lespmv_matrix_info(h, &m, &n, &nnz, &dtype);
    int32_t count = 0;
    const int32_t mask = ((1 << LESPMV_NUM_FORMATS) - 1) | LESPMV_TUNE_PANELS;
    lespmv_tune_candidates(mask, dtype, nullptr, 0, &count);
    std::vector<lespmv_tune_result_t> grid(count);
    lespmv_tune_candidates(mask, dtype, grid.data(), count, &count);
    LSP_REQUIRE(cand >= 0 && cand < count, LESPMV_ERR_DEVICE, "selector layout and tuning grid disagree");
    *choice = grid[cand];
    if (cls) *cls = c;
  });
}
