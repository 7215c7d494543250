// The DIA row engine of the per-call DIA SpMM (spmv_dia.hip): rows of a DIA
// matrix times k row-major right-hand sides, Y = alpha*A*X + beta*Y, summed
// in FP64 for FP32 data too (one rounding, at the store). A row's sum runs
// over the diagonals in storage order and depends on nothing else.
//
// Diagonal d stores its entry of row i at data[d * stride + i] (zero where
// the diagonal has no entry); its column is i + offs[d]. Against CSR there
// are no column indices to load, every address is known before the first
// load returns, and the diagonal's values of consecutive rows are
// consecutive in memory.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "rows_engine.hpp"

namespace lsp {

constexpr int kDiaChains = 8;  // diagonals whose loads are issued together

// Rows in one work item: passes of 64 / kk rows (kk lanes cover k right-hand
// sides at W per lane), enough passes for about `iters` loads per lane.
inline int dia_item_rows(int ndiag, int k, int W, int iters) {
  const int kw = (k + W - 1) / W;
  int kk = 1;
  while (kk < kw) kk <<= 1;
  const int passes = (int)std::clamp(std::ceil((double)iters / std::max(ndiag, 1)), 1.0, (double)kMaxPasses);
  return passes * (64 / kk);
}

// Rows [r0, r1) for all k right-hand sides. A row gets kk lanes (the lanes
// that cover k at W right-hand sides each, rounded up to a power of two);
// lane t of the row handles right-hand sides t*W .. t*W + W - 1 and every
// diagonal, so no lanes combine partial sums. The value and X loads of
// kDiaChains diagonals are issued before the first of their products.
template <class V, int W, bool VEC>
__device__ __forceinline__ void dia_rows(const int64_t* __restrict__ offs, const V* __restrict__ data,
                                         int64_t stride, int nd, int64_t n, int k, const V* __restrict__ X,
                                         int64_t ldx, V* __restrict__ Y, int64_t ldy, double alpha, double beta,
                                         int r0, int r1, int lane) {
  constexpr int CH = kDiaChains;
  const int kw = (k + W - 1) / W;
  int kk = 1;
  while (kk < kw) kk <<= 1;
  const int P = 64 / kk;  // rows per pass
  const int grp = lane / kk, j = (lane % kk) * W;
  const int nvalid = k - j;
  const LSP_GLOBAL_AS V* __restrict__ Xg = (const LSP_GLOBAL_AS V*)X + j;
  for (int base = r0; base < r1; base += P) {
    const int64_t row = base + grp;
    const bool ok = row < r1 && j < k;
    double acc[W];
#pragma unroll
    for (int u = 0; u < W; u++) acc[u] = 0;
    if (ok) {
      for (int d0 = 0; d0 < nd; d0 += CH) {
        V vv[CH], xx[CH][W];
#pragma unroll
        for (int c = 0; c < CH; c++) {
          const int d = d0 + c;
          const int64_t col = d < nd ? row + offs[d] : -1;
          const bool in = col >= 0 && col < n;
          vv[c] = in ? data[(int64_t)d * stride + row] : (V)0;
          if (in) {
            load_w<V, W, VEC>(Xg + col * ldx, nvalid, xx[c]);
          } else {
#pragma unroll
            for (int u = 0; u < W; u++) xx[c][u] = 0;
          }
        }
#pragma unroll
        for (int c = 0; c < CH; c++)
#pragma unroll
          for (int u = 0; u < W; u++) acc[u] += (double)vv[c] * (double)xx[c][u];
      }
      V* y = Y + row * ldy + j;
      if constexpr (VEC && W > 1) {
        typedef V vec_t __attribute__((ext_vector_type(W)));
        vec_t o;
        if (beta == 0.0) {
#pragma unroll
          for (int u = 0; u < W; u++) o[u] = (V)(alpha * acc[u]);
        } else {
          const vec_t old = *(const vec_t*)y;
#pragma unroll
          for (int u = 0; u < W; u++) o[u] = (V)(alpha * acc[u] + beta * (double)old[u]);
        }
        *(vec_t*)y = o;
      } else {
#pragma unroll
        for (int u = 0; u < W; u++)
          if (u < nvalid) y[u] = (V)(beta == 0.0 ? alpha * acc[u] : alpha * acc[u] + beta * (double)y[u]);
      }
    }
  }
}

}  // namespace lsp
