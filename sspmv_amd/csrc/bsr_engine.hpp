// BSR 4x4 FP64 SpMM on the matrix cores, shared by the per-call BSR SpMM
// (spmv_bsr.hip) and the packed serving kernel (serve.hip), so a request
// served in a batch and the same product launched on its own give bitwise
// equal results.
//
// One wavefront per block row. v_mfma_f64_4x4x4f64 runs four independent
// 4x4x4 products; product q uses the lanes 4q + x + 16y (x, y in 0..3) with
// A[i][k] at (x, y) = (i, k), B[k][j] at (j, k) and D[i][j] at (j, i)
// (measured on gfx950). The four products share the block (A) and take four
// consecutive groups of 4 right-hand sides, so one instruction covers 16
// right-hand sides; k <= 64 takes up to four instructions per block.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lsp {

constexpr int kBsrSteps = 4;  // blocks whose loads are issued before their products

// Block row `brow` of a 4x4 BSR matrix (rptr, bcol, data: 16 values per block,
// row-major) times k <= 64 row-major right-hand sides: Y = alpha*A*X +
// beta*Y for the block row's rows below m; block columns reaching past n read
// zeros. The loads of kBsrSteps blocks are issued before the first of their
// products; the products accumulate block by block in storage order, so the
// sums do not depend on the step count. Every lane of the wave must call it
// (the matrix-core instructions use the whole wave).
__device__ __forceinline__ void bsr44_mfma_row(const int* __restrict__ rptr, const int* __restrict__ bcol,
                                               const double* __restrict__ data, int64_t m, int64_t n, int k,
                                               const double* __restrict__ X, int64_t ldx, double* __restrict__ Y,
                                               int64_t ldy, double alpha, double beta, int64_t brow, int lane) {
  constexpr int U = kBsrSteps;
  const int x = lane & 3, q = (lane >> 2) & 3, y = lane >> 4;
  const int jb = 4 * q + x;       // this lane's right-hand side in each group of 16
  const int ng = (k + 15) >> 4;   // groups of 16 right-hand sides (wave-uniform)
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int e = rptr[brow + 1];
  for (int b = rptr[brow]; b < e; b += U) {
    double a[U];
    int64_t col[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool ok = b + u < e;
      a[u] = ok ? data[(int64_t)(b + u) * 16 + x * 4 + y] : 0.0;
      col[u] = ok ? (int64_t)bcol[b + u] * 4 + y : n;
    }
    double xv[U][4];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const int j = jb + 16 * g;
        xv[u][g] = (g < ng && col[u] < n && j < k) ? X[col[u] * ldx + j] : 0.0;
      }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (b + u >= e) break;  // wave-uniform
#pragma unroll
      for (int g = 0; g < 4; g++)
        if (g < ng) acc[g] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[u], xv[u][g], acc[g], 0, 0, 0);
    }
  }
  const int64_t row = brow * 4 + y;
  if (row < m) {
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const int j = jb + 16 * g;
      if (g < ng && j < k) {
        double* p = Y + row * ldy + j;
        *p = beta == 0.0 ? alpha * acc[g] : alpha * acc[g] + beta * *p;
      }
    }
  }
}

}  // namespace lsp
