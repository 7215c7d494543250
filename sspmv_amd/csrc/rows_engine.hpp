// The CSR row engine shared by the packed serving kernel (serve.hip) and the
// per-call SpMM variants (spmv_csr.hip): rows of a CSR matrix times k
// row-major right-hand sides, Y = alpha*A*X + beta*Y, with W right-hand sides
// per lane and kChains gathers in flight per lane, summed in FP64 for FP32
// data too (one rounding, at the store). Both callers use the same
// lane layout and summation order, so a request served in a batch and the
// same product launched on its own give bitwise equal results.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>

#include <hip/hip_runtime.h>

#include "buffer_loads.hpp"

namespace lsp {

constexpr int kItemIters = 64;  // target loads per lane in one work item
constexpr int kMaxPasses = 8;   // passes of 64 / G rows in one work item
constexpr int kChains = 8;      // nonzeros per lane whose loads are issued together
constexpr int kModeW[5] = {1, 2, 2, 4, 4};  // right-hand sides per lane of each mode

// Mode of a product: W right-hand sides per lane (1 for k = 1, 2 for k <= 3,
// else 4) and whether a lane's W values move as one vector, which needs k,
// ldx and ldy multiples of W and X, Y aligned to the vector. Depends on the
// product alone.
inline int rows_mode(int k, int64_t ldx, int64_t ldy, const void* X, const void* Y, int64_t vbytes) {
  const int W = k == 1 ? 1 : k <= 3 ? 2 : 4;
  if (W == 1) return 0;
  const uintptr_t align = (uintptr_t)W * (uintptr_t)vbytes;
  const bool vec = k % W == 0 && ldx % W == 0 && ldy % W == 0 && reinterpret_cast<uintptr_t>(X) % align == 0 &&
                   reinterpret_cast<uintptr_t>(Y) % align == 0;
  return W == 2 ? (vec ? 2 : 1) : (vec ? 4 : 3);
}

// Whether a product's loads fit 32-bit byte offsets from their arrays' bases
// (entries: nnz * value bytes, X: n rows of ldx), so that the row engine can
// issue them as buffer loads (run_rows' BUF).
inline bool rows_buffer_ok(int64_t nnz, int64_t n, int64_t ldx, int64_t vbytes) {
  return nnz * vbytes <= kMaxBufferBytes && n * ldx * vbytes <= kMaxBufferBytes;
}

// Rows in one work item: passes of 64 / G rows, enough passes for about
// `iters` loads per lane at the matrix's mean row length.
inline int rows_item(double avg_row, int lanes, int k, int W, int iters = kItemIters) {
  const int kw = (k + W - 1) / W;
  int kk = 1;
  while (kk < kw) kk <<= 1;
  const int G = std::max(lanes, kk), nzl = G / kk;
  const double per_pass = std::max(1.0, std::ceil(avg_row / nzl));
  const int passes = (int)std::clamp(std::ceil(iters / per_pass), 1.0, (double)kMaxPasses);
  return passes * (64 / G);
}

// Pointers taken from a request descriptor are generic to the compiler; in the
// global address space it issues global (not flat) loads, which wait on the
// vector-memory counter alone.
#define LSP_GLOBAL_AS __attribute__((address_space(1)))

// X and Y values of one lane's W right-hand sides: consecutive in memory; one
// vector access when the request is aligned for it (VEC), else W guarded
// scalar accesses.
template <class V, int W, bool VEC>
__device__ __forceinline__ void load_w(const LSP_GLOBAL_AS V* __restrict__ p, int nvalid, V* out) {
  if constexpr (VEC && W > 1) {
    typedef V vec_t __attribute__((ext_vector_type(W)));
    const vec_t v = *(const LSP_GLOBAL_AS vec_t*)p;
#pragma unroll
    for (int u = 0; u < W; u++) out[u] = v[u];
  } else {
#pragma unroll
    for (int u = 0; u < W; u++) out[u] = u < nvalid ? p[u] : (V)0;
  }
}

// The same from byte offset `off` of a buffer resource.
template <class V, int W, bool VEC>
__device__ __forceinline__ void load_w_buf(__amdgpu_buffer_rsrc_t r, uint32_t off, int nvalid, V* out) {
  constexpr int B = W * (int)sizeof(V);
  if constexpr (VEC && W > 1 && B == 8) {
    typedef V vec_t __attribute__((ext_vector_type(W)));
    const vec_t v = __builtin_bit_cast(vec_t, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
#pragma unroll
    for (int u = 0; u < W; u++) out[u] = v[u];
  } else if constexpr (VEC && W > 1 && B >= 16) {
    typedef V half_t __attribute__((ext_vector_type(16 / sizeof(V))));
    constexpr int H = 16 / (int)sizeof(V);
#pragma unroll
    for (int h = 0; h < W / H; h++) {
      const half_t v = __builtin_bit_cast(half_t, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, h * 16, 0));
#pragma unroll
      for (int u = 0; u < H; u++) out[h * H + u] = v[u];
    }
  } else {
#pragma unroll
    for (int u = 0; u < W; u++) out[u] = u < nvalid ? buf_load<V>(r, off, u * (uint32_t)sizeof(V)) : (V)0;
  }
}

// Rows [r0, r1) of request q for all k right-hand sides (positions [r0, r1)
// of q.order when q.order is set: rows sorted by length within windows, so
// that the rows of a pass need about the same number of steps). A row gets G
// lanes
// (G = max(lanes, kk), kk = lanes needed to cover k at W right-hand sides
// each, rounded up to a power of two); lane t handles right-hand sides
// (t mod kk)*W .. +W-1 and every (G/kk)-th nonzero of the row, and the
// partial sums meet in xor shuffles. A lane issues the index, value and x
// loads of kChains nonzeros before it uses any of them (out-of-row slots are
// masked), so that many gathers are in flight per lane. With BUF (the product
// passes rows_buffer_ok) the index, value and X loads are buffer loads: 32-bit
// offsets, the chain stride in a scalar offset.
template <class V, int W, bool VEC, bool BUF, class Q>
__device__ __forceinline__ void run_rows(const Q& q, int r0, int r1, int lane) {
  constexpr int CH = kChains;
  const int k = q.k;
  const int kw = (k + W - 1) / W;
  int kk = 1;
  while (kk < kw) kk <<= 1;
  const int G = q.lanes > kk ? q.lanes : kk;  // lanes per row (power of two <= 64)
  const int P = 64 / G;                       // rows per pass
  const int grp = lane / G, t = lane % G, j = (t % kk) * W, s = t / kk, nzl = G / kk;
  const int nvalid = k - j;
  const LSP_GLOBAL_AS V* __restrict__ val = (const LSP_GLOBAL_AS V*)q.val;
  const LSP_GLOBAL_AS V* __restrict__ X = (const LSP_GLOBAL_AS V*)q.X + j;
  LSP_GLOBAL_AS V* __restrict__ Y = (LSP_GLOBAL_AS V*)q.Y;
  const LSP_GLOBAL_AS int* __restrict__ rp = (const LSP_GLOBAL_AS int*)q.rp;
  const LSP_GLOBAL_AS int* __restrict__ ci = (const LSP_GLOBAL_AS int*)q.ci;
  const int64_t ldx = q.ldx;
  const double alpha = q.alpha, beta = q.beta;
  __amdgpu_buffer_rsrc_t cr, vr, xr;
  if constexpr (BUF) {
    cr = make_rsrc(q.ci, kMaxBufferBytes);
    vr = make_rsrc(q.val, kMaxBufferBytes);
    xr = make_rsrc(q.X, kMaxBufferBytes);
  }
  const LSP_GLOBAL_AS int* __restrict__ order = (const LSP_GLOBAL_AS int*)q.order;
  for (int base = r0; base < r1; base += P) {
    // position base + grp of the item; with an order, the rows of a pass are
    // taken from it (a row's sum does not depend on which pass computes it)
    const int pos = base + grp;
    const bool ok = pos < r1 && j < k;
    const int row = ok && order ? order[pos] : pos;
    double acc[W];  // FP64 sums for FP32 data too: products and sums round once, at the store
#pragma unroll
    for (int u = 0; u < W; u++) acc[u] = 0;
    if (ok) {
      const int e = rp[row + 1];
      for (int p = rp[row] + s; p < e; p += CH * nzl) {
        int cc[CH];
        V vv[CH], xx[CH][W];
#pragma unroll
        for (int c = 0; c < CH; c++) {
          const int pc = p + c * nzl;
          if constexpr (BUF) {
            const uint32_t cs = (uint32_t)(c * nzl);  // wave-uniform chain stride
            cc[c] = pc < e ? buf_load<int>(cr, (uint32_t)p * 4u, cs * 4u) : 0;
            vv[c] = pc < e ? buf_load<V>(vr, (uint32_t)p * (uint32_t)sizeof(V), cs * (uint32_t)sizeof(V)) : (V)0;
          } else {
            cc[c] = pc < e ? ci[pc] : 0;
            vv[c] = pc < e ? val[pc] : (V)0;
          }
        }
#pragma unroll
        for (int c = 0; c < CH; c++) {
          if (p + c * nzl < e) {
            if constexpr (BUF)
              load_w_buf<V, W, VEC>(xr, ((uint32_t)cc[c] * (uint32_t)ldx + (uint32_t)j) * (uint32_t)sizeof(V), nvalid,
                                    xx[c]);
            else
              load_w<V, W, VEC>(X + (int64_t)cc[c] * ldx, nvalid, xx[c]);
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
    }
    double sum[W];
#pragma unroll
    for (int u = 0; u < W; u++) {
      double a = acc[u];
      for (int off = kk; off < G; off <<= 1) a += __shfl_xor(a, off);
      sum[u] = a;
    }
    if (ok && s == 0) {
      LSP_GLOBAL_AS V* y = Y + (int64_t)row * q.ldy + j;
      if constexpr (VEC && W > 1) {
        typedef V vec_t __attribute__((ext_vector_type(W)));
        vec_t o;
        if (beta == 0.0) {
#pragma unroll
          for (int u = 0; u < W; u++) o[u] = (V)(alpha * sum[u]);
        } else {
          const vec_t old = *(const LSP_GLOBAL_AS vec_t*)y;
#pragma unroll
          for (int u = 0; u < W; u++) o[u] = (V)(alpha * sum[u] + beta * (double)old[u]);
        }
        *(LSP_GLOBAL_AS vec_t*)y = o;
      } else {
#pragma unroll
        for (int u = 0; u < W; u++)
          if (u < nvalid) y[u] = (V)(beta == 0.0 ? alpha * sum[u] : alpha * sum[u] + beta * (double)y[u]);
      }
    }
  }
}

}  // namespace lsp
