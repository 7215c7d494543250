// BSR on gfx950.
//
// Reference: csr_to_bsr (include/sparse_conversion.h:768-926): block rows of br
// rows, occupied block columns listed in ascending order per block row, each
// block stored row-major, later duplicates overwrite; br = bc = 1 copies CSR
// unchanged. LeSpMV_bsr (src/spmv_bsr.cpp:214-233) multiplies block rows.
//
// GPU conversion: 64-bit keys (block row, block column) per nonzero, a radix
// sort, unique + scan for row_ptr/block_colindex, then a thread-per-row fill
// (binary search of the block column in its block row) that keeps the
// reference's last-duplicate-wins order.
// Kernels:
//  * block (default when br*bc <= 64 and both are powers of two): one
//    wavefront per block row; 64/(br*bc) blocks are processed side by side, so
//    every load of block data is 64 consecutive values; row sums are reduced
//    with lane shuffles;
//  * row (variant 1, any shape): one thread per row, walking its block row.
#include <algorithm>

#include "bsr_engine.hpp"
#include "common.hpp"

namespace lsp {
namespace {

constexpr int kBlock = 256;

__global__ void block_keys(int64_t m, int br, int bc, int64_t nb, const int* __restrict__ rp,
                           const int* __restrict__ ci, uint64_t* __restrict__ keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint64_t b = (uint64_t)(i / br);
  for (int j = rp[i]; j < rp[i + 1]; j++) keys[j] = b * (uint64_t)nb + (uint64_t)(ci[j] / bc);
}

__global__ void unique_flags(int64_t n, const uint64_t* __restrict__ keys, int* __restrict__ flag) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) flag[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
}

__global__ void unique_scatter(int64_t n, int64_t nb, const uint64_t* __restrict__ keys, const int* __restrict__ flag,
                               const int64_t* __restrict__ pos, int* __restrict__ bcol, int* __restrict__ counts) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !flag[k]) return;
  bcol[pos[k]] = (int)(keys[k] % (uint64_t)nb);
  atomicAdd(counts + keys[k] / (uint64_t)nb, 1);
}

__global__ void widen_counts(int64_t n, const int* __restrict__ c, int64_t* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = c[k];
}

__global__ void narrow_ptr(int64_t n, const int64_t* __restrict__ in, int* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = (int)in[k];
}

template <class V>
__global__ void bsr_fill(int64_t m, int br, int bc, const int* __restrict__ rp, const int* __restrict__ ci,
                         const V* __restrict__ val, const int* __restrict__ rptr, const int* __restrict__ bcol,
                         V* __restrict__ data) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t b = i / br;
  const int lo0 = rptr[b], hi0 = rptr[b + 1];
  const int64_t bs = (int64_t)br * bc;
  for (int j = rp[i]; j < rp[i + 1]; j++) {
    const int target = ci[j] / bc;
    int lo = lo0, hi = hi0;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (bcol[mid] < target) lo = mid + 1; else hi = mid;
    }
    data[(int64_t)lo * bs + (int64_t)(i % br) * bc + ci[j] % bc] = val[j];
  }
}

template <class V>
__global__ __launch_bounds__(kBlock) void bsr_row_kernel(int64_t m, int64_t n, int br, int bc,
                                                         const int* __restrict__ rptr, const int* __restrict__ bcol,
                                                         const V* __restrict__ data, const V* __restrict__ x,
                                                         V* __restrict__ y, V alpha, V beta) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= m) return;
  const int64_t b = i / br, r = i - b * br;
  const int64_t bs = (int64_t)br * bc;
  V sum = 0;
  for (int k = rptr[b]; k < rptr[b + 1]; k++) {
    const V* blk = data + (int64_t)k * bs + r * bc;
    const int64_t c0 = (int64_t)bcol[k] * bc;
    for (int c = 0; c < bc; c++)
      if (c0 + c < n) sum += blk[c] * x[c0 + c];
  }
  y[i] = beta == V(0) ? alpha * sum : alpha * sum + beta * y[i];
}

// One wavefront per block row; BS = br*bc divides 64, BC = bc. G = 64 / BS
// blocks per step; U steps' block columns and values are loaded (streamed
// past the caches that keep x) before any of their x gathers, so a block row
// of the 27-point stencil (about 26 blocks of 4 x 4) has all its loads in
// flight at once. Workgroups are renumbered so each XCD takes a contiguous
// range of block rows (their x neighbourhoods stay in that XCD's L2).
template <class V, int BS, int BC>
__global__ __launch_bounds__(kBlock) void bsr_wave_kernel(int64_t m, int64_t n, int64_t mb,
                                                          const int* __restrict__ rptr, const int* __restrict__ bcol,
                                                          const V* __restrict__ data, const V* __restrict__ x,
                                                          V* __restrict__ y, V alpha, V beta) {
  constexpr int G = 64 / BS;  // blocks per wave step
  constexpr int BR = BS / BC;
  constexpr int U = BS <= 16 ? 8 : 4;  // steps in flight
  const int lane = threadIdx.x & 63;
  const int64_t b = xcd_block(blockIdx.x, gridDim.x) * (kBlock / 64) + (threadIdx.x >> 6);
  if (b >= mb) return;
  const int q = lane / BS, p = lane % BS, c = p % BC;
  const int k0 = rptr[b], k1 = rptr[b + 1];
  V acc = 0;
  for (int kb = k0 + q; kb < k1; kb += G * U) {
    int col[U];
    V a[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = kb + u * G;
      const bool ok = k < k1;
      col[u] = ok ? __builtin_nontemporal_load(bcol + k) : -1;
      a[u] = ok ? __builtin_nontemporal_load(data + (int64_t)k * BS + p) : V(0);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t xc = (int64_t)col[u] * BC + c;
      if (col[u] >= 0 && xc < n) acc += a[u] * x[xc];
    }
  }
  // sum over the BC columns of a block row, then over the G blocks
#pragma unroll
  for (int o = 1; o < BC; o <<= 1) acc += __shfl_xor(acc, o, 64);
#pragma unroll
  for (int o = BS; o < 64; o <<= 1) acc += __shfl_xor(acc, o, 64);
  if (q == 0 && c == 0) {
    const int64_t row = b * BR + p / BC;
    if (row < m) y[row] = beta == V(0) ? alpha * acc : alpha * acc + beta * y[row];
  }
}

template <class V, int KP>
__global__ __launch_bounds__(kBlock) void bsr_spmm_kernel(int64_t m, int64_t n, int k, int br, int bc,
                                                          const int* __restrict__ rptr, const int* __restrict__ bcol,
                                                          const V* __restrict__ data, const V* __restrict__ X,
                                                          int64_t ldx, V* __restrict__ Y, int64_t ldy, V alpha,
                                                          V beta) {
  const int cc = threadIdx.x & (KP - 1);
  const int64_t i = (int64_t)blockIdx.x * (kBlock / KP) + threadIdx.x / KP;
  if (i >= m || cc >= k) return;
  const int64_t b = i / br, r = i - b * br;
  const int64_t bs = (int64_t)br * bc;
  V sum = 0;
  for (int q = rptr[b]; q < rptr[b + 1]; q++) {
    const V* blk = data + (int64_t)q * bs + r * bc;
    const int64_t c0 = (int64_t)bcol[q] * bc;
    for (int c = 0; c < bc; c++)
      if (c0 + c < n) sum += blk[c] * X[(c0 + c) * ldx + cc];
  }
  V* yp = Y + i * ldy + cc;
  *yp = beta == V(0) ? alpha * sum : alpha * sum + beta * *yp;
}

// BSR 4x4 FP64 SpMM on the matrix cores (bsr_engine.hpp): one wave per block
// row.
__global__ __launch_bounds__(kBlock) void bsr44_spmm_mfma(int64_t m, int64_t n, int64_t mb, int k,
                                                          const int* __restrict__ rptr, const int* __restrict__ bcol,
                                                          const double* __restrict__ data,
                                                          const double* __restrict__ X, int64_t ldx,
                                                          double* __restrict__ Y, int64_t ldy, double alpha,
                                                          double beta) {
  const int64_t brow = (int64_t)blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
  if (brow >= mb) return;  // whole waves leave together: the MFMAs need every lane
  bsr44_mfma_row(rptr, bcol, data, m, n, k, X, ldx, Y, ldy, alpha, beta, brow, threadIdx.x & 63);
}

int bits_for(uint64_t v) {
  int b = 0;
  while (b < 64 && (v >> b)) b++;
  return b;
}

template <class V, int BS, int BC>
bool try_wave(Matrix& A, BsrData& d, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  if (d.br * d.bc != BS || d.bc != BC) return false;
  hipLaunchKernelGGL((bsr_wave_kernel<V, BS, BC>), grid1d(d.mb * 64, kBlock), dim3(kBlock), 0, s, A.m, A.n, d.mb,
                     d.row_ptr.as<int>(), d.bcol.as<int>(), d.data.as<V>(), static_cast<const V*>(x),
                     static_cast<V*>(y), (V)alpha, (V)beta);
  LSP_LAUNCHED();
  return true;
}

template <class V>
void run(Matrix& A, BsrData& d, int variant, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  if (variant != LESPMV_VARIANT_VECTOR) {
    if (try_wave<V, 4, 2>(A, d, alpha, x, beta, y, s) || try_wave<V, 16, 4>(A, d, alpha, x, beta, y, s) ||
        try_wave<V, 8, 4>(A, d, alpha, x, beta, y, s) || try_wave<V, 16, 8>(A, d, alpha, x, beta, y, s) ||
        try_wave<V, 16, 2>(A, d, alpha, x, beta, y, s) || try_wave<V, 32, 8>(A, d, alpha, x, beta, y, s) ||
        try_wave<V, 64, 8>(A, d, alpha, x, beta, y, s) || try_wave<V, 32, 4>(A, d, alpha, x, beta, y, s) ||
        try_wave<V, 64, 16>(A, d, alpha, x, beta, y, s) || try_wave<V, 16, 16>(A, d, alpha, x, beta, y, s))
      return;
  }
  hipLaunchKernelGGL((bsr_row_kernel<V>), grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, A.n, d.br, d.bc,
                     d.row_ptr.as<int>(), d.bcol.as<int>(), d.data.as<V>(), static_cast<const V*>(x),
                     static_cast<V*>(y), (V)alpha, (V)beta);
  LSP_LAUNCHED();
}

template <class V, int KP>
void spmm_launch(Matrix& A, BsrData& d, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
                 int64_t ldy, hipStream_t s) {
  hipLaunchKernelGGL((bsr_spmm_kernel<V, KP>), grid1d(A.m, kBlock / KP), dim3(kBlock), 0, s, A.m, A.n, k, d.br, d.bc,
                     d.row_ptr.as<int>(), d.bcol.as<int>(), d.data.as<V>(), static_cast<const V*>(X), ldx,
                     static_cast<V*>(Y), ldy, (V)alpha, (V)beta);
  LSP_LAUNCHED();
}

template <class V>
void spmm_dispatch(Matrix& A, BsrData& d, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
                   int64_t ldy, hipStream_t s) {
  if (k <= 1) spmm_launch<V, 1>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 2) spmm_launch<V, 2>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 4) spmm_launch<V, 4>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 8) spmm_launch<V, 8>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 16) spmm_launch<V, 16>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 32) spmm_launch<V, 32>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else spmm_launch<V, 64>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
}

}  // namespace

void convert_bsr(Matrix& A, const lespmv_params_t& p, hipStream_t s) {
  LSP_REQUIRE(p.block_rows > 0 && p.block_cols > 0, LESPMV_ERR_ARG, "block dimensions must be positive");
  auto d = std::make_unique<BsrData>();
  d->br = p.block_rows;
  d->bc = p.block_cols;
  d->mb = ceil_div(A.m, d->br);
  d->nb = ceil_div(A.n, d->bc);
  const size_t vs = vsize(A.dtype);
  if (d->br == 1 && d->bc == 1) {  // include/sparse_conversion.h:784-815: CSR copied as is
    d->nnzb = A.nnz;
    d->row_ptr.alloc(sizeof(int) * (A.m + 1));
    d->bcol.alloc(sizeof(int) * A.nnz);
    d->data.alloc(vs * A.nnz);
    LSP_HIP(hipMemcpyAsync(d->row_ptr.p, A.rp.p, sizeof(int) * (A.m + 1), hipMemcpyDeviceToDevice, s));
    if (A.nnz) {
      LSP_HIP(hipMemcpyAsync(d->bcol.p, A.ci.p, sizeof(int) * A.nnz, hipMemcpyDeviceToDevice, s));
      LSP_HIP(hipMemcpyAsync(d->data.p, A.val.p, vs * A.nnz, hipMemcpyDeviceToDevice, s));
    }
    LSP_HIP(hipStreamSynchronize(s));
    A.fmt[LESPMV_BSR] = std::move(d);
    return;
  }
  LSP_REQUIRE((double)d->mb * (double)d->nb < 1.8e19, LESPMV_ERR_UNSUPPORTED, "block grid too large");
  const int64_t nnz = A.nnz;
  DBuf keys, flag, pos, counts, rptr64;
  keys.alloc(sizeof(uint64_t) * std::max<int64_t>(nnz, 1));
  flag.alloc(sizeof(int) * std::max<int64_t>(nnz, 1));
  pos.alloc(sizeof(int64_t) * (nnz + 1));
  counts.alloc(sizeof(int) * (d->mb + 1));
  rptr64.alloc(sizeof(int64_t) * (d->mb + 1));
  LSP_HIP(hipMemsetAsync(counts.p, 0, sizeof(int) * (d->mb + 1), s));
  if (A.m) {
    hipLaunchKernelGGL(block_keys, grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, d->br, d->bc, d->nb,
                       A.rp.as<int>(), A.ci.as<int>(), keys.as<uint64_t>());
    LSP_LAUNCHED();
  }
  radix_sort_u64(keys.as<uint64_t>(), nnz, bits_for((uint64_t)d->mb * (uint64_t)d->nb), s);
  if (nnz) {
    hipLaunchKernelGGL(unique_flags, grid1d(nnz, kBlock), dim3(kBlock), 0, s, nnz, keys.as<uint64_t>(),
                       flag.as<int>());
    LSP_LAUNCHED();
  }
  d->nnzb = scan_exclusive_i32_to_i64(flag.as<int32_t>(), pos.as<int64_t>(), nnz, s);
  LSP_REQUIRE(d->nnzb < INT32_MAX, LESPMV_ERR_UNSUPPORTED, "too many blocks for 32-bit block indices");
  check_fill(p, d->nnzb * d->br * d->bc, A.nnz, A.m, "BSR layout");
  d->bcol.alloc(sizeof(int) * std::max<int64_t>(d->nnzb, 1));
  if (nnz) {
    hipLaunchKernelGGL(unique_scatter, grid1d(nnz, kBlock), dim3(kBlock), 0, s, nnz, d->nb, keys.as<uint64_t>(),
                       flag.as<int>(), pos.as<int64_t>(), d->bcol.as<int>(), counts.as<int>());
    LSP_LAUNCHED();
  }
  hipLaunchKernelGGL(widen_counts, grid1d(d->mb, kBlock), dim3(kBlock), 0, s, d->mb, counts.as<int>(),
                     rptr64.as<int64_t>());
  LSP_LAUNCHED();
  const int64_t total = scan_exclusive_i64(rptr64.as<int64_t>(), d->mb, s);
  LSP_HIP(hipMemcpyAsync(rptr64.as<int64_t>() + d->mb, &total, sizeof(int64_t), hipMemcpyHostToDevice, s));
  d->row_ptr.alloc(sizeof(int) * (d->mb + 1));
  hipLaunchKernelGGL(narrow_ptr, grid1d(d->mb + 1, kBlock), dim3(kBlock), 0, s, d->mb + 1, rptr64.as<int64_t>(),
                     d->row_ptr.as<int>());
  LSP_LAUNCHED();
  const int64_t slots = d->nnzb * d->br * d->bc;
  d->data.alloc(vs * slots);
  LSP_HIP(hipMemsetAsync(d->data.p, 0, vs * slots, s));
  if (A.m) {
    if (A.dtype == LESPMV_F64)
      hipLaunchKernelGGL((bsr_fill<double>), grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, d->br, d->bc,
                         A.rp.as<int>(), A.ci.as<int>(), A.val.as<double>(), d->row_ptr.as<int>(), d->bcol.as<int>(),
                         d->data.as<double>());
    else
      hipLaunchKernelGGL((bsr_fill<float>), grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, d->br, d->bc,
                         A.rp.as<int>(), A.ci.as<int>(), A.val.as<float>(), d->row_ptr.as<int>(), d->bcol.as<int>(),
                         d->data.as<float>());
    LSP_LAUNCHED();
  }
  LSP_HIP(hipStreamSynchronize(s));
  A.fmt[LESPMV_BSR] = std::move(d);
}

void bsr_spmv(Matrix& A, int variant, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  auto& d = *static_cast<BsrData*>(A.fmt[LESPMV_BSR].get());
  if (A.m == 0) return;
  if (A.dtype == LESPMV_F64) run<double>(A, d, variant, alpha, x, beta, y, s);
  else run<float>(A, d, variant, alpha, x, beta, y, s);
}

void bsr_spmm(Matrix& A, int variant, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
              int64_t ldy, hipStream_t s) {
  auto& d = *static_cast<BsrData*>(A.fmt[LESPMV_BSR].get());
  if (A.m == 0 || k == 0) return;
  // 4x4 blocks in FP64 with more than four right-hand sides: the block product
  // is a dense 4x4x4 contraction, run on the matrix cores (with four or fewer,
  // three of the instruction's four products would idle; variant VECTOR keeps
  // the scalar kernel)
  if (A.dtype == LESPMV_F64 && d.br == 4 && d.bc == 4 && k > 4 && variant != LESPMV_VARIANT_VECTOR) {
    hipLaunchKernelGGL(bsr44_spmm_mfma, grid1d(d.mb, kBlock / 64), dim3(kBlock), 0, s, A.m, A.n, d.mb, k,
                       d.row_ptr.as<int>(), d.bcol.as<int>(), d.data.as<double>(), static_cast<const double*>(X), ldx,
                       static_cast<double*>(Y), ldy, alpha, beta);
    LSP_LAUNCHED();
    return;
  }
  if (A.dtype == LESPMV_F64) spmm_dispatch<double>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else spmm_dispatch<float>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
}

}  // namespace lsp
