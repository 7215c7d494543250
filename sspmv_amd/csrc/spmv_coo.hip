// Row-sorted COO on gfx950.
//
// Reference: csr_to_coo (include/sparse_conversion.h:144-168) expands rows in
// row order; LeSpMV_coo (src/spmv_coo.cpp:160-182) scatters
// y[row] += alpha*v*x[col] (serial) or with `omp atomic` (flag 1).
//
// GPU kernels:
//  * default: each wavefront owns kItems = 8 rounds x 64 consecutive nonzeros;
//    a wave64 segmented inclusive scan (keys = row ids, DPP/permute shuffles)
//    reduces each run of equal rows; runs that start and end inside the
//    wavefront are written directly, the row open at a wavefront edge leaves a
//    partial that a second, deterministic fix-up kernel adds in order.
//  * atomic (variant 3): y = beta*y, then every run adds alpha*sum with a
//    hardware FP64/FP32 atomic, like the reference's flag-1 kernel.
// Empty rows never appear in COO; their y = beta*y comes from a list built at
// conversion time.
#include <algorithm>

#include "common.hpp"

namespace lsp {
namespace {

constexpr int kBlock = 256;
constexpr int kRounds = 8;
constexpr int kWaveItems = kRounds * 64;

__global__ void expand_rows(int64_t m, const int* __restrict__ rp, int* __restrict__ row) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  for (int j = rp[i]; j < rp[i + 1]; j++) row[j] = (int)i;
}

__global__ void mark_empty(int64_t m, const int* __restrict__ rp, int* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) flag[i] = rp[i] == rp[i + 1];
}

__global__ void compact_empty(int64_t m, const int* __restrict__ rp, const int64_t* __restrict__ pos,
                              int* __restrict__ list) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && rp[i] == rp[i + 1]) list[pos[i]] = (int)i;
}

template <class V>
__global__ void scale_rows(int64_t n, const int* __restrict__ list, V* __restrict__ y, V beta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t r = list ? list[i] : i;
  y[r] = beta == V(0) ? V(0) : beta * y[r];
}

template <class V>
__device__ __forceinline__ V seg_scan(V v, int key, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const V up = __shfl_up(v, d, 64);
    const int kup = __shfl_up(key, d, 64);
    if (lane >= d && kup == key) v += up;
  }
  return v;
}

template <class V, bool kAtomic>
__global__ __launch_bounds__(kBlock) void coo_kernel(int64_t nnz, const int* __restrict__ row,
                                                     const int* __restrict__ col, const V* __restrict__ val,
                                                     const V* __restrict__ x, V* __restrict__ y, V alpha, V beta,
                                                     int* __restrict__ head_row, V* __restrict__ head_val,
                                                     int* __restrict__ tail_row, V* __restrict__ tail_val,
                                                     int* __restrict__ through) {
  const int lane = threadIdx.x & 63;
  const int64_t w = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t w0 = w * kWaveItems;
  if (w0 >= nnz) return;  // whole wavefronts leave together
  const int64_t w1 = min(nnz, w0 + kWaveItems);
  const int first_row = row[w0], last_row = row[w1 - 1];
  const bool first_cont = w0 > 0 && row[w0 - 1] == first_row;   // row started in an earlier wave
  const bool last_cont = w1 < nnz && row[w1] == last_row;       // row goes on in a later wave
  if (!kAtomic && lane == 0) {
    head_row[w] = -1;
    tail_row[w] = -1;
    through[w] = 0;
  }
  // all loads of the wavefront's 512 nonzeros are issued before the scans
  int keys[kRounds], cols[kRounds];
  V prods[kRounds];
#pragma unroll
  for (int r = 0; r < kRounds; r++) {
    const int64_t i = w0 + (int64_t)r * 64 + lane;
    const bool valid = i < w1;
    // read once: non-temporal loads keep x in the caches
    keys[r] = valid ? __builtin_nontemporal_load(row + i) : -2 - lane;  // distinct keys for lanes past the end
    cols[r] = valid ? __builtin_nontemporal_load(col + i) : 0;
    prods[r] = valid ? __builtin_nontemporal_load(val + i) : V(0);
  }
#pragma unroll
  for (int r = 0; r < kRounds; r++) {
    const int64_t i = w0 + (int64_t)r * 64 + lane;
    if (i < w1) prods[r] *= x[cols[r]];
  }
  int carry_row = -1;
  V carry = 0;
#pragma unroll
  for (int r = 0; r < kRounds; r++) {
    const int64_t i = w0 + (int64_t)r * 64 + lane;
    if (w0 + (int64_t)r * 64 >= w1) break;  // uniform across the wave
    const bool valid = i < w1;
    const int key = keys[r];
    V v = seg_scan(prods[r], key, lane);
    if (key == carry_row) v += carry;
    int next = __shfl_down(key, 1, 64);
    const int next_round = r + 1 < kRounds ? __shfl(keys[r + 1 < kRounds ? r + 1 : r], 0, 64) : -1;
    if (lane == 63) next = next_round;
    const bool run_end = valid && (next != key || i + 1 == w1);
    carry_row = __shfl(key, 63, 64);
    carry = __shfl(v, 63, 64);
    if (!run_end) continue;
    if (kAtomic) {
      atomicAdd(y + key, alpha * v);
      continue;
    }
    const bool ends_here = !(i + 1 == w1 && last_cont);
    const bool started_here = !(key == first_row && first_cont);
    if (ends_here && started_here) {
      y[key] = beta == V(0) ? alpha * v : alpha * v + beta * y[key];
    } else if (ends_here) {  // first row of the wave, begun earlier
      head_row[w] = key;
      head_val[w] = v;
    } else {  // last row of the wave, goes on later
      tail_row[w] = key;
      tail_val[w] = v;
      through[w] = started_here ? 0 : 1;
    }
  }
}

// The wavefront where a split row begins adds the partials of the following
// wavefronts in order and writes the row once.
template <class V>
__global__ void coo_fixup(int64_t waves, const int* __restrict__ head_row, const V* __restrict__ head_val,
                          const int* __restrict__ tail_row, const V* __restrict__ tail_val,
                          const int* __restrict__ through, V* __restrict__ y, V alpha, V beta) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= waves) return;
  const int r = tail_row[w];
  if (r < 0 || through[w]) return;
  V total = tail_val[w];
  for (int64_t u = w + 1; u < waves; u++) {
    if (through[u] && tail_row[u] == r) {
      total += tail_val[u];
      continue;
    }
    if (head_row[u] == r) total += head_val[u];
    break;
  }
  y[r] = beta == V(0) ? alpha * total : alpha * total + beta * y[r];
}

// SpMM: KP lanes per nonzero, lane c adds alpha*v*X[col, c] atomically after
// Y = beta*Y.
template <class V, int KP>
__global__ __launch_bounds__(kBlock) void coo_spmm_kernel(int64_t nnz, int k, const int* __restrict__ row,
                                                          const int* __restrict__ col, const V* __restrict__ val,
                                                          const V* __restrict__ X, int64_t ldx, V* __restrict__ Y,
                                                          int64_t ldy, V alpha) {
  const int c = threadIdx.x & (KP - 1);
  const int64_t i = (int64_t)blockIdx.x * (kBlock / KP) + threadIdx.x / KP;
  if (i >= nnz || c >= k) return;
  atomicAdd(Y + (int64_t)row[i] * ldy + c, alpha * val[i] * X[(int64_t)col[i] * ldx + c]);
}

template <class V>
__global__ void scale_matrix(int64_t m, int k, V* __restrict__ Y, int64_t ldy, V beta) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * k) return;
  const int64_t r = t / k, c = t - r * k;
  V* p = Y + r * ldy + c;
  *p = beta == V(0) ? V(0) : beta * *p;
}

struct CooScratch : CooData {
  DBuf empty;  // int32[num_empty]
  int64_t num_empty = 0;
  DBuf head_row, head_val, tail_row, tail_val, through;
  int64_t waves = 0;
  int64_t device_bytes() const override {
    return row.bytes + empty.bytes + head_row.bytes + head_val.bytes + tail_row.bytes + tail_val.bytes +
           through.bytes;
  }
};

template <class V>
void run(Matrix& A, CooScratch& d, int variant, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  V* yv = static_cast<V*>(y);
  const bool atomic = variant == LESPMV_VARIANT_ATOMIC;
  if (atomic) {
    hipLaunchKernelGGL((scale_rows<V>), grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, (const int*)nullptr, yv,
                       (V)beta);
    LSP_LAUNCHED();
  } else if (d.num_empty) {
    hipLaunchKernelGGL((scale_rows<V>), grid1d(d.num_empty, kBlock), dim3(kBlock), 0, s, d.num_empty,
                       d.empty.as<int>(), yv, (V)beta);
    LSP_LAUNCHED();
  }
  if (A.nnz == 0) return;
  const int64_t blocks = ceil_div(d.waves, kBlock / 64);
  if (atomic) {
    hipLaunchKernelGGL((coo_kernel<V, true>), dim3((unsigned)blocks), dim3(kBlock), 0, s, A.nnz, d.row.as<int>(),
                       A.ci.as<int>(), A.val.as<V>(), static_cast<const V*>(x), yv, (V)alpha, (V)beta, nullptr,
                       nullptr, nullptr, nullptr, nullptr);
    LSP_LAUNCHED();
    return;
  }
  hipLaunchKernelGGL((coo_kernel<V, false>), dim3((unsigned)blocks), dim3(kBlock), 0, s, A.nnz, d.row.as<int>(),
                     A.ci.as<int>(), A.val.as<V>(), static_cast<const V*>(x), yv, (V)alpha, (V)beta,
                     d.head_row.as<int>(), d.head_val.as<V>(), d.tail_row.as<int>(), d.tail_val.as<V>(),
                     d.through.as<int>());
  LSP_LAUNCHED();
  hipLaunchKernelGGL((coo_fixup<V>), grid1d(d.waves, kBlock), dim3(kBlock), 0, s, d.waves, d.head_row.as<int>(),
                     d.head_val.as<V>(), d.tail_row.as<int>(), d.tail_val.as<V>(), d.through.as<int>(), yv,
                     (V)alpha, (V)beta);
  LSP_LAUNCHED();
}

template <class V, int KP>
void spmm_launch(Matrix& A, CooScratch& d, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
                 int64_t ldy, hipStream_t s) {
  hipLaunchKernelGGL((scale_matrix<V>), grid1d(A.m * k, kBlock), dim3(kBlock), 0, s, A.m, k, static_cast<V*>(Y), ldy,
                     (V)beta);
  LSP_LAUNCHED();
  if (A.nnz == 0) return;
  hipLaunchKernelGGL((coo_spmm_kernel<V, KP>), grid1d(A.nnz, kBlock / KP), dim3(kBlock), 0, s, A.nnz, k,
                     d.row.as<int>(), A.ci.as<int>(), A.val.as<V>(), static_cast<const V*>(X), ldx,
                     static_cast<V*>(Y), ldy, (V)alpha);
  LSP_LAUNCHED();
}

template <class V>
void spmm_dispatch(Matrix& A, CooScratch& d, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
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

void convert_coo(Matrix& A, const lespmv_params_t& p, hipStream_t s) {
  (void)p;
  auto d = std::make_unique<CooScratch>();
  d->nnz = A.nnz;
  d->row.alloc(sizeof(int) * A.nnz);
  if (A.m) {
    hipLaunchKernelGGL(expand_rows, grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, A.rp.as<int>(), d->row.as<int>());
    LSP_LAUNCHED();
    DBuf flag, pos;
    flag.alloc(sizeof(int) * A.m);
    pos.alloc(sizeof(int64_t) * (A.m + 1));
    hipLaunchKernelGGL(mark_empty, grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, A.rp.as<int>(), flag.as<int>());
    LSP_LAUNCHED();
    d->num_empty = scan_exclusive_i32_to_i64(flag.as<int32_t>(), pos.as<int64_t>(), A.m, s);
    d->empty.alloc(sizeof(int) * std::max<int64_t>(d->num_empty, 1));
    if (d->num_empty) {
      hipLaunchKernelGGL(compact_empty, grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, A.rp.as<int>(),
                         pos.as<int64_t>(), d->empty.as<int>());
      LSP_LAUNCHED();
    }
    LSP_HIP(hipStreamSynchronize(s));
  }
  d->waves = ceil_div(A.nnz, kWaveItems);
  const int64_t w = std::max<int64_t>(d->waves, 1);
  d->head_row.alloc(sizeof(int) * w);
  d->tail_row.alloc(sizeof(int) * w);
  d->through.alloc(sizeof(int) * w);
  d->head_val.alloc(vsize(A.dtype) * w);
  d->tail_val.alloc(vsize(A.dtype) * w);
  A.fmt[LESPMV_COO] = std::move(d);
}

void coo_spmv(Matrix& A, int variant, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  auto& d = *static_cast<CooScratch*>(A.fmt[LESPMV_COO].get());
  if (A.m == 0) return;
  if (A.dtype == LESPMV_F64) run<double>(A, d, variant, alpha, x, beta, y, s);
  else run<float>(A, d, variant, alpha, x, beta, y, s);
}

void coo_spmm(Matrix& A, int variant, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
              int64_t ldy, hipStream_t s) {
  (void)variant;
  auto& d = *static_cast<CooScratch*>(A.fmt[LESPMV_COO].get());
  if (A.m == 0 || k == 0) return;
  if (A.dtype == LESPMV_F64) spmm_dispatch<double>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else spmm_dispatch<float>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
}

}  // namespace lsp
