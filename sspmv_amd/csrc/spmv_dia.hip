// DIA on gfx950.
//
// Reference: csr_to_dia (include/sparse_conversion.h:611-766) numbers the
// occupied diagonals in order of first appearance in a row-major scan, stores
// offsets col-row and data[d*stride + row] with stride = M rounded up to
// `alignment`, and refuses more than max_diags diagonals. LeSpMV_dia
// (src/spmv_dia.cpp:176-197) accumulates diagonal by diagonal.
//
// GPU conversion: atomicMin of the first nonzero index per diagonal, a radix
// sort of the occupied diagonals by that index (= first-appearance order),
// then a thread-per-row fill that keeps the reference's last-duplicate-wins.
// GPU kernels own rows instead of diagonals, so there are no atomics:
//  * vector (variant 1): one thread per row, x through L1/L2;
//  * window (variant 2): a workgroup owns 256 rows and stages the x windows it
//    needs in LDS, one window per cluster of nearby offsets (a 27-point stencil
//    has three clusters); falls back to variant 1 when the windows do not fit;
//  * SpMM: the DIA row engine (dia_engine.hpp): W right-hand sides per lane
//    and the loads of 8 diagonals in flight.
#include <algorithm>

#include "common.hpp"
#include "dia_engine.hpp"

namespace lsp {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxClusters = 16;
constexpr int kLdsBytes = 48 * 1024;

// first[m + c - i] = the smallest entry index on diagonal c - i (its first
// appearance in row-major order). Lane i takes row i; at step q every lane
// holds the q-th entry of its row, and entry indices grow with the lane. The
// lanes agreeing on a diagonal are served by one atomic from the lowest of
// them (the smallest index), so a banded matrix costs one atomic per diagonal
// per wave and step; and as the stored minimum only decreases, a plain read
// that is already <= that index (possibly an older, larger value: then the
// atomic runs anyway) skips it. (27-point stencil: 109M atomics on 27
// addresses before.)
__global__ __launch_bounds__(kBlock) void diag_first(int64_t m, const int* __restrict__ rp,
                                                     const int* __restrict__ ci, int* first) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool row_ok = i < m;
  const int s = row_ok ? rp[i] : 0, e = row_ok ? rp[i + 1] : 0;
  int len = e - s;
  // the wave's longest row bounds the steps (every lane takes part in the ballots)
  for (int off = 32; off > 0; off >>= 1) len = max(len, __shfl_xor(len, off));
  const int lane = threadIdx.x & 63;
  for (int q = 0; q < len; q++) {
    const int j = s + q;
    const bool has = j < e;
    const int64_t k = has ? (m - i) + ci[j] : -1;
    uint64_t todo = __ballot(has);
    while (todo) {
      const int leader = __ffsll((unsigned long long)todo) - 1;
      const long long kl = __shfl((long long)k, leader);
      const uint64_t same = __ballot(has && k == kl) & todo;
      if (lane == leader) {
        int* f = first + kl;
        if (*f > j) atomicMin(f, j);
      }
      todo &= ~same;
    }
  }
}

__global__ void fill_i32(int64_t n, int value, int* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = value;
}

__global__ void diag_flags(int64_t len, const int* __restrict__ first, int* __restrict__ flag) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < len) flag[k] = first[k] != INT32_MAX;
}

__global__ void diag_compact(int64_t len, const int* __restrict__ first, const int64_t* __restrict__ pos,
                             uint32_t* __restrict__ keys, uint32_t* __restrict__ idx) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= len || first[k] == INT32_MAX) return;
  keys[pos[k]] = (uint32_t)first[k];
  idx[pos[k]] = (uint32_t)k;
}

__global__ void diag_rank(int64_t nd, int64_t m, const uint32_t* __restrict__ idx, int* __restrict__ rank,
                          int64_t* __restrict__ offs) {
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nd) return;
  rank[idx[d]] = (int)d;
  offs[d] = (int64_t)idx[d] - m;
}

template <class V>
__global__ void dia_fill(int64_t m, const int* __restrict__ rp, const int* __restrict__ ci, const V* __restrict__ val,
                         const int* __restrict__ rank, int64_t stride, V* __restrict__ data) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  for (int j = rp[i]; j < rp[i + 1]; j++) data[(int64_t)rank[(m - i) + ci[j]] * stride + i] = val[j];
}

// R consecutive rows per thread; the data and x loads of U diagonals are all
// issued before the first product (a row's sum still runs over the diagonals
// in order). With NT the diagonals are read with non-temporal loads. One row a thread; U = 8 for FP32 (27-point stencil 160^3
// 148 -> 95 us, 2-D Laplacian 2048^2 26 -> 22 us; 4 or 16 diagonals, or 2 or
// 4 rows a thread, were slower), U = 1 for FP64 (205 / 40 us; U = 8 gave
// 208 / 44 us).
template <class V>
constexpr int kDiaRows = 1;
template <class V>
constexpr int kDiaUnroll = sizeof(V) == 8 ? 1 : 8;

template <class V, int R, int U, bool NT = false>
__global__ __launch_bounds__(kBlock) void dia_vector(int64_t m, int64_t n, int nd, int64_t stride,
                                                     const int64_t* __restrict__ offs, const V* __restrict__ data,
                                                     const V* __restrict__ x, V* __restrict__ y, V alpha, V beta) {
  const int64_t i0 = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * R;
  if (i0 >= m) return;
  if constexpr (R == 1 && U == 1) {
    V sum = 0;
    for (int d = 0; d < nd; d++) {
      const int64_t j = i0 + offs[d];
      if (j >= 0 && j < n) {
        const V* dp = data + (int64_t)d * stride + i0;
        sum += (NT ? __builtin_nontemporal_load(dp) : *dp) * x[j];
      }
    }
    y[i0] = beta == V(0) ? alpha * sum : alpha * sum + beta * y[i0];
    return;
  }
  V sum[R];
#pragma unroll
  for (int r = 0; r < R; r++) sum[r] = 0;
  for (int d0 = 0; d0 < nd; d0 += U) {
    V a[U][R], b[U][R];
    bool ok[U][R];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int d = d0 + u;
      const int64_t off = d < nd ? offs[d] : 0;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int64_t i = i0 + r, j = i + off;
        ok[u][r] = d < nd && i < m && j >= 0 && j < n;
        a[u][r] = ok[u][r] ? (NT ? __builtin_nontemporal_load(data + (int64_t)d * stride + i)
                                 : data[(int64_t)d * stride + i]) : V(0);
        b[u][r] = ok[u][r] ? x[j] : V(0);
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int r = 0; r < R; r++)
        if (ok[u][r]) sum[r] += a[u][r] * b[u][r];
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int64_t i = i0 + r;
    if (i < m) y[i] = beta == V(0) ? alpha * sum[r] : alpha * sum[r] + beta * y[i];
  }
}

struct Cluster {
  int64_t lo, hi;  // offset range
  int base;        // LDS element base of this cluster's window
};

template <class V, bool NT>
__global__ __launch_bounds__(kBlock) void dia_window(int64_t m, int64_t n, int nd, int64_t stride,
                                                     const int64_t* __restrict__ offs,
                                                     const int* __restrict__ lds_off, const V* __restrict__ data,
                                                     const V* __restrict__ x, V* __restrict__ y, V alpha, V beta,
                                                     int nclusters, const int64_t* __restrict__ clo,
                                                     const int* __restrict__ cwidth, const int* __restrict__ cbase) {
  extern __shared__ unsigned char smem[];
  V* win = reinterpret_cast<V*>(smem);
  const int64_t r0 = (int64_t)blockIdx.x * kBlock;
  for (int c = 0; c < nclusters; c++) {
    const int64_t start = r0 + clo[c];
    const int w = cwidth[c], base = cbase[c];
    for (int t = threadIdx.x; t < w; t += kBlock) {
      const int64_t j = start + t;
      win[base + t] = (j >= 0 && j < n) ? x[j] : V(0);
    }
  }
  __syncthreads();
  const int64_t i = r0 + threadIdx.x;
  if (i >= m) return;
  V sum = 0;
  for (int d = 0; d < nd; d++) {
    const int64_t j = i + offs[d];
    if (j >= 0 && j < n) {
      const V* dp = data + (int64_t)d * stride + i;
      sum += (NT ? __builtin_nontemporal_load(dp) : *dp) * win[lds_off[d] + threadIdx.x];
    }
  }
  y[i] = beta == V(0) ? alpha * sum : alpha * sum + beta * y[i];
}

template <class V, int W, bool VEC>
__global__ __launch_bounds__(kBlock) void dia_rows_kernel(const int64_t* __restrict__ offs,
                                                          const V* __restrict__ data, int64_t stride, int nd,
                                                          int64_t m, int64_t n, int k, const V* __restrict__ X,
                                                          int64_t ldx, V* __restrict__ Y, int64_t ldy, double alpha,
                                                          double beta, int rpi, int64_t nitems) {
  const int64_t item = xcd_block(blockIdx.x, gridDim.x) * (kBlock / 64) + threadIdx.x / 64;
  if (item >= nitems) return;  // whole waves leave together
  const int r0 = (int)(item * rpi);
  const int r1 = r0 + rpi < m ? r0 + rpi : (int)m;
  dia_rows<V, W, VEC>(offs, data, stride, nd, n, k, X, ldx, Y, ldy, alpha, beta, r0, r1, threadIdx.x & 63);
}

struct DiaDev : DiaData {
  // LDS window plan (variant 2)
  int nclusters = 0;
  int lds_elems = 0;
  DBuf lds_off, clo, cwidth, cbase;
  int64_t device_bytes() const override {
    return offsets.bytes + data.bytes + lds_off.bytes + clo.bytes + cwidth.bytes + cbase.bytes;
  }
};

int bits_for(uint64_t v) {
  int b = 0;
  while (b < 64 && (v >> b)) b++;
  return b;
}

// Group sorted offsets into clusters whose windows fit the LDS budget.
void plan_windows(DiaDev& d, const std::vector<int64_t>& offs, size_t vbytes, hipStream_t s) {
  d.nclusters = 0;
  if (offs.empty()) return;
  std::vector<int64_t> sorted(offs);
  std::sort(sorted.begin(), sorted.end());
  std::vector<std::pair<int64_t, int64_t>> cl;
  for (int64_t o : sorted) {
    if (!cl.empty() && o - cl.back().second <= kBlock) cl.back().second = o;
    else cl.push_back({o, o});
  }
  if ((int)cl.size() > kMaxClusters) return;
  std::vector<int64_t> lo;
  std::vector<int> width, base;
  int total = 0;
  for (auto& c : cl) {
    lo.push_back(c.first);
    width.push_back((int)(c.second - c.first + kBlock));
    base.push_back(total);
    total += width.back();
  }
  if ((size_t)total * vbytes > (size_t)kLdsBytes) return;
  std::vector<int> loff(offs.size());
  for (size_t q = 0; q < offs.size(); q++) {
    for (size_t c = 0; c < cl.size(); c++)
      if (offs[q] >= cl[c].first && offs[q] <= cl[c].second) loff[q] = base[c] + (int)(offs[q] - cl[c].first);
  }
  d.lds_off.alloc(sizeof(int) * loff.size());
  d.clo.alloc(sizeof(int64_t) * lo.size());
  d.cwidth.alloc(sizeof(int) * width.size());
  d.cbase.alloc(sizeof(int) * base.size());
  LSP_HIP(hipMemcpyAsync(d.lds_off.p, loff.data(), sizeof(int) * loff.size(), hipMemcpyHostToDevice, s));
  LSP_HIP(hipMemcpyAsync(d.clo.p, lo.data(), sizeof(int64_t) * lo.size(), hipMemcpyHostToDevice, s));
  LSP_HIP(hipMemcpyAsync(d.cwidth.p, width.data(), sizeof(int) * width.size(), hipMemcpyHostToDevice, s));
  LSP_HIP(hipMemcpyAsync(d.cbase.p, base.data(), sizeof(int) * base.size(), hipMemcpyHostToDevice, s));
  LSP_HIP(hipStreamSynchronize(s));
  d.nclusters = (int)cl.size();
  d.lds_elems = total;
}

template <class V>
void run(Matrix& A, DiaDev& d, int variant, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  // diagonals beyond the 256 MB last-level cache are streamed with
  // non-temporal loads, so they do not push x out of the caches (27-point
  // stencil 160^3: FP64 192 -> 179 us, FP32 94 -> 89 us); smaller ones keep
  // normal loads, which let repeated calls find part of them in that cache
  // (2-D Laplacian 1024^2 over four rotated copies: 11.7 us against 13.7)
  const bool nt = (double)d.ndiag * (double)d.stride * (double)sizeof(V) > 256.0 * (1 << 20);
  if (variant == LESPMV_VARIANT_MERGE || (variant == LESPMV_VARIANT_AUTO && d.nclusters > 0)) {
    if (d.nclusters > 0) {
      hipLaunchKernelGGL((nt ? dia_window<V, true> : dia_window<V, false>), grid1d(A.m, kBlock), dim3(kBlock),
                         d.lds_elems * sizeof(V), s, A.m, A.n,
                         d.ndiag, d.stride, d.offsets.as<int64_t>(), d.lds_off.as<int>(), d.data.as<V>(),
                         static_cast<const V*>(x), static_cast<V*>(y), (V)alpha, (V)beta, d.nclusters,
                         d.clo.as<int64_t>(), d.cwidth.as<int>(), d.cbase.as<int>());
      LSP_LAUNCHED();
      return;
    }
  }
  constexpr int R = kDiaRows<V>;
  hipLaunchKernelGGL((nt ? dia_vector<V, R, kDiaUnroll<V>, true> : dia_vector<V, R, kDiaUnroll<V>, false>),
                     grid1d(ceil_div(A.m, R), kBlock), dim3(kBlock), 0, s, A.m, A.n, d.ndiag, d.stride,
                     d.offsets.as<int64_t>(), d.data.as<V>(), static_cast<const V*>(x), static_cast<V*>(y), (V)alpha,
                     (V)beta);
  LSP_LAUNCHED();
}

// One pass of rows per work item (a wave): the DIA rows carry all their loads
// in one pass already.
constexpr int kDiaCallIters = 8;

template <class V>
void spmm_dispatch(Matrix& A, DiaDev& d, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
                   int64_t ldy, hipStream_t s) {
  const int mode = rows_mode(k, ldx, ldy, X, Y, (int64_t)sizeof(V));
  const int rpi = dia_item_rows(d.ndiag, k, kModeW[mode], kDiaCallIters);
  const int64_t nitems = ceil_div(A.m, rpi);
  void (*kern)(const int64_t*, const V*, int64_t, int, int64_t, int64_t, int, const V*, int64_t, V*, int64_t, double,
               double, int, int64_t);
  switch (mode) {
    case 0: kern = dia_rows_kernel<V, 1, false>; break;
    case 1: kern = dia_rows_kernel<V, 2, false>; break;
    case 2: kern = dia_rows_kernel<V, 2, true>; break;
    case 3: kern = dia_rows_kernel<V, 4, false>; break;
    default: kern = dia_rows_kernel<V, 4, true>; break;
  }
  hipLaunchKernelGGL(kern, grid1d(nitems, kBlock / 64), dim3(kBlock), 0, s, d.offsets.as<int64_t>(), d.data.as<V>(),
                     d.stride, d.ndiag, A.m, A.n, k, static_cast<const V*>(X), ldx, static_cast<V*>(Y), ldy, alpha,
                     beta, rpi, nitems);
  LSP_LAUNCHED();
}

}  // namespace

void convert_dia(Matrix& A, const lespmv_params_t& p, hipStream_t s) {
  LSP_REQUIRE(p.alignment > 0, LESPMV_ERR_ARG, "alignment must be positive");
  auto d = std::make_unique<DiaDev>();
  const int64_t len = A.m + A.n;
  DBuf first, flag, pos;
  first.alloc(sizeof(int) * std::max<int64_t>(len, 1));
  flag.alloc(sizeof(int) * std::max<int64_t>(len, 1));
  pos.alloc(sizeof(int64_t) * (len + 1));
  // INT32_MAX marks an unoccupied diagonal
  hipLaunchKernelGGL(fill_i32, grid1d(std::max<int64_t>(len, 1), kBlock), dim3(kBlock), 0, s,
                     std::max<int64_t>(len, 1), INT32_MAX, first.as<int>());
  LSP_LAUNCHED();
  if (A.m) {
    hipLaunchKernelGGL(diag_first, grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, A.rp.as<int>(), A.ci.as<int>(),
                       first.as<int>());
    LSP_LAUNCHED();
  }
  hipLaunchKernelGGL(diag_flags, grid1d(len, kBlock), dim3(kBlock), 0, s, len, first.as<int>(), flag.as<int>());
  LSP_LAUNCHED();
  const int64_t nd = scan_exclusive_i32_to_i64(flag.as<int32_t>(), pos.as<int64_t>(), len, s);
  d->ndiag = (int32_t)nd;
  check_fill(p, nd * align_up(A.m, p.alignment), A.nnz, A.m, "DIA layout");
  if (nd > p.max_diags) {
    throw Error(LESPMV_ERR_MAX_DIAGS,
                "Number of diagonals (" + std::to_string(nd) + ") excedes limit (" + std::to_string(p.max_diags) + ")");
  }
  DBuf keys, idx, rank;
  keys.alloc(sizeof(uint32_t) * std::max<int64_t>(nd, 1));
  idx.alloc(sizeof(uint32_t) * std::max<int64_t>(nd, 1));
  rank.alloc(sizeof(int) * std::max<int64_t>(len, 1));
  hipLaunchKernelGGL(diag_compact, grid1d(len, kBlock), dim3(kBlock), 0, s, len, first.as<int>(), pos.as<int64_t>(),
                     keys.as<uint32_t>(), idx.as<uint32_t>());
  LSP_LAUNCHED();
  radix_sort_pairs(keys.as<uint32_t>(), idx.as<uint32_t>(), nd, bits_for((uint64_t)std::max<int64_t>(A.nnz, 1)), s);
  d->offsets.alloc(sizeof(int64_t) * std::max<int64_t>(nd, 1));
  if (nd) {
    hipLaunchKernelGGL(diag_rank, grid1d(nd, kBlock), dim3(kBlock), 0, s, nd, A.m, idx.as<uint32_t>(),
                       rank.as<int>(), d->offsets.as<int64_t>());
    LSP_LAUNCHED();
  }
  d->stride = align_up(A.m, p.alignment);
  d->data.alloc(vsize(A.dtype) * nd * d->stride);
  LSP_HIP(hipMemsetAsync(d->data.p, 0, vsize(A.dtype) * nd * d->stride, s));
  if (A.m) {
    if (A.dtype == LESPMV_F64)
      hipLaunchKernelGGL((dia_fill<double>), grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, A.rp.as<int>(),
                         A.ci.as<int>(), A.val.as<double>(), rank.as<int>(), d->stride, d->data.as<double>());
    else
      hipLaunchKernelGGL((dia_fill<float>), grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, A.rp.as<int>(),
                         A.ci.as<int>(), A.val.as<float>(), rank.as<int>(), d->stride, d->data.as<float>());
    LSP_LAUNCHED();
  }
  auto offs = to_host<int64_t>(d->offsets.p, nd, s);
  if (nd) {
    d->min_off = *std::min_element(offs.begin(), offs.end());
    d->max_off = *std::max_element(offs.begin(), offs.end());
  }
  plan_windows(*d, offs, vsize(A.dtype), s);
  A.fmt[LESPMV_DIA] = std::move(d);
}

void dia_spmv(Matrix& A, int variant, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  auto& d = *static_cast<DiaDev*>(A.fmt[LESPMV_DIA].get());
  if (A.m == 0) return;
  if (A.dtype == LESPMV_F64) run<double>(A, d, variant, alpha, x, beta, y, s);
  else run<float>(A, d, variant, alpha, x, beta, y, s);
}

void dia_spmm(Matrix& A, int variant, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
              int64_t ldy, hipStream_t s) {
  (void)variant;
  auto& d = *static_cast<DiaDev*>(A.fmt[LESPMV_DIA].get());
  if (A.m == 0 || k == 0) return;
  if (A.dtype == LESPMV_F64) spmm_dispatch<double>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else spmm_dispatch<float>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
}

}  // namespace lsp
