// Device primitives used by the format conversions: reductions, exclusive
// scans and a stable LSD radix sort (8-bit digits, wave64 ballot ranking).
#include <algorithm>

#include "common.hpp"

namespace lsp {
namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 4;                       // items per thread in scans
constexpr int kScanTile = kBlock * kScanItems;      // 1024
constexpr int kSortRounds = 16;                     // rounds of 256 items per sort tile
constexpr int kSortTile = kBlock * kSortRounds;     // 4096 keys per workgroup

__global__ void max_rowlen_kernel(int64_t m, const int32_t* __restrict__ rp, int* __restrict__ out) {
  int best = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    best = max(best, rp[i + 1] - rp[i]);
  for (int off = 32; off > 0; off >>= 1) best = max(best, __shfl_xor(best, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, best);
}

// ---- exclusive scan of int64 ------------------------------------------------

__device__ int64_t block_exclusive_scan(int64_t v, int64_t* lds, int64_t* total) {
  // wave inclusive scan
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int64_t t = __shfl_up(x, off, 64);
    if (lane >= off) x += t;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t acc = 0;
    for (int i = 0; i < kBlock / 64; i++) { int64_t t = lds[i]; lds[i] = acc; acc += t; }
    lds[kBlock / 64] = acc;
  }
  __syncthreads();
  const int64_t excl = x - v + lds[w];
  *total = lds[kBlock / 64];
  __syncthreads();
  return excl;
}

// pass 1: per-tile sums
__global__ void scan_tile_sums(const int64_t* __restrict__ in, int64_t n, int64_t* __restrict__ sums) {
  __shared__ int64_t lds[kBlock / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++)
    if (base + i < n) s += in[base + i];
  int64_t total;
  block_exclusive_scan(s, lds, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// pass 2: scan each tile, adding the tile's offset
__global__ void scan_tiles_apply(const int64_t* __restrict__ in, int64_t* __restrict__ out, int64_t n,
                                 const int64_t* __restrict__ offs) {
  __shared__ int64_t lds[kBlock / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = base + i < n ? in[base + i] : 0;
    s += v[i];
  }
  int64_t total;
  int64_t run = block_exclusive_scan(s, lds, &total) + offs[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
}

__global__ void widen_i32(const int32_t* __restrict__ in, int64_t* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

// Exclusive scan of `n` int64 in place; returns total (host, synchronising).
int64_t scan_inplace(int64_t* d, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  const int64_t tiles = ceil_div(n, kScanTile);
  DBuf sums;
  sums.alloc(sizeof(int64_t) * (tiles + 1));
  hipLaunchKernelGGL(scan_tile_sums, dim3((unsigned)tiles), dim3(kBlock), 0, s, d, n, sums.as<int64_t>());
  LSP_LAUNCHED();
  int64_t total;
  if (tiles == 1) {
    LSP_HIP(hipMemcpyAsync(&total, sums.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    LSP_HIP(hipStreamSynchronize(s));
    LSP_HIP(hipMemsetAsync(sums.p, 0, sizeof(int64_t), s));
  } else {
    total = scan_inplace(sums.as<int64_t>(), tiles, s);
  }
  hipLaunchKernelGGL(scan_tiles_apply, dim3((unsigned)tiles), dim3(kBlock), 0, s, d, d, n, sums.as<int64_t>());
  LSP_LAUNCHED();
  LSP_HIP(hipStreamSynchronize(s));  // keep `sums` alive until the apply pass ran
  return total;
}

// ---- stable radix sort --------------------------------------------------------

template <class K>
__device__ __forceinline__ unsigned digit_of(K k, int shift) {
  return (unsigned)((k >> shift) & (K)0xFF);
}

template <class K>
__global__ __launch_bounds__(kBlock) void radix_hist(const K* __restrict__ keys, int64_t n, int shift,
                                                     int64_t* __restrict__ hist, int64_t nblocks) {
  __shared__ int cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortRounds; r++) {
    const int64_t i = base + (int64_t)r * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&cnt[digit_of(keys[i], shift)], 1);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Rank of this lane among lower lanes of the wave holding the same digit, and
// the mask of lanes with that digit.
__device__ __forceinline__ uint64_t match_digit(unsigned d, bool valid) {
  uint64_t match = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const uint64_t bal = __ballot(valid && ((d >> b) & 1));
    match &= ((d >> b) & 1) ? bal : ~bal;
  }
  return match;
}

template <class K, bool kVals>
__global__ __launch_bounds__(kBlock) void radix_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                        K* __restrict__ kout, uint32_t* __restrict__ vout, int64_t n,
                                                        int shift, const int64_t* __restrict__ offs,
                                                        int64_t nblocks) {
  __shared__ int64_t running[256];
  __shared__ int wave_cnt[kBlock / 64][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  running[threadIdx.x] = offs[(int64_t)threadIdx.x * nblocks + blockIdx.x];
  for (int i = 0; i < kBlock / 64; i++) wave_cnt[i][threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kSortRounds; r++) {
    const int64_t i = base + (int64_t)r * kBlock + threadIdx.x;
    const bool valid = i < n;
    K key = valid ? kin[i] : (K)0;
    const unsigned d = digit_of(key, shift);
    const uint64_t match = match_digit(d, valid);
    const int below = __popcll(match & lt);
    if (valid && below == 0) wave_cnt[w][d] = __popcll(match);
    __syncthreads();
    if (valid) {
      int64_t dest = running[d] + below;
      for (int ww = 0; ww < w; ww++) dest += wave_cnt[ww][d];
      kout[dest] = key;
      if (kVals) vout[dest] = vin[i];
    }
    __syncthreads();
    int add = 0;
    for (int ww = 0; ww < kBlock / 64; ww++) {
      add += wave_cnt[ww][threadIdx.x];
      wave_cnt[ww][threadIdx.x] = 0;
    }
    running[threadIdx.x] += add;
    __syncthreads();
  }
}

template <class K, bool kVals>
void radix_sort_impl(K* keys, uint32_t* vals, int64_t n, int bits, hipStream_t s) {
  if (n <= 1 || bits <= 0) return;
  const int64_t nblocks = ceil_div(n, kSortTile);
  DBuf kbuf, vbuf, hist;
  kbuf.alloc(sizeof(K) * n);
  if (kVals) vbuf.alloc(sizeof(uint32_t) * n);
  hist.alloc(sizeof(int64_t) * 256 * nblocks);
  K* ka = keys;
  K* kb = kbuf.as<K>();
  uint32_t* va = vals;
  uint32_t* vb = kVals ? vbuf.as<uint32_t>() : nullptr;
  int passes = 0;
  for (int shift = 0; shift < bits; shift += 8, passes++) {
    hipLaunchKernelGGL((radix_hist<K>), dim3((unsigned)nblocks), dim3(kBlock), 0, s, ka, n, shift,
                       hist.as<int64_t>(), nblocks);
    LSP_LAUNCHED();
    scan_inplace(hist.as<int64_t>(), 256 * nblocks, s);
    hipLaunchKernelGGL((radix_scatter<K, kVals>), dim3((unsigned)nblocks), dim3(kBlock), 0, s, ka, va, kb, vb, n,
                       shift, hist.as<int64_t>(), nblocks);
    LSP_LAUNCHED();
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (passes & 1) {  // result sits in the scratch buffers
    LSP_HIP(hipMemcpyAsync(keys, ka, sizeof(K) * n, hipMemcpyDeviceToDevice, s));
    if (kVals) LSP_HIP(hipMemcpyAsync(vals, va, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
  }
  LSP_HIP(hipStreamSynchronize(s));
}

// Plain streaming copy (16-byte vectors): the reference point for HBM byte
// counters and for the achievable copy bandwidth.
__global__ __launch_bounds__(kBlock) void copy_kernel(const int4* __restrict__ src, int4* __restrict__ dst,
                                                      int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += (int64_t)gridDim.x * kBlock)
    dst[i] = src[i];
}

}  // namespace

void device_copy(void* dst, const void* src, int64_t bytes, hipStream_t s) {
  LSP_REQUIRE(bytes % 16 == 0, LESPMV_ERR_ARG, "copy size must be a multiple of 16 bytes");
  const int64_t n16 = bytes / 16;
  if (n16 == 0) return;
  // four workgroups per CU, each striding over the buffer: on MI355X this
  // moved 5.66 TB/s (1 GiB in, 1 GiB out) against 4.55 TB/s with 64 per CU
  const int64_t blocks = std::min<int64_t>(ceil_div(n16, kBlock), 256 * 4);
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, static_cast<const int4*>(src),
                     static_cast<int4*>(dst), n16);
  LSP_LAUNCHED();
}

int64_t reduce_max_rowlen(const int32_t* rp, int64_t m, hipStream_t s) {
  DBuf out;
  out.alloc(sizeof(int));
  LSP_HIP(hipMemsetAsync(out.p, 0, sizeof(int), s));
  const int64_t blocks = std::min<int64_t>(ceil_div(m, kBlock), 4096);
  hipLaunchKernelGGL(max_rowlen_kernel, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(kBlock), 0, s, m, rp,
                     out.as<int>());
  LSP_LAUNCHED();
  int h = 0;
  LSP_HIP(hipMemcpyAsync(&h, out.p, sizeof(int), hipMemcpyDeviceToHost, s));
  LSP_HIP(hipStreamSynchronize(s));
  return h;
}

int64_t scan_exclusive_i64(int64_t* d, int64_t n, hipStream_t s) { return scan_inplace(d, n, s); }

int64_t scan_exclusive_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(widen_i32, grid1d(std::min<int64_t>(n, 1 << 22), kBlock), dim3(kBlock), 0, s, in, out, n);
  LSP_LAUNCHED();
  return scan_inplace(out, n, s);
}

void radix_sort_pairs(uint32_t* keys, uint32_t* vals, int64_t n, int bits, hipStream_t s) {
  radix_sort_impl<uint32_t, true>(keys, vals, n, std::min(bits, 32), s);
}

void radix_sort_u32(uint32_t* keys, int64_t n, int bits, hipStream_t s) {
  radix_sort_impl<uint32_t, false>(keys, nullptr, n, std::min(bits, 32), s);
}

void radix_sort_u64(uint64_t* keys, int64_t n, int bits, hipStream_t s) {
  radix_sort_impl<uint64_t, false>(keys, nullptr, n, std::min(bits, 64), s);
}

}  // namespace lsp
