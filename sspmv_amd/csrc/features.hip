// Matrix features on the GPU: the MTX extractor of LeSpMV
// (src/sparse_features.cpp) as histogram and reduction kernels.
//
// Reference steps and where they happen here:
//   MtxLoad per-entry histograms (:336-472)   feat_tiles: one wave per row of a
//       256-row-block chunk; the row's column-block counts go to a per-wave LDS
//       histogram (Rows_cnt[row][*] without storing it), tile counts, distinct
//       rows per tile (uniq_RB), per-tile row maxima and sums of squares are
//       accumulated in LDS per workgroup; column counts, the set of occupied
//       diagonals and the (row block, column) occupancy bitmap behind uniq_CB
//       are global.
//   Diag_Dom, per-row log10 ranges, distance  feat_rows: one thread per row,
//       the row's entries in CSR order (file order for files read by
//       lespmv_read_mtx), so Diag_Dom's sign is the reference's.
//   column statistics and log10 ranges        feat_cols: one thread per column.
//   P_ratioAndGini (:611-640)                 radix sort of the count vector,
//       exclusive scan, one kernel for the trapezoid sum and the smallest k
//       with k/len + (sum of the k largest)/nnz >= 1 (monotone in k).
//   pattern / value symmetry (:726-762)       stable two-key radix sort of the
//       entries by (row, col); for each entry a binary search of its mirror; the
//       values compared are those of the first occurrence of each position, and
//       they compare equal when their "%lf" renderings are equal (sign, and the
//       value rounded half-to-even at six decimals).
// Final ratios are formed on the host from the exact integer aggregates.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace lsp {

namespace {

constexpr int kT = LESPMV_FEAT_TILES;
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

struct Geo {
  int64_t m, n;
  int qR, modR, thrR, qC, modC, thrC;
  int64_t close_thr;
};

__host__ __device__ __forceinline__ int rb_of(const Geo& g, int r) {
  return r < g.thrR ? r / (g.qR + 1) : g.modR + (r - g.thrR) / g.qR;
}
__host__ __device__ __forceinline__ int cb_of(const Geo& g, int c) {
  return c < g.thrC ? c / (g.qC + 1) : g.modC + (c - g.thrC) / g.qC;
}

enum Acc {
  A_LOWER, A_UPPER, A_DIAG, A_CLOSE, A_NZ_ROWS, A_ROW_MIN, A_ROW_MAX, A_ROW_SQ, A_DOMINANT, A_ROW_DIV, A_ROW_DIV_SET,
  A_MAX_OFF, A_MAX_DIAG, A_NZ_COLS, A_COL_MIN, A_COL_MAX, A_COL_SQ, A_COL_DIV, A_COL_DIV_SET, A_NDIAGS, A_PATTERN,
  A_VALUE, A_COUNT
};
constexpr unsigned long long kMinInit = 100000000ull;  // the reference's initial minima

using u64 = unsigned long long;

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
  for (int o = 32; o; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}
template <class T>
__device__ __forceinline__ T wave_min(T v) {
  for (int o = 32; o; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ u64 dbits(double d) { return (u64)__double_as_longlong(d); }
__device__ __forceinline__ double bitsd(u64 b) { return __longlong_as_double((long long)b); }

// Adds (sum, max, sum of squares) of a per-thread count to its block's slot;
// one set of atomics per wave when the wave's threads share the block.
__device__ void block_stats(bool ok, int blk, uint32_t cnt, uint32_t* nnz, uint32_t* mx, u64* sq) {
  const int lane = threadIdx.x & 63;
  const int b0 = __shfl(ok ? blk : -1, 0);
  const bool uniform = __all(!ok || blk == b0) && b0 >= 0;
  if (uniform) {
    const u64 s = wave_sum<u64>(ok ? cnt : 0);
    const uint32_t m = wave_max<uint32_t>(ok ? cnt : 0);
    const u64 q = wave_sum<u64>(ok ? (u64)cnt * cnt : 0);
    if (lane == 0) {
      atomicAdd(&nnz[b0], (uint32_t)s);
      atomicMax(&mx[b0], m);
      atomicAdd(&sq[b0], q);
    }
  } else if (ok) {
    atomicAdd(&nnz[blk], cnt);
    atomicMax(&mx[blk], cnt);
    atomicAdd(&sq[blk], (u64)cnt * cnt);
  }
}

// One thread per row: length, Diag_Dom (sequential, CSR order), log10 range,
// gap sum / length, row-block sums. The distance partial of each block is
// reduced in a fixed order.
template <class V>
__global__ __launch_bounds__(kBlock) void feat_rows(Geo g, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ci, const V* __restrict__ val,
                                                    uint32_t* __restrict__ rowlen, u64* __restrict__ acc,
                                                    uint32_t* __restrict__ rb_nnz, uint32_t* __restrict__ rb_max,
                                                    u64* __restrict__ rb_sq, double* __restrict__ dist_part) {
  __shared__ double s_dist[kWaves];
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool ok = r < g.m;
  uint32_t len = 0;
  double dom = 0.0, mx = 0.0, mn = 0.0, dist = 0.0;
  if (ok) {
    const int s = rp[r], e = rp[r + 1];
    len = (uint32_t)(e - s);
    for (int j = s; j < e; j++) {
      const double a = fabs((double)val[j]);
      if (ci[j] == r) dom += a;
      else dom -= a;
      const double l = log10(a);
      if (mx < l) mx = l;
      if (a > 0.0 && mn > l) mn = l;
    }
    if (len) dist = (double)(ci[e - 1] - ci[s]) / (double)len;
    rowlen[r] = len;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const u64 nz = wave_sum<u64>(ok && len ? 1 : 0);
  const u64 lmin = wave_min<u64>(ok ? len : kMinInit);
  const u64 lmax = wave_max<u64>(ok ? len : 0);
  const u64 lsq = wave_sum<u64>(ok ? (u64)len * len : 0);
  const u64 dominant = wave_sum<u64>(ok && dom > 0 ? 1 : 0);
  const bool has_div = ok && mn != 0.0;
  const u64 div = wave_max<u64>(has_div ? dbits(mx - mn) : 0);
  const u64 div_set = wave_sum<u64>(has_div ? 1 : 0);
  const double dsum = wave_sum<double>(dist);
  if (lane == 0) {
    atomicAdd(&acc[A_NZ_ROWS], nz);
    atomicMin(&acc[A_ROW_MIN], lmin);
    atomicMax(&acc[A_ROW_MAX], lmax);
    atomicAdd(&acc[A_ROW_SQ], lsq);
    atomicAdd(&acc[A_DOMINANT], dominant);
    if (div_set) {
      atomicMax(&acc[A_ROW_DIV], div);
      atomicAdd(&acc[A_ROW_DIV_SET], div_set);
    }
    s_dist[w] = dsum;
  }
  block_stats(ok, ok ? rb_of(g, (int)r) : 0, len, rb_nnz, rb_max, rb_sq);
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kWaves; i++) t += s_dist[i];
    dist_part[blockIdx.x] = t;
  }
}

// Fixed-order sum of the per-block distance partials (one block).
__global__ __launch_bounds__(kBlock) void sum_partials(const double* __restrict__ part, int64_t n,
                                                       double* __restrict__ out) {
  __shared__ double s[kBlock];
  double t = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kBlock) t += part[i];
  s[threadIdx.x] = t;
  __syncthreads();
  for (int o = kBlock / 2; o; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

struct TileBufs {
  uint32_t *tiles, *uniq_rb, *uniq_cb, *tile_max;
  u64* tile_sq;
  uint32_t* col_cnt;
  unsigned char* dflag;
  u64 *colmax, *colminneg;
  uint32_t* bitmap;
  int64_t words;  // bitmap words per row block
  u64* acc;
};

// One wave per row; a workgroup covers `rows_per_wg` rows of one row block.
template <class V>
__global__ __launch_bounds__(kBlock) void feat_tiles(Geo g, const int32_t* __restrict__ rp,
                                                     const int32_t* __restrict__ ci, const V* __restrict__ val,
                                                     int chunks_per_rb, int rows_per_wg, TileBufs b) {
  __shared__ int hist[kWaves][kT];
  __shared__ uint32_t t_cnt[kT], t_urb[kT], t_ucb[kT], t_max[kT];
  __shared__ u64 t_sq[kT];
  const int rb = blockIdx.x / chunks_per_rb, chunk = blockIdx.x % chunks_per_rb;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < kT; i += kBlock) {
    t_cnt[i] = t_urb[i] = t_ucb[i] = t_max[i] = 0;
    t_sq[i] = 0;
  }
  for (int i = lane; i < kT; i += 64) hist[w][i] = 0;
  __syncthreads();
  const int64_t rb_begin = rb < g.modR ? (int64_t)rb * (g.qR + 1) : g.thrR + (int64_t)(rb - g.modR) * g.qR;
  const int64_t rb_rows = rb < g.modR ? g.qR + 1 : g.qR;
  const int64_t r0 = rb_begin + (int64_t)chunk * rows_per_wg;
  const int64_t r1 = min(rb_begin + rb_rows, r0 + rows_per_wg);
  uint32_t n_lower = 0, n_upper = 0, n_diag = 0, n_close = 0;
  double max_off = 0.0, max_diag = 0.0;
  const int64_t diag_base = g.m - 1;
  uint32_t* bm = b.bitmap + (int64_t)rb * b.words;
  for (int64_t r = r0 + w; r < r1; r += kWaves) {
    const int s = rp[r], e = rp[r + 1];
    for (int j = s + lane; j < e; j += 64) {
      const int c = ci[j];
      const double a = fabs((double)val[j]);
      const int cb = cb_of(g, c);
      atomicAdd(&hist[w][cb], 1);
      atomicAdd(&b.col_cnt[c], 1u);
      const int64_t off = (int64_t)c - r;
      unsigned char* fl = b.dflag + (off + diag_base);
      if (*fl == 0) *fl = 1;
      if ((off < 0 ? -off : off) <= g.close_thr) n_close++;
      if (off == 0) {
        n_diag++;
        max_diag = fmax(max_diag, a);
      } else {
        if (off < 0) n_lower++;
        else n_upper++;
        max_off = fmax(max_off, a);
      }
      const double l = log10(a);
      if (l > 0.0) {
        if (b.colmax[c] < dbits(l)) atomicMax(&b.colmax[c], dbits(l));
      } else if (a > 0.0 && l < 0.0) {
        if (b.colminneg[c] < dbits(-l)) atomicMax(&b.colminneg[c], dbits(-l));
      }
      uint32_t* word = bm + (c >> 5);
      const uint32_t bit = 1u << (c & 31);
      if (!(*word & bit)) {
        const uint32_t old = atomicOr(word, bit);
        if (!(old & bit)) atomicAdd(&t_ucb[cb], 1u);
      }
    }
    // the LDS histogram updates of this wave complete before it reads them
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int k = lane; k < kT; k += 64) {
      const int cnt = hist[w][k];
      if (cnt) {
        hist[w][k] = 0;
        atomicAdd(&t_cnt[k], (uint32_t)cnt);
        atomicAdd(&t_urb[k], 1u);
        atomicMax(&t_max[k], (uint32_t)cnt);
        atomicAdd(&t_sq[k], (u64)cnt * cnt);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  const u64 lo = wave_sum<u64>(n_lower), up = wave_sum<u64>(n_upper), dg = wave_sum<u64>(n_diag),
            cl = wave_sum<u64>(n_close);
  const double mo = wave_max<double>(max_off), md = wave_max<double>(max_diag);
  if (lane == 0) {
    if (lo) atomicAdd(&b.acc[A_LOWER], lo);
    if (up) atomicAdd(&b.acc[A_UPPER], up);
    if (dg) atomicAdd(&b.acc[A_DIAG], dg);
    if (cl) atomicAdd(&b.acc[A_CLOSE], cl);
    if (lo + up) atomicMax(&b.acc[A_MAX_OFF], dbits(mo));
    if (dg) atomicMax(&b.acc[A_MAX_DIAG], dbits(md));
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kT; k += kBlock) {
    const int64_t t = (int64_t)rb * kT + k;
    if (t_cnt[k]) {
      atomicAdd(&b.tiles[t], t_cnt[k]);
      atomicAdd(&b.uniq_rb[t], t_urb[k]);
      atomicMax(&b.tile_max[t], t_max[k]);
      atomicAdd(&b.tile_sq[t], t_sq[k]);
    }
    if (t_ucb[k]) atomicAdd(&b.uniq_cb[t], t_ucb[k]);
  }
}

// One thread per column: counts, log10 range, column-block sums.
__global__ __launch_bounds__(kBlock) void feat_cols(Geo g, const uint32_t* __restrict__ col_cnt,
                                                    const u64* __restrict__ colmax, const u64* __restrict__ colminneg,
                                                    u64* __restrict__ acc, uint32_t* __restrict__ cb_nnz,
                                                    uint32_t* __restrict__ cb_max, u64* __restrict__ cb_sq) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool ok = c < g.n;
  const uint32_t cnt = ok ? col_cnt[c] : 0;
  const double mx = ok ? bitsd(colmax[c]) : 0.0;
  const double mn = ok ? -bitsd(colminneg[c]) : 0.0;
  const bool has_div = ok && mn != 0.0;
  const u64 nz = wave_sum<u64>(ok && cnt ? 1 : 0);
  const u64 cmin = wave_min<u64>(ok ? cnt : kMinInit);
  const u64 cmax = wave_max<u64>(ok ? cnt : 0);
  const u64 csq = wave_sum<u64>(ok ? (u64)cnt * cnt : 0);
  const u64 div = wave_max<u64>(has_div ? dbits(mx - mn) : 0);
  const u64 div_set = wave_sum<u64>(has_div ? 1 : 0);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&acc[A_NZ_COLS], nz);
    atomicMin(&acc[A_COL_MIN], cmin);
    atomicMax(&acc[A_COL_MAX], cmax);
    atomicAdd(&acc[A_COL_SQ], csq);
    if (div_set) {
      atomicMax(&acc[A_COL_DIV], div);
      atomicAdd(&acc[A_COL_DIV_SET], div_set);
    }
  }
  block_stats(ok, ok ? cb_of(g, (int)c) : 0, cnt, cb_nnz, cb_max, cb_sq);
}

__global__ __launch_bounds__(kBlock) void count_flags(const unsigned char* __restrict__ f, int64_t n,
                                                      u64* __restrict__ out) {
  u64 t = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) t += f[i] != 0;
  t = wave_sum<u64>(t);
  if ((threadIdx.x & 63) == 0 && t) atomicAdd(out, t);
}

// min, max, number of non-zeros, sum and sum of squares of a count vector.
enum { S_MIN, S_MAX, S_NZ, S_SUM, S_SQ, S_COUNT };
__global__ __launch_bounds__(kBlock) void vec_stats(const uint32_t* __restrict__ x, int64_t n, u64* __restrict__ out) {
  u64 mn = kMinInit, mx = 0, nz = 0, s = 0, q = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const u64 v = x[i];
    mn = min(mn, v);
    mx = max(mx, v);
    nz += v != 0;
    s += v;
    q += v * v;
  }
  mn = wave_min(mn); mx = wave_max(mx); nz = wave_sum(nz); s = wave_sum(s); q = wave_sum(q);
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&out[S_MIN], mn);
    atomicMax(&out[S_MAX], mx);
    atomicAdd(&out[S_NZ], nz);
    atomicAdd(&out[S_SUM], s);
    atomicAdd(&out[S_SQ], q);
  }
}

// Trapezoid sum over the cumulative counts, and the p-ratio stopping index.
// sorted: ascending counts; excl: their exclusive prefix sums.
__global__ __launch_bounds__(kBlock) void pratio_gini_kernel(const uint32_t* __restrict__ sorted,
                                                             const int64_t* __restrict__ excl, int64_t len,
                                                             int64_t num_nnzs, u64* __restrict__ out) {
  const int64_t total = excl[len - 1] + sorted[len - 1];
  u64 cum = 0, kmin = ~0ull;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len; i += (int64_t)gridDim.x * kBlock) {
    cum += (u64)(excl[i] + sorted[i]);
    const int64_t k = len - i;  // the k largest counts are sorted[i .. len)
    const double p = (double)k / (double)len;
    const double q = (double)(total - excl[i]) / (double)num_nnzs;
    if (!(p + q < 1.0)) kmin = min(kmin, (u64)k);
  }
  cum = wave_sum(cum);
  kmin = wave_min(kmin);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&out[0], cum);
    atomicMin(&out[1], kmin);
  }
}

// (n * Q - S^2) / n = sum of squared deviations from the mean S / n, from the
// exact integer moments.
__host__ __device__ inline double sq_dev(int64_t n, u64 S, u64 Q) {
  if (n <= 0) return 0.0;
  const unsigned __int128 a = (unsigned __int128)(u64)n * Q, b = (unsigned __int128)S * S;
  const unsigned __int128 d = a > b ? a - b : 0;
  const double dd = (double)(u64)(d >> 64) * 18446744073709551616.0 + (double)(u64)d;
  return dd / (double)n;
}

// Per-tile and per-block images: ave and std of the row (column) counts over
// the block's rows (columns).
__global__ __launch_bounds__(kBlock) void feat_images(Geo g, const uint32_t* __restrict__ tiles,
                                                      const u64* __restrict__ tile_sq, const uint32_t* __restrict__ rb_nnz,
                                                      const u64* __restrict__ rb_sq, const uint32_t* __restrict__ cb_nnz,
                                                      const u64* __restrict__ cb_sq, double* __restrict__ tile_ave,
                                                      double* __restrict__ tile_std, double* __restrict__ rb_ave,
                                                      double* __restrict__ rb_std, double* __restrict__ cb_ave,
                                                      double* __restrict__ cb_std) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < (int64_t)kT * kT) {
    const int tr = (int)(i / kT);
    const int64_t rows = tr < g.modR ? g.qR + 1 : g.qR;
    const double ave = rows ? (double)tiles[i] / (double)rows : 0.0;
    tile_ave[i] = ave;
    tile_std[i] = rows ? sqrt(sq_dev(rows, tiles[i], tile_sq[i]) / (double)rows) : 0.0;
  } else if (i < (int64_t)kT * kT + 2 * kT) {
    const bool col = i >= (int64_t)kT * kT + kT;
    const int k = (int)(i - (int64_t)kT * kT - (col ? kT : 0));
    const int64_t cnt = col ? (k < g.modC ? g.qC + 1 : g.qC) : (k < g.modR ? g.qR + 1 : g.qR);
    const uint32_t s = col ? cb_nnz[k] : rb_nnz[k];
    const u64 q = col ? cb_sq[k] : rb_sq[k];
    (col ? cb_ave : rb_ave)[k] = cnt ? (double)s / (double)cnt : 0.0;
    (col ? cb_std : rb_std)[k] = cnt ? sqrt(sq_dev(cnt, s, q) / (double)cnt) : 0.0;
  }
}

// ---- pattern / value symmetry ------------------------------------------------

__global__ void row_ids(const int32_t* __restrict__ rp, int64_t m, uint32_t* __restrict__ rid) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < m)
    for (int j = rp[r]; j < rp[r + 1]; j++) rid[j] = (uint32_t)r;
}

__global__ void iota_u32(int64_t n, uint32_t* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

__global__ void gather_u32(int64_t n, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ src,
                           uint32_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

template <class V>
__global__ void sorted_entries(int64_t n, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ rows,
                               const int32_t* __restrict__ ci, const V* __restrict__ val, u64* __restrict__ key,
                               double* __restrict__ sval) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const uint32_t p = perm[i];
    key[i] = ((u64)rows[i] << 32) | (uint32_t)ci[p];
    sval[i] = (double)val[p];
  }
}

struct Lf6 {
  u64 hi, lo;
  uint32_t tag;  // bit 0 sign, bit 1 compared as the raw magnitude
};

// The text printf("%lf") gives for a: sign and the value rounded half-to-even
// at six decimals, computed exactly from the binary significand.
__device__ Lf6 lf6_key(double a) {
  const u64 bits = dbits(a);
  const uint32_t sign = (uint32_t)(bits >> 63);
  const u64 mag = bits & 0x7fffffffffffffffull;
  const int ex = (int)(mag >> 52);
  if (ex == 0x7ff) return {(mag & 0x000fffffffffffffull) ? 1ull : 0ull, 0, sign | 2u | 4u};  // nan / inf
  if (ex >= 1023 + 52) return {mag, 0, sign | 2u};  // |a| >= 2^52: an integer, printed exactly
  const u64 m = ex ? ((mag & 0x000fffffffffffffull) | (1ull << 52)) : mag;
  const int s = ex ? 1075 - ex : 1074;  // a = m * 2^-s, s >= 1
  const unsigned __int128 t = (unsigned __int128)m * 1000000ull;  // < 2^73
  if (s >= 74) return {0, 0, sign};
  unsigned __int128 q = t >> s;
  const unsigned __int128 rem = t & ((((unsigned __int128)1) << s) - 1);
  const unsigned __int128 half = ((unsigned __int128)1) << (s - 1);
  if (rem > half || (rem == half && (q & 1))) q += 1;
  return {(u64)(q >> 64), (u64)q, sign};
}

__device__ __forceinline__ int64_t lower_bound_u64(const u64* __restrict__ k, int64_t n, u64 x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (k[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void symmetry_pairs(int64_t n, const u64* __restrict__ key,
                                                         const double* __restrict__ sval, u64* __restrict__ acc) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  u64 pat = 0, vs = 0;
  if (i < n) {
    const u64 k = key[i];
    const uint32_t r = (uint32_t)(k >> 32), c = (uint32_t)k;
    if (r != c) {
      const u64 mirror = ((u64)c << 32) | r;
      const int64_t lm = lower_bound_u64(key, n, mirror);
      if (lm < n && key[lm] == mirror) {
        pat = 1;
        const int64_t ls = lower_bound_u64(key, n, k);
        const Lf6 a = lf6_key(sval[lm]), b = lf6_key(sval[ls]);
        vs = a.hi == b.hi && a.lo == b.lo && a.tag == b.tag;
      }
    }
  }
  pat = wave_sum(pat);
  vs = wave_sum(vs);
  if ((threadIdx.x & 63) == 0) {
    if (pat) atomicAdd(&acc[A_PATTERN], pat);
    if (vs) atomicAdd(&acc[A_VALUE], vs);
  }
}

double bits_to_double(u64 b) {
  double d;
  std::memcpy(&d, &b, sizeof d);
  return d;
}

int bits_for(u64 maxval) {
  int b = 0;
  while (b < 64 && (maxval >> b)) b++;
  return b;
}

struct PG {
  double p, gini;
};

// P_ratioAndGini of a device count vector of `len` entries (len >= 1).
PG pratio_gini(const uint32_t* counts, int64_t len, u64 maxval, int64_t num_nnzs, hipStream_t s) {
  DBuf sorted, excl, out;
  sorted.alloc(sizeof(uint32_t) * len);
  excl.alloc(sizeof(int64_t) * len);
  out.alloc(sizeof(u64) * 2);
  LSP_HIP(hipMemcpyAsync(sorted.p, counts, sizeof(uint32_t) * len, hipMemcpyDeviceToDevice, s));
  radix_sort_u32(sorted.as<uint32_t>(), len, bits_for(maxval), s);
  scan_exclusive_i32_to_i64(sorted.as<int32_t>(), excl.as<int64_t>(), len, s);
  const u64 init[2] = {0, ~0ull};
  LSP_HIP(hipMemcpyAsync(out.p, init, sizeof init, hipMemcpyHostToDevice, s));
  const int grid = (int)std::min<int64_t>(ceil_div(len, kBlock), 4096);
  hipLaunchKernelGGL(pratio_gini_kernel, dim3(grid), dim3(kBlock), 0, s, sorted.as<uint32_t>(),
                     excl.as<int64_t>(), len, num_nnzs, out.as<u64>());
  LSP_LAUNCHED();
  const auto h = to_host<u64>(out.p, 2, s);
  const auto last = to_host<int64_t>(excl.as<int64_t>() + len - 1, 1, s);
  const auto slast = to_host<uint32_t>(sorted.as<uint32_t>() + len - 1, 1, s);
  const int64_t total = last[0] + slast[0];
  PG r;
  r.p = (double)(int64_t)h[1] / (double)len;
  // trapezoid over the cumulative sums with unit steps: sum(cum) - cum[last]/2
  const double area = (double)(int64_t)(2 * h[0] - (u64)total) / 2.0;
  const double area_total = (double)num_nnzs * (double)len / 2.0;
  r.gini = (area_total - area) / area_total;
  return r;
}

std::vector<u64> stats_of(const uint32_t* x, int64_t n, hipStream_t s) {
  DBuf out;
  out.alloc(sizeof(u64) * S_COUNT);
  const u64 init[S_COUNT] = {kMinInit, 0, 0, 0, 0};
  LSP_HIP(hipMemcpyAsync(out.p, init, sizeof init, hipMemcpyHostToDevice, s));
  const int grid = (int)std::min<int64_t>(ceil_div(n, kBlock), 2048);
  hipLaunchKernelGGL(vec_stats, dim3(grid), dim3(kBlock), 0, s, x, n, out.as<u64>());
  LSP_LAUNCHED();
  return to_host<u64>(out.p, S_COUNT, s);
}

template <class V>
void symmetry_counts(Matrix& A, u64* acc, hipStream_t s) {
  const int64_t nnz = A.nnz;
  DBuf rows, keys, perm, skey, sval;
  rows.alloc(sizeof(uint32_t) * nnz);
  keys.alloc(sizeof(uint32_t) * nnz);
  perm.alloc(sizeof(uint32_t) * nnz);
  skey.alloc(sizeof(u64) * nnz);
  sval.alloc(sizeof(double) * nnz);
  hipLaunchKernelGGL(row_ids, grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.rp.as<int32_t>(), A.m,
                     rows.as<uint32_t>());
  LSP_LAUNCHED();
  hipLaunchKernelGGL(iota_u32, grid1d(nnz, kBlock), dim3(kBlock), 0, s, nnz, perm.as<uint32_t>());
  LSP_LAUNCHED();
  // stable by column, then stable by row: (row, col, original position)
  LSP_HIP(hipMemcpyAsync(keys.p, A.ci.p, sizeof(uint32_t) * nnz, hipMemcpyDeviceToDevice, s));
  radix_sort_pairs(keys.as<uint32_t>(), perm.as<uint32_t>(), nnz, bits_for((u64)A.n), s);
  hipLaunchKernelGGL(gather_u32, grid1d(nnz, kBlock), dim3(kBlock), 0, s, nnz, perm.as<uint32_t>(),
                     rows.as<uint32_t>(), keys.as<uint32_t>());
  LSP_LAUNCHED();
  radix_sort_pairs(keys.as<uint32_t>(), perm.as<uint32_t>(), nnz, bits_for((u64)A.m), s);
  hipLaunchKernelGGL((sorted_entries<V>), grid1d(nnz, kBlock), dim3(kBlock), 0, s, nnz, perm.as<uint32_t>(),
                     keys.as<uint32_t>(), A.ci.as<int32_t>(), A.val.as<V>(), skey.as<u64>(), sval.as<double>());
  LSP_LAUNCHED();
  hipLaunchKernelGGL(symmetry_pairs, grid1d(nnz, kBlock), dim3(kBlock), 0, s, nnz, skey.as<u64>(),
                     sval.as<double>(), acc);
  LSP_LAUNCHED();
}

template <class V>
void compute(Matrix& A, int32_t flags, lespmv_features_t& F, lespmv_feature_images_t* img) {
  hipStream_t s = A.stream;
  const int64_t M = A.m, N = A.n, nnz = A.nnz;
  LSP_REQUIRE(M >= 1 && N >= 1, LESPMV_ERR_ARG, "features need at least one row and one column");
  const bool symmetric_file = (flags & LESPMV_FEAT_SYMMETRIC_FILE) != 0;
  LSP_REQUIRE(!symmetric_file || M == N, LESPMV_ERR_ARG, "a symmetric MatrixMarket matrix is square");
  Geo g;
  g.m = M;
  g.n = N;
  g.qR = (int)(M / kT); g.modR = (int)(M % kT); g.thrR = g.modR * (g.qR + 1);
  g.qC = (int)(N / kT); g.modC = (int)(N % kT); g.thrC = g.modC * (g.qC + 1);
  g.close_thr = (M + N - 1) / 20;

  LSP_HIP(hipEventRecord(A.ev0, s));
  const int64_t T2 = (int64_t)kT * kT;
  DBuf tiles, uniq_rb, uniq_cb, tile_max, tile_sq, col_cnt, dflag, colmax, colminneg, bitmap, acc;
  DBuf rowlen, rbn, rbm, rbq, cbn, cbm, cbq, dist_part, dist;
  tiles.alloc(4 * T2); uniq_rb.alloc(4 * T2); uniq_cb.alloc(4 * T2); tile_max.alloc(4 * T2); tile_sq.alloc(8 * T2);
  col_cnt.alloc(4 * N);
  dflag.alloc(M + N - 1);
  colmax.alloc(8 * N); colminneg.alloc(8 * N);
  const int64_t words = ceil_div(N, 32);
  bitmap.alloc(4 * words * kT);
  acc.alloc(8 * A_COUNT);
  rowlen.alloc(4 * M);
  rbn.alloc(4 * kT); rbm.alloc(4 * kT); rbq.alloc(8 * kT); cbn.alloc(4 * kT); cbm.alloc(4 * kT); cbq.alloc(8 * kT);
  const int64_t row_blocks = ceil_div(M, kBlock);
  dist_part.alloc(8 * row_blocks);
  dist.alloc(8);
  for (DBuf* d : {&tiles, &uniq_rb, &uniq_cb, &tile_max, &tile_sq, &col_cnt, &dflag, &colmax, &colminneg, &bitmap,
                  &rbn, &rbm, &rbq, &cbn, &cbm, &cbq})
    LSP_HIP(hipMemsetAsync(d->p, 0, d->bytes, s));
  std::vector<u64> acc0(A_COUNT, 0);
  acc0[A_ROW_MIN] = acc0[A_COL_MIN] = kMinInit;
  LSP_HIP(hipMemcpyAsync(acc.p, acc0.data(), 8 * A_COUNT, hipMemcpyHostToDevice, s));

  hipLaunchKernelGGL((feat_rows<V>), dim3((unsigned)row_blocks), dim3(kBlock), 0, s, g, A.rp.as<int32_t>(),
                     A.ci.as<int32_t>(), A.val.as<V>(), rowlen.as<uint32_t>(), acc.as<u64>(), rbn.as<uint32_t>(),
                     rbm.as<uint32_t>(), rbq.as<u64>(), dist_part.as<double>());
  LSP_LAUNCHED();
  hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kBlock), 0, s, dist_part.as<double>(), row_blocks,
                     dist.as<double>());
  LSP_LAUNCHED();

  const int64_t rb_rows = g.qR + 1;
  const int rows_per_wg = (int)std::clamp<int64_t>(ceil_div(rb_rows, 16), 4, 1024);
  const int chunks = (int)ceil_div(rb_rows, rows_per_wg);
  TileBufs tb{tiles.as<uint32_t>(), uniq_rb.as<uint32_t>(), uniq_cb.as<uint32_t>(), tile_max.as<uint32_t>(),
              tile_sq.as<u64>(), col_cnt.as<uint32_t>(), dflag.as<unsigned char>(), colmax.as<u64>(),
              colminneg.as<u64>(), bitmap.as<uint32_t>(), words, acc.as<u64>()};
  hipLaunchKernelGGL((feat_tiles<V>), dim3((unsigned)(kT * chunks)), dim3(kBlock), 0, s, g, A.rp.as<int32_t>(),
                     A.ci.as<int32_t>(), A.val.as<V>(), chunks, rows_per_wg, tb);
  LSP_LAUNCHED();
  hipLaunchKernelGGL(feat_cols, grid1d(N, kBlock), dim3(kBlock), 0, s, g, col_cnt.as<uint32_t>(), colmax.as<u64>(),
                     colminneg.as<u64>(), acc.as<u64>(), cbn.as<uint32_t>(), cbm.as<uint32_t>(), cbq.as<u64>());
  LSP_LAUNCHED();
  hipLaunchKernelGGL(count_flags, dim3((unsigned)std::min<int64_t>(ceil_div(M + N - 1, kBlock), 2048)),
                     dim3(kBlock), 0, s, dflag.as<unsigned char>(), M + N - 1, acc.as<u64>() + A_NDIAGS);
  LSP_LAUNCHED();
  if (!symmetric_file && M == N && nnz > 0) symmetry_counts<V>(A, acc.as<u64>(), s);

  DBuf tile_ave, tile_std, rb_ave, rb_std, cb_ave, cb_std;
  tile_ave.alloc(8 * T2); tile_std.alloc(8 * T2);
  rb_ave.alloc(8 * kT); rb_std.alloc(8 * kT); cb_ave.alloc(8 * kT); cb_std.alloc(8 * kT);
  hipLaunchKernelGGL(feat_images, grid1d(T2 + 2 * kT, kBlock), dim3(kBlock), 0, s, g, tiles.as<uint32_t>(),
                     tile_sq.as<u64>(), rbn.as<uint32_t>(), rbq.as<u64>(), cbn.as<uint32_t>(), cbq.as<u64>(),
                     tile_ave.as<double>(), tile_std.as<double>(), rb_ave.as<double>(), rb_std.as<double>(),
                     cb_ave.as<double>(), cb_std.as<double>());
  LSP_LAUNCHED();

  const auto a = to_host<u64>(acc.p, A_COUNT, s);
  const auto st_t = stats_of(tiles.as<uint32_t>(), T2, s);
  const auto st_rb = stats_of(rbn.as<uint32_t>(), kT, s);
  const auto st_cb = stats_of(cbn.as<uint32_t>(), kT, s);
  const auto st_urb = stats_of(uniq_rb.as<uint32_t>(), T2, s);
  const auto st_ucb = stats_of(uniq_cb.as<uint32_t>(), T2, s);
  const PG pg_row = pratio_gini(rowlen.as<uint32_t>(), M, a[A_ROW_MAX], nnz, s);
  const PG pg_col = pratio_gini(col_cnt.as<uint32_t>(), N, a[A_COL_MAX], nnz, s);
  const PG pg_t = pratio_gini(tiles.as<uint32_t>(), T2, st_t[S_MAX], nnz, s);
  const PG pg_rb = pratio_gini(rbn.as<uint32_t>(), kT, st_rb[S_MAX], nnz, s);
  const PG pg_cb = pratio_gini(cbn.as<uint32_t>(), kT, st_cb[S_MAX], nnz, s);
  const double dist_sum = to_host<double>(dist.p, 1, s)[0];

  if (img) {
    auto copy = [&](void* dst, const DBuf& src, size_t bytes) {
      if (dst) LSP_HIP(hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToHost, s));
    };
    copy(img->tile_nnz, tiles, 4 * T2); copy(img->tile_max, tile_max, 4 * T2);
    copy(img->tile_ave, tile_ave, 8 * T2); copy(img->tile_std, tile_std, 8 * T2);
    copy(img->tile_uniq_rows, uniq_rb, 4 * T2); copy(img->tile_uniq_cols, uniq_cb, 4 * T2);
    copy(img->rb_nnz, rbn, 4 * kT); copy(img->rb_max, rbm, 4 * kT); copy(img->rb_ave, rb_ave, 8 * kT);
    copy(img->rb_std, rb_std, 8 * kT);
    copy(img->cb_nnz, cbn, 4 * kT); copy(img->cb_max, cbm, 4 * kT); copy(img->cb_ave, cb_ave, 8 * kT);
    copy(img->cb_std, cb_std, 8 * kT);
  }
  LSP_HIP(hipEventRecord(A.ev1, s));
  LSP_HIP(hipEventSynchronize(A.ev1));
  float ms = 0.f;
  LSP_HIP(hipEventElapsedTime(&ms, A.ev0, A.ev1));

  // host-side ratios from the exact aggregates (the reference's formulas)
  std::memset(&F, 0, sizeof F);
  const double nn = (double)nnz;
  F.num_rows = M; F.num_cols = N; F.num_nnzs = nnz;
  F.nnz_lower = (int64_t)a[A_LOWER]; F.nnz_upper = (int64_t)a[A_UPPER]; F.nnz_diagonal = (int64_t)a[A_DIAG];
  F.nnz_mtx = symmetric_file ? F.nnz_diagonal + F.nnz_lower : nnz;
  F.complete_ndiags = (int64_t)a[A_NDIAGS]; F.close_nnz = (int64_t)a[A_CLOSE];
  F.min_nnz_each_row = (int64_t)a[A_ROW_MIN]; F.max_nnz_each_row = (int64_t)a[A_ROW_MAX];
  F.min_nnz_each_col = (int64_t)a[A_COL_MIN]; F.max_nnz_each_col = (int64_t)a[A_COL_MAX];
  F.nz_rows = (int64_t)a[A_NZ_ROWS]; F.nz_cols = (int64_t)a[A_NZ_COLS]; F.diag_dominant_rows = (int64_t)a[A_DOMINANT];
  F.t_num_blocks = kT; F.t_num_RB = g.qR; F.t_num_CB = g.qC; F.t_mod_RB = g.modR; F.t_mod_CB = g.modC;
  F.t_min_nnz_all_tiles = (int64_t)st_t[S_MIN]; F.t_max_nnz_all_tiles = (int64_t)st_t[S_MAX];
  F.t_min_nnz_each_RB = (int64_t)st_rb[S_MIN]; F.t_max_nnz_each_RB = (int64_t)st_rb[S_MAX];
  F.t_min_nnz_each_CB = (int64_t)st_cb[S_MIN]; F.t_max_nnz_each_CB = (int64_t)st_cb[S_MAX];
  F.uniq_rows_sum = (int64_t)st_urb[S_SUM]; F.uniq_cols_sum = (int64_t)st_ucb[S_SUM];
  F.symm_pattern_pairs = (int64_t)a[A_PATTERN]; F.symm_value_pairs = (int64_t)a[A_VALUE];

  if (symmetric_file) {
    F.pattern_symm = F.value_symm = 1.0;
  } else if (M == N) {
    const double offd = (double)(F.nnz_mtx - F.nnz_diagonal);
    F.pattern_symm = (double)F.symm_pattern_pairs / offd;
    F.value_symm = (double)F.symm_value_pairs / offd;
  }
  F.is_symmetric = F.pattern_symm == 1.0 && F.value_symm == 1.0;
  F.nnz_ratio = nn / ((double)M * (double)N);
  F.diag_close_ratio = (double)F.close_nnz / nn;
  F.distance_per_row = dist_sum / (double)M;
  F.nz_row_ratio = (double)F.nz_rows / (double)M;
  F.ave_nnz_each_row = nn / (double)M;
  F.var_nnz_each_row = sq_dev(M, (u64)nnz, a[A_ROW_SQ]) / (double)M;
  F.standard_dev_row = std::sqrt(F.var_nnz_each_row);
  F.P_ratio_row = pg_row.p; F.Gini_row = pg_row.gini;
  F.nz_col_ratio = (double)F.nz_cols / (double)N;
  F.ave_nnz_each_col = nn / (double)N;
  F.var_nnz_each_col = sq_dev(N, (u64)nnz, a[A_COL_SQ]) / (double)N;
  F.standard_dev_col = std::sqrt(F.var_nnz_each_col);
  F.P_ratio_col = pg_col.p; F.Gini_col = pg_col.gini;
  F.diagonal_dominant_ratio = (double)F.diag_dominant_rows / (double)std::min(M, N);
  F.max_value_offdiag = (a[A_LOWER] + a[A_UPPER]) ? bits_to_double(a[A_MAX_OFF]) : -DBL_MAX;
  F.max_value_diagonal = a[A_DIAG] ? bits_to_double(a[A_MAX_DIAG]) : -DBL_MAX;
  F.row_variability = a[A_ROW_DIV_SET] ? bits_to_double(a[A_ROW_DIV]) : -1.0;
  F.col_variability = a[A_COL_DIV_SET] ? bits_to_double(a[A_COL_DIV]) : -1.0;
  F.t_ave_nnz_all_tiles = nn / ((double)kT * kT);
  F.t_ave_nnz_RB = nn / kT;
  F.t_ave_nnz_CB = nn / kT;
  F.t_var_nnz_all_tiles = (-1.0 + sq_dev(T2, st_t[S_SUM], st_t[S_SQ])) / ((double)kT * kT);
  F.t_var_nnz_RB = (-1.0 + sq_dev(kT, st_rb[S_SUM], st_rb[S_SQ])) / (double)kT;
  F.t_var_nnz_CB = (-1.0 + sq_dev(kT, st_cb[S_SUM], st_cb[S_SQ])) / (double)kT;
  F.t_standard_dev_all_tiles = std::sqrt(std::max(F.t_var_nnz_all_tiles, 0.0));
  F.t_standard_dev_RB = std::sqrt(std::max(F.t_var_nnz_RB, 0.0));
  F.t_standard_dev_CB = std::sqrt(std::max(F.t_var_nnz_CB, 0.0));
  F.t_Gini_all_tiles = pg_t.gini; F.t_Gini_RB = pg_rb.gini; F.t_Gini_CB = pg_cb.gini;
  F.t_P_ratio_all_tiles = pg_t.p; F.t_P_ratio_RB = pg_rb.p; F.t_P_ratio_CB = pg_cb.p;
  F.t_nz_ratio_tiles = (-1.0 + (double)st_t[S_NZ]) / ((double)kT * kT);
  F.t_nz_ratio_RB = (-1.0 + (double)st_rb[S_NZ]) / (double)kT;
  F.t_nz_ratio_CB = (-1.0 + (double)st_cb[S_NZ]) / (double)kT;
  F.uniqR = (double)F.uniq_rows_sum / nn;
  F.uniqC = (double)F.uniq_cols_sum / nn;
  F.potReuseR = (double)F.uniq_rows_sum / (double)M;
  F.potReuseC = (double)F.uniq_cols_sum / (double)N;
  F.GrX_uniqR = F.GrX_uniqC = F.GrX_potReuseR = F.GrX_potReuseC = -1.0;
  F.device_ms = ms;
}

// ---- block occupancy (selector input, no reference counterpart) -------------

__global__ void blk_keys(int64_t m, int br, int bc, u64 nb, const int32_t* __restrict__ rp,
                         const int32_t* __restrict__ ci, u64* __restrict__ keys) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < m) {
    const u64 row_key = (u64)(r / br) * nb;
    for (int j = rp[r]; j < rp[r + 1]; j++) keys[j] = row_key + (u64)(ci[j] / bc);
  }
}

__global__ __launch_bounds__(kBlock) void count_unique(const u64* __restrict__ keys, int64_t n, u64* __restrict__ out) {
  u64 t = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    t += (i == 0 || keys[i] != keys[i - 1]);
  t = wave_sum(t);
  if ((threadIdx.x & 63) == 0 && t) atomicAdd(out, t);
}

// .features line names in file order (src/sparse_features.cpp:1347-1400).
const char* const kNames[LESPMV_NUM_FEATURES] = {
    "num_rows", "num_cols", "num_nnz", "nnz_ratio", "nnz_on_maindiag_ratio", "nnz_row_distance_ratio",
    "nz_row_ratio", "max_nnz_each_row", "ave_nnz_each_row", "standard_dev_row", "P-ratio_row", "Gini_coeff_row",
    "nz_col_ratio", "max_nnz_each_col", "ave_nnz_each_col", "standard_dev_col", "P-ratio_col", "Gini_coeff_col",
    "nz_tiles_ratio", "max_nnz_each_tiles", "ave_nnz_each_tiles", "standard_dev_tiles", "P-ratio_tiles",
    "Gini_coeff_tiles", "nz_ratio_RB", "max_nnz_each_RB", "ave_nnz_each_RB", "standard_dev_RB", "P-ratio_RB",
    "Gini_coeff_RB", "nz_ratio_CB", "max_nnz_each_CB", "ave_nnz_each_CB", "standard_dev_CB", "P-ratio_CB",
    "Gini_coeff_CB", "uniq_R", "uniq_C", "potReuseR", "potReuseC"};

}  // namespace

}  // namespace lsp

using namespace lsp;

extern "C" int lespmv_features(lespmv_matrix_t h, int32_t flags, lespmv_features_t* out,
                               lespmv_feature_images_t* images) {
  return lsp::guarded_call([&] {
    LSP_REQUIRE(h && out, LESPMV_ERR_ARG, "null argument");
    Matrix& A = *reinterpret_cast<Matrix*>(h);
    LSP_REQUIRE(!A.idx64, LESPMV_ERR_UNSUPPORTED, "a handle with 64-bit row offsets runs CSR SpMV and SpMM only");
    if (A.dtype == LESPMV_F64) compute<double>(A, flags, *out, images);
    else compute<float>(A, flags, *out, images);
  });
}

extern "C" int lespmv_features_vector(const lespmv_features_t* f, double* o) {
  return lsp::guarded_call([&] {
    LSP_REQUIRE(f && o, LESPMV_ERR_ARG, "null argument");
    const double v[LESPMV_NUM_FEATURES] = {
        (double)f->num_rows, (double)f->num_cols, (double)f->num_nnzs, f->nnz_ratio, f->diag_close_ratio,
        f->distance_per_row / (double)f->num_cols, f->nz_row_ratio, (double)f->max_nnz_each_row,
        f->ave_nnz_each_row, f->standard_dev_row, f->P_ratio_row, f->Gini_row, f->nz_col_ratio,
        (double)f->max_nnz_each_col, f->ave_nnz_each_col, f->standard_dev_col, f->P_ratio_col, f->Gini_col,
        f->t_nz_ratio_tiles, (double)f->t_max_nnz_all_tiles, f->t_ave_nnz_all_tiles, f->t_standard_dev_all_tiles,
        f->t_P_ratio_all_tiles, f->t_Gini_all_tiles, f->t_nz_ratio_RB, (double)f->t_max_nnz_each_RB,
        f->t_ave_nnz_RB, f->t_standard_dev_RB, f->t_P_ratio_RB, f->t_Gini_RB, f->t_nz_ratio_CB,
        (double)f->t_max_nnz_each_CB, f->t_ave_nnz_CB, f->t_standard_dev_CB, f->t_P_ratio_CB, f->t_Gini_CB,
        f->uniqR, f->uniqC, f->potReuseR, f->potReuseC};
    std::memcpy(o, v, sizeof v);
  });
}

extern "C" int lespmv_block_counts(lespmv_matrix_t h, const int32_t* dims, int32_t n, int64_t* nnzb) {
  return lsp::guarded_call([&] {
    LSP_REQUIRE(h && (n == 0 || (dims && nnzb)), LESPMV_ERR_ARG, "null argument");
    Matrix& A = *reinterpret_cast<Matrix*>(h);
    LSP_REQUIRE(!A.idx64, LESPMV_ERR_UNSUPPORTED, "a handle with 64-bit row offsets runs CSR SpMV and SpMM only");
    hipStream_t s = A.stream;
    DBuf keys, out;
    keys.alloc(sizeof(u64) * std::max<int64_t>(A.nnz, 1));
    out.alloc(sizeof(u64));
    for (int32_t i = 0; i < n; i++) {
      const int br = dims[2 * i], bc = dims[2 * i + 1];
      LSP_REQUIRE(br > 0 && bc > 0, LESPMV_ERR_ARG, "block dimensions must be positive");
      const u64 mb = (u64)ceil_div(A.m, br), nb = (u64)ceil_div(A.n, bc);
      if (A.nnz == 0) {
        nnzb[i] = 0;
        continue;
      }
      hipLaunchKernelGGL(blk_keys, grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, br, bc, nb, A.rp.as<int32_t>(),
                         A.ci.as<int32_t>(), keys.as<u64>());
      LSP_LAUNCHED();
      radix_sort_u64(keys.as<uint64_t>(), A.nnz, bits_for(mb * nb), s);
      LSP_HIP(hipMemsetAsync(out.p, 0, sizeof(u64), s));
      hipLaunchKernelGGL(count_unique, dim3((unsigned)std::min<int64_t>(ceil_div(A.nnz, kBlock), 2048)),
                         dim3(kBlock), 0, s, keys.as<u64>(), A.nnz, out.as<u64>());
      LSP_LAUNCHED();
      nnzb[i] = (int64_t)to_host<u64>(out.p, 1, s)[0];
    }
  });
}

extern "C" const char* lespmv_feature_name(int32_t i) {
  return (i >= 0 && i < LESPMV_NUM_FEATURES) ? kNames[i] : nullptr;
}
