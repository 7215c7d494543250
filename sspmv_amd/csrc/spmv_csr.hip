// CSR SpMV / SpMM kernels for gfx950.
//
// Reference semantics: LeSpMV_csr (src/spmv_csr.cpp:164-185) computes
// y = alpha*A*x + beta*y row by row; kernel_flag 2 balances rows by nonzeros
// (src/spmv_csr.cpp:137-162, utils/partition.cpp:58-75). On the GPU the two
// kernels are:
//  * vector: L lanes of a 64-wide wavefront per row (L = 2..64 tuned from the
//    row-length distribution), unit-stride loads of col/val, DPP/permute
//    reduction inside the L-lane group;
//  * merge: merge-path tiles of 1024 (row end + nonzero) items, so every
//    workgroup gets the same amount of work whatever the row lengths. The tile's
//    products are staged in LDS, complete rows are reduced from LDS, and the row
//    open at a tile edge is carried to a small deterministic fix-up kernel.
// SpMM has a plain kernel (a group of k lanes per row) and the row engine of
// rows_engine.hpp (variants 16 + log2 L).
#include <algorithm>

#include "buffer_loads.hpp"
#include "common.hpp"
#include "rows_engine.hpp"

namespace lsp {
namespace {

constexpr int kBlock = 256;
// merge-path items (row ends + entries) per tile of the CSR kernel:
// kMergeItems = 2048 on 256 threads, eight entries per thread (with x
// gathered through a buffer resource; 27-point stencil 160^3 FP64 305 us
// against 323 with 1024 on 256 and 313 with 2048 on 512, FP32 226 against
// 260; banded rows FP32 206 against 262; lognormal rows 851 against 847)
// items per tile of the column-panel kernel and threads of its workgroup:
// 2048 items on 256 threads, eight entries per thread (FP64 lognormal rows
// with x gathered through a buffer resource: 487 against 500 us on 512
// threads; FP32 378 either way; 2048 is also the most a tile's row-end
// words, one per lane of a wave, can cover)
template <class V>
constexpr int kPanelItems = 2048;
template <class V>
constexpr int kPanelThreads = 256;

template <int L, class V>
__device__ __forceinline__ V group_sum(V v) {
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, L);
  return v;
}

template <class V>
__device__ __forceinline__ V axpby(V alpha, V sum, V beta, const V* y) {
  return beta == V(0) ? alpha * sum : alpha * sum + beta * *y;
}

template <class V, int L>
__global__ __launch_bounds__(kBlock) void csr_vector_kernel(int64_t m, const int* __restrict__ rp,
                                                            const int* __restrict__ ci, const V* __restrict__ val,
                                                            const V* __restrict__ x, V* __restrict__ y, V alpha,
                                                            V beta) {
  const int lane = threadIdx.x & (L - 1);
  const int64_t row = xcd_block(blockIdx.x, gridDim.x) * (kBlock / L) + threadIdx.x / L;
  if (row >= m) return;  // whole L-lane groups leave together
  const int s = rp[row], e = rp[row + 1];
  V sum = 0, sum2 = 0;
  int j = s + lane;
  // whole rounds of 4L entries (the same count for every lane of the row),
  // then the rest of the row in one round with its loads masked past the end:
  // short rows get all their loads in flight at once
  for (; j - lane + 4 * L <= e; j += 4 * L) {
    const int c0 = ci[j], c1 = ci[j + L], c2 = ci[j + 2 * L], c3 = ci[j + 3 * L];
    const V v0 = val[j], v1 = val[j + L], v2 = val[j + 2 * L], v3 = val[j + 3 * L];
    sum += v0 * x[c0] + v1 * x[c1];
    sum2 += v2 * x[c2] + v3 * x[c3];
  }
  if (j < e) {
    const bool b1 = j + L < e, b2 = j + 2 * L < e, b3 = j + 3 * L < e;
    const int c0 = ci[j], c1 = b1 ? ci[j + L] : 0, c2 = b2 ? ci[j + 2 * L] : 0, c3 = b3 ? ci[j + 3 * L] : 0;
    const V v0 = val[j], v1 = b1 ? val[j + L] : V(0), v2 = b2 ? val[j + 2 * L] : V(0);
    const V v3 = b3 ? val[j + 3 * L] : V(0);
    const V x0 = x[c0], x1 = b1 ? x[c1] : V(0), x2 = b2 ? x[c2] : V(0), x3 = b3 ? x[c3] : V(0);
    sum += v0 * x0 + v1 * x1;
    sum2 += v2 * x2 + v3 * x3;
  }
  sum = group_sum<L>(sum + sum2);
  if (lane == 0) y[row] = axpby(alpha, sum, beta, y + row);
}

// Number of row ends consumed at merge-path diagonal d (row end r precedes
// nonzero j iff rp[r+1] <= j).
__device__ int64_t merge_search(int64_t d, const int* rp, int64_t m, int64_t nnz) {
  int64_t lo = d - nnz > 0 ? d - nnz : 0, hi = d < m ? d : m;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int64_t)rp[mid + 1] <= d - mid - 1) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void merge_partition_kernel(int64_t tiles, int64_t m, int64_t nnz, const int* __restrict__ rp,
                                       int* __restrict__ coord, int64_t items) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > tiles) return;
  int64_t d = t * items;
  if (d > m + nnz) d = m + nnz;
  const int64_t r = merge_search(d, rp, m, nnz);
  coord[2 * t] = (int)r;
  coord[2 * t + 1] = (int)(d - r);
}

// Tiles cut at row starts: the merge-path cut of tile t (every `items`
// diagonals) moved back to the start of the row it falls in. A tile then
// holds whole rows only, at most items + the longest row + 1 merge-path items.
__global__ void row_partition_kernel(int64_t tiles, int64_t m, int64_t nnz, const int* __restrict__ rp,
                                     int* __restrict__ coord, int64_t items) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > tiles) return;
  int64_t d = t * items;
  if (d > m + nnz) d = m + nnz;
  const int64_t r = merge_search(d, rp, m, nnz);
  coord[2 * t] = (int)r;
  coord[2 * t + 1] = rp[r];
}

template <class V, int BT = kBlock>
__device__ V block_sum(V v, V* red) {
  v = group_sum<64>(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  V t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < BT / 64; i++) t += red[i];
  return t;
}

constexpr int kLongRow = 48;                     // rows longer than this in a tile are reduced by a wavefront

// Loads of data that is read once: with NT the line is marked for early
// eviction so it does not push reused data (x) out of the L2.
template <bool NT, class T>
__device__ __forceinline__ T stream_load(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// One merge-path tile b of a CSR matrix (rp, ci, val are the matrix's own
// arrays, coord its tile starts). With BITS the tile's row boundaries come
// from `ends`, one bit per merge-path item set at row ends (row r ends at item
// r + rp[r + 1]), instead of from the row pointers: ITEMS / 32 words per tile
// against four bytes per row. With COMPACT the column of entry e (an index
// into the panel run, x already offset to the panel's first column) is
// ci16[e] | (nibble e of cihi) << 16: 2.5 bytes per entry instead of 4.
// When x's extent xbytes (bytes from x to the last column any entry names)
// is below 2^32, x is gathered through a buffer resource.
template <class V, bool NT, int ITEMS, bool BITS = false, bool COMPACT = false, int BT = kBlock>
__device__ __forceinline__ void merge_tile(int b, int64_t m, const int* __restrict__ rp,
                                           const int* __restrict__ ci, const V* __restrict__ val,
                                           const V* __restrict__ x, V* __restrict__ y, V alpha, V beta,
                                           const int* __restrict__ coord, int* __restrict__ carry_row,
                                           V* __restrict__ carry_val, V* __restrict__ head_val,
                                           const uint32_t* __restrict__ ends = nullptr,
                                           const uint16_t* __restrict__ ci16 = nullptr,
                                           const uint8_t* __restrict__ cihi = nullptr, int64_t off = 0,
                                           int64_t xbytes = 0) {
  constexpr int Q = ITEMS / BT;  // entries staged per thread (BT threads in the workgroup)
  __shared__ V prod[ITEMS];
  __shared__ int rows[ITEMS + 2];
  __shared__ V red[BT / 64];
  __shared__ int long_rows[ITEMS / kLongRow + 2];
  __shared__ int nlong;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r0 = coord[2 * b], z0 = coord[2 * b + 1];
  const int r1 = coord[2 * b + 2], z1 = coord[2 * b + 3];
  const int nz = z1 - z0;
  const int last = r1 + 1 < m ? r1 + 1 : (int)m;  // rp[r0 .. last] staged
  if (tid == 0) nlong = 0;
  // every load of the tile is issued before the first gather returns
  int cidx[Q];
  V v[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int i = tid + q * BT;
    const bool ok = i < nz;
    if constexpr (COMPACT) {
      const int64_t e = off + z0 + i;
      const uint32_t lo = ok ? (uint32_t)stream_load<NT>(ci16 + e) : 0u;
      const uint32_t hb = ok ? (uint32_t)stream_load<NT>(cihi + (e >> 1)) : 0u;
      cidx[q] = (int)((((hb >> ((e & 1) << 2)) & 15u) << 16) | lo);
    } else {
      cidx[q] = ok ? stream_load<NT>(ci + z0 + i) : 0;
    }
    v[q] = ok ? stream_load<NT>(val + z0 + i) : V(0);
  }
  if constexpr (BITS) {
    // wave 0: the tile's items start at r0 + z0 = b * ITEMS (a whole word);
    // the j-th row end of the tile, at item k of it, closes row r0 + j, whose
    // successor starts at entry z0 + k - j
    static_assert(ITEMS % 32 == 0 && ITEMS / 32 <= 64, "one wave holds the tile's words");
    if (wave == 0) {
      const int64_t w0 = ((int64_t)r0 + z0) >> 5;
      uint32_t w = lane < ITEMS / 32 ? stream_load<NT>(ends + w0 + lane) : 0u;
      const int cnt = __popc(w);
      int pre = cnt;
      for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(pre, d);
        if (lane >= d) pre += t;
      }
      int j = pre - cnt;
      while (w) {
        const int q = __ffs(w) - 1;
        w &= w - 1;
        rows[j + 1] = z0 + lane * 32 + q - j;
        ++j;
      }
      if (lane == 0) rows[0] = stream_load<NT>(rp + r0);
    }
  } else {
    for (int i = tid; i <= last - r0; i += BT) rows[i] = stream_load<NT>(rp + r0 + i);
  }
  V xv[Q];
  if (xbytes > 0 && xbytes <= kMaxBufferBytes) {  // uniform over the workgroup
    const __amdgpu_buffer_rsrc_t xr = make_rsrc(x, xbytes);
#pragma unroll
    for (int q = 0; q < Q; q++) xv[q] = (tid + q * BT < nz) ? buf_at<V>(xr, cidx[q]) : V(0);
  } else {
#pragma unroll
    for (int q = 0; q < Q; q++) xv[q] = (tid + q * BT < nz) ? x[cidx[q]] : V(0);
  }
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int i = tid + q * BT;
    if (i < nz) prod[i] = v[q] * xv[q];
  }
  __syncthreads();
  // rows whose end lies in this tile: short ones by rl threads each (thread
  // q of the group adds entries q, q + rl, ...; xor shuffles combine the
  // group); rl = 4 when the tile's rows average 16 entries or more, else 1
  // (tiles of column panels hold rows of one or two entries). The loop bound
  // is uniform, so every lane takes part in the shuffles.
  const int nrows = r1 - r0;
  const int rl = nz >= 16 * nrows ? 4 : 1;
  for (int base = 0; base < nrows; base += BT / rl) {
    const int i = base + tid / rl, q = tid % rl;
    V sum = 0;
    bool mine = false;
    int s = 0;
    if (i < nrows) {
      s = rows[i];
      const int e = rows[i + 1];
      const int sl = (s > z0 ? s : z0) - z0, el = e - z0;
      if (el - sl > kLongRow) {
        if (q == 0) long_rows[atomicAdd(&nlong, 1)] = i;
      } else {
        for (int k = sl + q; k < el; k += rl) sum += prod[k];
        mine = true;
      }
    }
    for (int off = 1; off < rl; off <<= 1) sum += __shfl_xor(sum, off);
    if (mine && q == 0) {
      const int r = r0 + i;
      if (s >= z0) y[r] = axpby(alpha, sum, beta, y + r);
      else head_val[b] = sum;  // i == 0: row started in an earlier tile
    }
  }
  __syncthreads();
  // long rows: one wavefront each
  for (int q = wave; q < nlong; q += BT / 64) {
    const int i = long_rows[q];
    const int s = rows[i], e = rows[i + 1];
    const int sl = (s > z0 ? s : z0) - z0, el = e - z0;
    V sum = 0;
    for (int k = sl + lane; k < el; k += 64) sum += prod[k];
    sum = group_sum<64>(sum);
    if (lane == 0) {
      const int r = r0 + i;
      if (s >= z0) y[r] = axpby(alpha, sum, beta, y + r);
      else head_val[b] = sum;
    }
  }
  // the row still open at the tile end
  if (r1 < m) {
    const int s = rows[r1 - r0];
    const int sl = (s > z0 ? s : z0) - z0;
    V part = 0;
    for (int k = sl + tid; k < nz; k += BT) part += prod[k];
    part = block_sum<V, BT>(part, red);
    if (tid == 0) {
      carry_row[b] = r1;
      carry_val[b] = part;
    }
  } else if (tid == 0) {
    carry_row[b] = -1;
  }
}

template <class V, bool NT>
__global__ __launch_bounds__(kMergeThreads) void csr_merge_kernel(int64_t m, int64_t n, const int* __restrict__ rp,
                                                           const int* __restrict__ ci, const V* __restrict__ val,
                                                           const V* __restrict__ x, V* __restrict__ y, V alpha,
                                                           V beta, const int* __restrict__ coord,
                                                           int* __restrict__ carry_row, V* __restrict__ carry_val,
                                                           V* __restrict__ head_val) {
  merge_tile<V, NT, kMergeItems, false, false, kMergeThreads>(
      (int)xcd_block(blockIdx.x, gridDim.x), m, rp, ci, val, x, y, alpha, beta, coord, carry_row, carry_val, head_val,
      nullptr, nullptr, nullptr, 0, n * (int64_t)sizeof(V));
}

// Column panels: the matrix split into P column ranges (P a multiple of the 8
// XCDs), each a CSR of its own (panel-local row pointers, entries in row
// order). Workgroups are dealt to the XCDs round-robin, so workgroup b runs on
// XCD b % 8; it takes panel b % 8 + 8 * ((b / 8) / maxT), tile (b / 8) % maxT.
// Each XCD thus gathers only from its own panels' slice of x, one panel after
// the other (the grid is dispatched roughly in order), and the slice stays in
// that XCD's L2. Each panel writes a partial y; panel_reduce adds them in
// panel order.
struct PanelArgs {
  int P;
  int share;  // XCDs per panel when P < 8 (P divides 8), else 1
  int64_t maxT;
  const int64_t* nzoff;  // [P + 1] first entry of each panel
  const int* prp;        // [P][m + 1] panel-local row pointers
  const int* tiles;      // [P] merge tiles per panel
  const int* coord;      // [P][2 (maxT + 1)]
  int* carry_row;        // [P][maxT]
  const uint32_t* ends;  // row-end bits of every panel's merge-path items
  const int64_t* endoff; // [P] first word of each panel's bits
  int64_t width;         // columns per panel
  int64_t n;             // columns of the matrix
  const uint16_t* ci16;  // compact columns (panel-local): low 16 bits
  const uint8_t* cihi;   // and high nibbles, two entries per byte
};

template <class V, bool COMPACT>
__global__ __launch_bounds__(kPanelThreads<V>) void panel_merge_kernel(int64_t m, PanelArgs a, const int* __restrict__ ci,
                                                             const V* __restrict__ val, const V* __restrict__ x,
                                                             V* __restrict__ partial, V* __restrict__ carry_val,
                                                             V* __restrict__ head_val) {
  const int64_t u = blockIdx.x / kXcds;
  const int xcd = (int)(blockIdx.x % kXcds);
  // P >= 8: XCD b % 8 runs panels b % 8, b % 8 + 8, ...; P < 8: `share`
  // XCDs run one panel, taking its tiles in turn
  const int p = a.share > 1 ? xcd / a.share : (int)(xcd + kXcds * (u / a.maxT));
  const int64_t t = a.share > 1 ? u * a.share + xcd % a.share : u % a.maxT;
  if (p >= a.P || t >= a.tiles[p]) return;
  const int64_t off = a.nzoff[p];
  // compact columns are panel-local: x starts at the panel's first column
  // (a panel past the last column has no entries; its base is clamped to n)
  const int64_t first = (int64_t)p * a.width < a.n ? (int64_t)p * a.width : a.n;
  const V* xs = COMPACT ? x + first : x;
  merge_tile<V, true, kPanelItems<V>, true, COMPACT, kPanelThreads<V>>(
      (int)t, m, a.prp + (int64_t)p * (m + 1), COMPACT ? nullptr : ci + off, val + off, xs, partial + (int64_t)p * m,
      V(1), V(0),
      a.coord + (int64_t)p * 2 * (a.maxT + 1), a.carry_row + (int64_t)p * a.maxT, carry_val + (int64_t)p * a.maxT,
      head_val + (int64_t)p * a.maxT, a.ends + a.endoff[p], a.ci16, a.cihi, off,
      (COMPACT ? (a.n - first < a.width ? a.n - first : a.width) : a.n) * (int64_t)sizeof(V));
}

template <class V>
__global__ void panel_fixup_kernel(int64_t m, PanelArgs a, const V* __restrict__ carry_val,
                                   const V* __restrict__ head_val, V* __restrict__ partial) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)a.P * a.maxT) return;
  const int p = (int)(g / a.maxT);
  const int64_t b = g % a.maxT;
  if (b >= a.tiles[p]) return;
  const int* coord = a.coord + (int64_t)p * 2 * (a.maxT + 1);
  const int* rp = a.prp + (int64_t)p * (m + 1);
  const int* crow = a.carry_row + (int64_t)p * a.maxT;
  const V* cval = carry_val + (int64_t)p * a.maxT;
  const int r0 = coord[2 * b], z0 = coord[2 * b + 1], r1 = coord[2 * b + 2];
  if (!(r0 < r1 && rp[r0] < z0)) return;
  V total = head_val[(int64_t)p * a.maxT + b];
  for (int64_t t = b - 1; t >= 0 && crow[t] == r0; --t) total += cval[t];
  partial[(int64_t)p * m + r0] = total;
}

template <class V>
__global__ void panel_reduce(int64_t m, int P, const V* __restrict__ partial, V* __restrict__ y, V alpha, V beta) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  V s = 0;
  for (int p = 0; p < P; p++) s += __builtin_nontemporal_load(partial + (int64_t)p * m + r);
  y[r] = axpby(alpha, s, beta, y + r);
}

// The same sum with 16-byte loads of each partial: for M a multiple of
// 16 / sizeof(V), so that every panel's partial starts 16-byte aligned.
template <class V>
__global__ void panel_reduce_vec(int64_t m, int P, const V* __restrict__ partial, V* __restrict__ y, V alpha,
                                 V beta) {
  constexpr int W = 16 / sizeof(V);
  typedef V vec_t __attribute__((ext_vector_type(W)));
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g * W >= m) return;
  vec_t s = __builtin_nontemporal_load(reinterpret_cast<const vec_t*>(partial) + g);
  for (int p = 1; p < P; p++) s += __builtin_nontemporal_load(reinterpret_cast<const vec_t*>(partial + (int64_t)p * m) + g);
#pragma unroll
  for (int u = 0; u < W; u++) y[g * W + u] = axpby(alpha, s[u], beta, y + g * W + u);
}

// row-end bits of the panels' merge paths: row i of panel p ends at item
// i + rp_p[i + 1]. Consecutive threads take consecutive rows, so the bits of
// one word come from a run of neighbouring lanes: the run is OR-combined with
// shuffles and its first lane sets the word with one atomic.
__global__ __launch_bounds__(kBlock) void panel_ends(int64_t m, int P, const int* __restrict__ prp,
                                                     const int64_t* __restrict__ endoff, uint32_t* __restrict__ ends) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int64_t word = -1;  // global word index; runs of equal words are contiguous along the lanes
  uint32_t bits = 0;
  if (g < (int64_t)P * m) {
    const int p = (int)(g / m);
    const int64_t i = g - (int64_t)p * m;
    const int64_t k = i + prp[(int64_t)p * (m + 1) + i + 1];
    word = endoff[p] + (k >> 5);
    bits = 1u << (k & 31);
  }
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t b = __shfl_down(bits, off);
    const long long w = __shfl_down((long long)word, off);
    if (lane + off < 64 && w == word) bits |= b;
  }
  const long long prev = __shfl_up((long long)word, 1);
  if (word >= 0 && (lane == 0 || prev != word)) atomicOr(ends + word, bits);
}

// panel-local columns of entries 2t and 2t + 1: low 16 bits, and the high
// nibbles packed into byte t (entry 2t in the low nibble)
__global__ void panel_compact(int64_t nnz, int P, int64_t width, const int64_t* __restrict__ nzoff,
                              const int* __restrict__ pci, uint16_t* __restrict__ lo, uint8_t* __restrict__ hi) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (2 * t >= nnz) return;
  uint32_t byte = 0;
  for (int u = 0; u < 2; u++) {
    const int64_t e = 2 * t + u;
    if (e >= nnz) break;
    int p = 0;
    while (p + 1 < P && nzoff[p + 1] <= e) p++;
    const int64_t local = pci[e] - (int64_t)p * width;
    lo[e] = (uint16_t)(local & 0xFFFF);
    byte |= (uint32_t)((local >> 16) & 15) << (4 * u);
  }
  hi[t] = (uint8_t)byte;
}

// Panel build: a wave takes 64 consecutive rows, whose entries are
// contiguous in the CSR arrays, and walks them 64 at a time (coalesced).
constexpr int kMaxPanels = 16;

// Row (0..63) of the wave's rows holding entry `idx`: the number of rows whose
// end is <= idx; lane l holds the end of row l (INT32_MAX past the last row).
__device__ __forceinline__ int row_in_wave(int rend, int idx) {
  int lo = 0;
  for (int step = 32; step > 0; step >>= 1)
    if (__shfl(rend, lo + step - 1) <= idx) lo += step;
  return lo;
}

// per (panel, row) entry counts: LDS counters for the wave's 64 rows x P
// panels, written out panel by panel
__global__ __launch_bounds__(kBlock) void panel_count(int64_t m, int P, int64_t width, const int* __restrict__ rp,
                                                      const int* __restrict__ ci, int64_t* __restrict__ cnt) {
  __shared__ int counts[kBlock / 64][64 * kMaxPanels];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int* c = counts[w];
  for (int i = lane; i < 64 * kMaxPanels; i += 64) c[i] = 0;
  __syncthreads();
  const int64_t r0 = ((int64_t)blockIdx.x * (kBlock / 64) + w) * 64;
  const int nr = r0 < m ? (int)(m - r0 < 64 ? m - r0 : 64) : 0;
  if (nr > 0) {
    const int rend = lane < nr ? rp[r0 + lane + 1] : INT32_MAX;
    const int s = rp[r0], e = rp[r0 + nr];
    for (int e0 = s; e0 < e; e0 += 64) {
      const int idx = e0 + lane;
      const int row = row_in_wave(rend, idx);
      if (idx < e) atomicAdd(&c[row * P + (int)(ci[idx] / width)], 1);
    }
  }
  __syncthreads();
  if (lane < nr)
    for (int q = 0; q < P; q++) cnt[(int64_t)q * (m + 1) + r0 + lane] = c[lane * P + q];
}

__global__ void panel_rp(int64_t n, const int64_t* __restrict__ in, int* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int)in[i];
}

// Entries to their panels, order kept: panel q's entries of the wave's rows go
// to consecutive slots from nzoff[q] + prp_q[r0] (held by lane q); an entry's
// slot adds the number of earlier lanes of its step bound for the same panel.
template <class V>
__global__ __launch_bounds__(kBlock) void panel_scatter(int64_t m, int P, int64_t width, const int* __restrict__ rp,
                                                        const int* __restrict__ ci, const V* __restrict__ val,
                                                        const int64_t* __restrict__ nzoff,
                                                        const int* __restrict__ prp, int* __restrict__ pci,
                                                        V* __restrict__ pval) {
  const int lane = threadIdx.x & 63;
  const int64_t r0 = ((int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * 64;
  if (r0 >= m) return;  // whole waves leave together
  const int nr = (int)(m - r0 < 64 ? m - r0 : 64);
  long long next = lane < P ? nzoff[lane] + prp[(int64_t)lane * (m + 1) + r0] : 0;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int s = rp[r0], e = rp[r0 + nr];
  for (int e0 = s; e0 < e; e0 += 64) {
    const int idx = e0 + lane;
    const bool ok = idx < e;
    const int c = ok ? ci[idx] : 0;
    const int p = ok ? (int)(c / width) : -1;
    long long dst = 0;
    int taken = 0;
    for (int q = 0; q < P; q++) {
      const uint64_t mk = __ballot(p == q);
      const long long base = __shfl(next, q);
      if (p == q) dst = base + __popcll(mk & below);
      if (lane == q) taken = __popcll(mk);
    }
    next += taken;
    if (ok) {
      pci[dst] = c;
      pval[dst] = val[idx];
    }
  }
}

// A row that started in an earlier tile and ends in tile b: add the partials of
// the preceding tiles that left it open, in tile order from the nearest.
template <class V>
__global__ void csr_merge_fixup_kernel(int64_t tiles, const int* __restrict__ coord, const int* __restrict__ rp,
                                       const int* __restrict__ carry_row, const V* __restrict__ carry_val,
                                       const V* __restrict__ head_val, V* __restrict__ y, V alpha, V beta) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= tiles) return;
  const int r0 = coord[2 * b], z0 = coord[2 * b + 1], r1 = coord[2 * b + 2];
  if (!(r0 < r1 && rp[r0] < z0)) return;
  V total = head_val[b];
  for (int64_t t = b - 1; t >= 0 && carry_row[t] == r0; --t) total += carry_val[t];
  y[r0] = axpby(alpha, total, beta, y + r0);
}

// SpMM: Y(M x k) = alpha*A*X(N x k) + beta*Y, row-major X/Y. A group of KP
// lanes (KP = k rounded up to a power of two) owns one row; lane c owns
// column c, so each gathered X row is one contiguous load of k values.
template <class V, int KP>
__global__ __launch_bounds__(kBlock) void csr_spmm_kernel(int64_t m, int k, const int* __restrict__ rp,
                                                          const int* __restrict__ ci, const V* __restrict__ val,
                                                          const V* __restrict__ X, int64_t ldx,
                                                          V* __restrict__ Y, int64_t ldy, V alpha, V beta) {
  const int c = threadIdx.x & (KP - 1);
  const int64_t row = (int64_t)blockIdx.x * (kBlock / KP) + threadIdx.x / KP;
  if (row >= m || c >= k) return;
  const int s = rp[row], e = rp[row + 1];
  V sum = 0;
  int j = s;
  for (; j + 1 < e; j += 2) {
    const int c0 = ci[j], c1 = ci[j + 1];
    const V v0 = val[j], v1 = val[j + 1];
    sum += v0 * X[(int64_t)c0 * ldx + c];
    sum += v1 * X[(int64_t)c1 * ldx + c];
  }
  if (j < e) sum += val[j] * X[(int64_t)ci[j] * ldx + c];
  V* yp = Y + row * ldy + c;
  *yp = axpby(alpha, sum, beta, yp);
}

template <class V, int L>
void launch_vector(Matrix& A, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  const int64_t rows_per_block = kBlock / L;
  hipLaunchKernelGGL((csr_vector_kernel<V, L>), grid1d(A.m, (int)rows_per_block), dim3(kBlock), 0, s, A.m,
                     A.rp.as<int>(), A.ci.as<int>(), A.val.as<V>(), static_cast<const V*>(x), static_cast<V*>(y),
                     (V)alpha, (V)beta);
  LSP_LAUNCHED();
}

template <class V>
void vector_dispatch(Matrix& A, int lanes, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  switch (lanes) {
    case 2: launch_vector<V, 2>(A, alpha, x, beta, y, s); break;
    case 4: launch_vector<V, 4>(A, alpha, x, beta, y, s); break;
    case 8: launch_vector<V, 8>(A, alpha, x, beta, y, s); break;
    case 16: launch_vector<V, 16>(A, alpha, x, beta, y, s); break;
    case 32: launch_vector<V, 32>(A, alpha, x, beta, y, s); break;
    default: launch_vector<V, 64>(A, alpha, x, beta, y, s); break;
  }
}

// The merge-path tiles and their carry fix-up over a partition `coord` of
// `tiles` tiles; carry_row / carry_val / head_val hold one entry per tile.
template <class V>
void merge_launch(int64_t m, int64_t n, int64_t tiles, const int* rp, const int* ci, const V* val, const void* x,
                  void* y,
                  double alpha, double beta, const int* coord, int* carry_row, V* carry_val, V* head_val,
                  hipStream_t s, bool fixup = true) {
  if (tiles == 0) return;
  // a matrix beyond the 256 MB last-level cache is streamed with
  // non-temporal loads, so its arrays do not push x out of the caches
  // (27-point stencil 160^3 FP64 284 -> 278 us); one that fits keeps normal
  // loads, which let repeated calls find it in that cache (matrices of 10k to
  // 2M rows: 5 % slower with non-temporal loads)
  const int64_t nnz = tiles * (int64_t)kMergeItems - m;  // upper bound (the last tile may be short)
  const bool nt = (double)nnz * (double)(sizeof(V) + 4) + 4.0 * (double)m > 256.0 * (1 << 20);
  hipLaunchKernelGGL((nt ? csr_merge_kernel<V, true> : csr_merge_kernel<V, false>), dim3((unsigned)tiles),
                     dim3(kMergeThreads), 0, s, m, n, rp, ci, val,
                     static_cast<const V*>(x), static_cast<V*>(y), (V)alpha, (V)beta, coord, carry_row, carry_val,
                     head_val);
  LSP_LAUNCHED();
  if (!fixup) return;  // tiles cut at row starts: no row continues into the next tile
  hipLaunchKernelGGL((csr_merge_fixup_kernel<V>), grid1d(tiles, kBlock), dim3(kBlock), 0, s, tiles, coord, rp,
                     carry_row, carry_val, head_val, static_cast<V*>(y), (V)alpha, (V)beta);
  LSP_LAUNCHED();
}

template <class V>
void merge_run(Matrix& A, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  CsrMeta& mt = A.meta;
  merge_launch<V>(A.m, A.n, mt.merge_tiles, A.rp.as<int>(), A.ci.as<int>(), A.val.as<V>(), x, y, alpha, beta,
                  mt.merge_coord.as<int>(), mt.carry_row.as<int>(), mt.carry_val.as<V>(), mt.head_val.as<V>(), s);
}

// Splits the matrix into P column panels (row order kept inside each panel).
template <class V>
void build_panels(Matrix& A, int P, hipStream_t s) {
  CsrMeta& mt = A.meta;
  const int64_t m = A.m, width = ceil_div(A.n, P);
  DBuf cnt;
  cnt.alloc(sizeof(int64_t) * P * (m + 1));
  LSP_HIP(hipMemsetAsync(cnt.p, 0, cnt.bytes, s));
  LSP_REQUIRE(P <= kMaxPanels, LESPMV_ERR_UNSUPPORTED, "at most 16 column panels");
  hipLaunchKernelGGL(panel_count, grid1d(m, kBlock), dim3(kBlock), 0, s, m, P, width, A.rp.as<int>(), A.ci.as<int>(),
                     cnt.as<int64_t>());
  LSP_LAUNCHED();
  std::vector<int64_t> nzoff(P + 1, 0);
  for (int p = 0; p < P; p++) nzoff[p + 1] = nzoff[p] + scan_exclusive_i64(cnt.as<int64_t>() + p * (m + 1), m + 1, s);
  mt.p_rp.alloc(sizeof(int) * P * (m + 1));
  hipLaunchKernelGGL(panel_rp, grid1d(P * (m + 1), kBlock), dim3(kBlock), 0, s, (int64_t)P * (m + 1),
                     cnt.as<int64_t>(), mt.p_rp.as<int>());
  LSP_LAUNCHED();
  mt.p_nzoff.alloc(sizeof(int64_t) * (P + 1));
  LSP_HIP(hipMemcpyAsync(mt.p_nzoff.p, nzoff.data(), sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, s));
  mt.p_ci.alloc(sizeof(int) * std::max<int64_t>(A.nnz, 1));
  mt.p_val.alloc(sizeof(V) * std::max<int64_t>(A.nnz, 1));
  hipLaunchKernelGGL((panel_scatter<V>), grid1d(m, kBlock), dim3(kBlock), 0, s, m, P, width, A.rp.as<int>(),
                     A.ci.as<int>(), A.val.as<V>(), mt.p_nzoff.as<int64_t>(), mt.p_rp.as<int>(), mt.p_ci.as<int>(),
                     mt.p_val.as<V>());
  LSP_LAUNCHED();
  std::vector<int> tiles(P);
  int64_t maxT = 1;
  for (int p = 0; p < P; p++) {
    tiles[p] = (int)ceil_div(m + nzoff[p + 1] - nzoff[p], kPanelItems<V>);
    maxT = std::max<int64_t>(maxT, tiles[p]);
  }
  mt.p_maxT = maxT;
  mt.p_tiles.alloc(sizeof(int) * P);
  LSP_HIP(hipMemcpyAsync(mt.p_tiles.p, tiles.data(), sizeof(int) * P, hipMemcpyHostToDevice, s));
  mt.p_coord.alloc(sizeof(int) * P * 2 * (maxT + 1));
  for (int p = 0; p < P; p++)
    if (tiles[p]) {
      hipLaunchKernelGGL(merge_partition_kernel, grid1d(tiles[p] + 1, kBlock), dim3(kBlock), 0, s, (int64_t)tiles[p],
                         m, nzoff[p + 1] - nzoff[p], mt.p_rp.as<int>() + (int64_t)p * (m + 1),
                         mt.p_coord.as<int>() + (int64_t)p * 2 * (maxT + 1), (int64_t)kPanelItems<V>);
      LSP_LAUNCHED();
    }
  // row-end bits: tiles[p] * ITEMS / 32 words per panel, so every tile reads
  // whole words inside its panel's run
  std::vector<int64_t> endoff(P + 1, 0);
  for (int p = 0; p < P; p++) endoff[p + 1] = endoff[p] + (int64_t)tiles[p] * (kPanelItems<V> / 32);
  mt.p_ends.alloc(sizeof(uint32_t) * std::max<int64_t>(endoff[P], 1));
  LSP_HIP(hipMemsetAsync(mt.p_ends.p, 0, mt.p_ends.bytes, s));
  mt.p_endoff.alloc(sizeof(int64_t) * (P + 1));
  LSP_HIP(hipMemcpyAsync(mt.p_endoff.p, endoff.data(), sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, s));
  if (m) {
    hipLaunchKernelGGL(panel_ends, grid1d((int64_t)P * m, kBlock), dim3(kBlock), 0, s, m, P, mt.p_rp.as<int>(),
                       mt.p_endoff.as<int64_t>(), mt.p_ends.as<uint32_t>());
    LSP_LAUNCHED();
  }
  // FP64 panels of at most 2^20 columns keep their (panel-local) columns in
  // 2.5 bytes per entry (lognormal matrix, 8 panels: 532 against 542 us). In
  // FP32 the extra 16-bit and nibble loads cost more than the 1.5 bytes per
  // entry they save (398 against 393 us), so FP32 keeps 32-bit columns.
  mt.p_width = width;
  mt.p_compact = sizeof(V) == 8 && width <= (int64_t{1} << 20);
  mt.p_ci16.reset();
  mt.p_cihi.reset();
  if (mt.p_compact) {
    mt.p_ci16.alloc(sizeof(uint16_t) * std::max<int64_t>(A.nnz, 1));
    mt.p_cihi.alloc(std::max<int64_t>((A.nnz + 1) / 2, 1));
    if (A.nnz) {
      hipLaunchKernelGGL(panel_compact, grid1d((A.nnz + 1) / 2, kBlock), dim3(kBlock), 0, s, A.nnz, P, width,
                         mt.p_nzoff.as<int64_t>(), mt.p_ci.as<int>(), mt.p_ci16.as<uint16_t>(),
                         mt.p_cihi.as<uint8_t>());
      LSP_LAUNCHED();
    }
  }
  mt.p_carry_row.alloc(sizeof(int) * P * maxT);
  mt.p_carry_val.alloc(sizeof(V) * P * maxT);
  mt.p_head.alloc(sizeof(V) * P * maxT);
  mt.p_partial.alloc(sizeof(V) * P * m);
  LSP_HIP(hipStreamSynchronize(s));
  if (mt.p_compact) mt.p_ci.reset();
  mt.panels = P;
}

template <class V>
void panel_run(Matrix& A, int P, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  CsrMeta& mt = A.meta;
  if (mt.panels != P) build_panels<V>(A, P, s);  // first call without csr_prepare
  const int share = P < kXcds ? kXcds / P : 1;
  PanelArgs a{P,
              share,
              mt.p_maxT,
              mt.p_nzoff.as<int64_t>(),
              mt.p_rp.as<int>(),
              mt.p_tiles.as<int>(),
              mt.p_coord.as<int>(),
              mt.p_carry_row.as<int>(),
              mt.p_ends.as<uint32_t>(),
              mt.p_endoff.as<int64_t>(),
              mt.p_width,
              A.n,
              mt.p_ci16.as<uint16_t>(),
              mt.p_cihi.as<uint8_t>()};
  const int64_t blocks = share > 1 ? kXcds * ceil_div(mt.p_maxT, share) : P * mt.p_maxT;
  auto kern = mt.p_compact ? panel_merge_kernel<V, true> : panel_merge_kernel<V, false>;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kPanelThreads<V>), 0, s, A.m, a, mt.p_ci.as<int>(),
                     mt.p_val.as<V>(),
                     static_cast<const V*>(x), mt.p_partial.as<V>(), mt.p_carry_val.as<V>(), mt.p_head.as<V>());
  LSP_LAUNCHED();
  hipLaunchKernelGGL((panel_fixup_kernel<V>), grid1d(P * mt.p_maxT, kBlock), dim3(kBlock), 0, s, A.m, a,
                     mt.p_carry_val.as<V>(), mt.p_head.as<V>(), mt.p_partial.as<V>());
  LSP_LAUNCHED();
  constexpr int W = 16 / sizeof(V);
  if (A.m % W == 0)
    hipLaunchKernelGGL((panel_reduce_vec<V>), grid1d(A.m / W, kBlock), dim3(kBlock), 0, s, A.m, P,
                       mt.p_partial.as<V>(), static_cast<V*>(y), (V)alpha, (V)beta);
  else
    hipLaunchKernelGGL((panel_reduce<V>), grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, P, mt.p_partial.as<V>(),
                       static_cast<V*>(y), (V)alpha, (V)beta);
  LSP_LAUNCHED();
}

template <class V, int KP>
void launch_spmm(Matrix& A, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y, int64_t ldy,
                 hipStream_t s) {
  hipLaunchKernelGGL((csr_spmm_kernel<V, KP>), grid1d(A.m, kBlock / KP), dim3(kBlock), 0, s, A.m, k,
                     A.rp.as<int>(), A.ci.as<int>(), A.val.as<V>(), static_cast<const V*>(X), ldx,
                     static_cast<V*>(Y), ldy, (V)alpha, (V)beta);
  LSP_LAUNCHED();
}

// SpMM through the row engine: one work item (rpi rows) per wave. A launch
// of its own wants more, smaller items than a packed batch (which already
// fills the GPU with many requests).
constexpr int kCallIters = 8;
struct RowsArgs {
  const int* rp;
  const int* ci;
  const void* val;
  const void* X;
  void* Y;
  const int* order;  // none: rows in order
  int64_t ldx, ldy;
  double alpha, beta;
  int32_t m, k, lanes, rpi;
};

template <class V, int W, bool VEC, bool BUF>
__global__ __launch_bounds__(kBlock) void csr_rows_kernel(RowsArgs a, int64_t nitems) {
  const int64_t item = xcd_block(blockIdx.x, gridDim.x) * (kBlock / 64) + threadIdx.x / 64;
  if (item >= nitems) return;  // whole waves leave together
  const int r0 = (int)(item * a.rpi);
  const int r1 = r0 + a.rpi < a.m ? r0 + a.rpi : a.m;
  run_rows<V, W, VEC, BUF>(a, r0, r1, threadIdx.x & 63);
}

template <class V>
void rows_dispatch(Matrix& A, int lanes, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
                   int64_t ldy, hipStream_t s) {
  const int mode = rows_mode(k, ldx, ldy, X, Y, (int64_t)sizeof(V));
  const int rpi = rows_item(A.meta.avg_row, lanes, k, kModeW[mode], kCallIters);
  const int64_t nitems = ceil_div(A.m, rpi);
  const RowsArgs a{A.rp.as<int>(), A.ci.as<int>(), A.val.p, X, Y,
                   A.row_order.p ? A.row_order.as<int>() : nullptr, ldx, ldy, alpha, beta, (int32_t)A.m, k, lanes,
                   rpi};
  const dim3 grid = grid1d(nitems, kBlock / 64);
  const bool buf = rows_buffer_ok(A.nnz, A.n, ldx, (int64_t)sizeof(V));
  void (*kern)(RowsArgs, int64_t);
  switch (mode) {
    case 0: kern = buf ? csr_rows_kernel<V, 1, false, true> : csr_rows_kernel<V, 1, false, false>; break;
    case 1: kern = buf ? csr_rows_kernel<V, 2, false, true> : csr_rows_kernel<V, 2, false, false>; break;
    case 2: kern = buf ? csr_rows_kernel<V, 2, true, true> : csr_rows_kernel<V, 2, true, false>; break;
    case 3: kern = buf ? csr_rows_kernel<V, 4, false, true> : csr_rows_kernel<V, 4, false, false>; break;
    default: kern = buf ? csr_rows_kernel<V, 4, true, true> : csr_rows_kernel<V, 4, true, false>; break;
  }
  hipLaunchKernelGGL(kern, grid, dim3(kBlock), 0, s, a, nitems);
  LSP_LAUNCHED();
}

template <class V>
void spmm_dispatch(Matrix& A, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y, int64_t ldy,
                   hipStream_t s) {
  if (k <= 1) launch_spmm<V, 1>(A, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 2) launch_spmm<V, 2>(A, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 4) launch_spmm<V, 4>(A, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 8) launch_spmm<V, 8>(A, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 16) launch_spmm<V, 16>(A, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 32) launch_spmm<V, 32>(A, k, alpha, X, ldx, beta, Y, ldy, s);
  else launch_spmm<V, 64>(A, k, alpha, X, ldx, beta, Y, ldy, s);
}

// Entries whose column lies more than `window` columns from the row's
// diagonal position r * n / m (counted once, for the kernel choice).
__global__ __launch_bounds__(kBlock) void far_count_kernel(int64_t m, int64_t n, int64_t window, int line_shift,
                                                           const int* __restrict__ rp, const int* __restrict__ ci,
                                                           unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[kBlock / 64];
  unsigned long long c = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < m; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = r * n / m;
    int prev = -1;  // x line of the row's previous entry
    for (int j = rp[r]; j < rp[r + 1]; j++) {
      const int col = ci[j];
      const int64_t dc = col - d;
      const int line = col >> line_shift;
      c += ((dc > window || dc < -window) && line != prev) ? 1ull : 0ull;
      prev = line;
    }
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < kBlock / 64; i++) t += red[i];
    if (t) atomicAdd(out, t);
  }
}

int panels_of(int variant) {
  return variant == LESPMV_VARIANT_PANEL ? 8 : variant == LESPMV_VARIANT_PANEL16 ? 16
                                                : variant == LESPMV_VARIANT_PANEL4 ? 4 : 0;
}

// Lanes per row of the vector kernel: each lane keeps four loads in flight,
// so uniform rows want about avg/4 lanes; long-tailed rows want about avg
// lanes so a wave's rows finish together.
int lanes_for(double avg, int64_t max_row) {
  const bool uniform = (double)max_row <= 2.0 * avg + 8.0;
  const double want = uniform ? avg / 4.0 : avg;
  int L = 2;
  while (L < 64 && L < want) L *= 2;
  return L;
}

}  // namespace

void csr_merge_partition(int64_t tiles, int64_t m, int64_t nnz, const int* rp, int* coord, hipStream_t s) {
  hipLaunchKernelGGL(merge_partition_kernel, grid1d(tiles + 1, kBlock), dim3(kBlock), 0, s, tiles, m, nnz, rp, coord,
                     (int64_t)kMergeItems);
  LSP_LAUNCHED();
}

void csr_merge_tiles(int dtype, int64_t m, int64_t n, int64_t tiles, const int* rp, const int* ci, const void* val,
                     const void* x, void* y, double alpha, double beta, const int* coord, int* carry_row,
                     void* carry_val, void* head_val, hipStream_t s) {
  if (dtype == LESPMV_F64)
    merge_launch<double>(m, n, tiles, rp, ci, static_cast<const double*>(val), x, y, alpha, beta, coord, carry_row,
                         static_cast<double*>(carry_val), static_cast<double*>(head_val), s);
  else
    merge_launch<float>(m, n, tiles, rp, ci, static_cast<const float*>(val), x, y, alpha, beta, coord, carry_row,
                        static_cast<float*>(carry_val), static_cast<float*>(head_val), s);
}

void csr_analyse(Matrix& A, hipStream_t s) {
  CsrMeta& mt = A.meta;
  mt.max_row = A.m ? reduce_max_rowlen(A.rp.as<int32_t>(), A.m, s) : 0;
  mt.avg_row = A.m ? (double)A.nnz / (double)A.m : 0.0;
  mt.lanes = lanes_for(mt.avg_row, mt.max_row);
  // far x lines per entry: entries more than 2 MB of x away from the
  // diagonal that open a new 128-byte line of x in their row (columns in
  // runs, as in small dense blocks, share a line and count once)
  mt.far_frac = 0.0;
  if (A.m && A.nnz) {
    DBuf cnt;
    cnt.alloc(sizeof(unsigned long long));
    LSP_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), s));
    const int64_t window = (int64_t{2} << 20) / (int64_t)vsize(A.dtype);
    hipLaunchKernelGGL(far_count_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(A.m, kBlock), 4096)), dim3(kBlock),
                       0, s, A.m, A.n, window, vsize(A.dtype) == 8 ? 4 : 5, A.rp.as<int>(), A.ci.as<int>(),
                       cnt.as<unsigned long long>());
    LSP_LAUNCHED();
    unsigned long long far = 0;
    LSP_HIP(hipMemcpyAsync(&far, cnt.p, sizeof(far), hipMemcpyDeviceToHost, s));
    LSP_HIP(hipStreamSynchronize(s));
    mt.far_frac = (double)far / (double)A.nnz;
  }
  // Column panels when x is at least four times an XCD's 4 MB L2 and there
  // are at least 0.3 far x lines per entry: every XCD then keeps its own slice
  // of x (lognormal matrix of 4M columns, 70 % uniform columns: FP64 495 us
  // against 830 us for the vector kernel, FP32 378 against 697; random 4x4
  // blocks, a quarter line per entry, run faster without). Otherwise the
  // merge path, which bounds every workgroup's work, when the longest row is
  // far beyond the mean, and also for a matrix that does not fit the 256 MB
  // last-level cache whose rows are short or long-tailed: read from HBM,
  // the merge tiles' staged loads win there (27-point stencils of 37M-109M
  // entries 0.88-0.90x the vector kernel's time, banded and block-diagonal
  // rows of 13-30 entries 0.76-0.98x; rows of 50 entries and more 1.1-1.3x),
  // while on cache-resident matrices the vector kernel is faster (2-D
  // Laplacians of 7M-14M entries 1.12-1.23x for the merge path, and the
  // selector's synthetic families).
  const double sv = (double)vsize(A.dtype);
  const double xbytes = (double)A.n * sv;
  const double bytes = (double)A.nnz * (4.0 + sv) + 4.0 * (double)(A.m + 1) + sv * (double)(A.n + A.m);
  if (mt.far_frac >= 0.3 && xbytes >= 16.0 * (1 << 20))
    mt.auto_variant = LESPMV_VARIANT_PANEL;
  else if (mt.max_row > 32 * std::max<double>(mt.avg_row, 4.0))
    mt.auto_variant = LESPMV_VARIANT_MERGE;
  else if (bytes > 256.0 * (1 << 20) && (mt.avg_row <= 32.0 || mt.max_row > 8.0 * mt.avg_row))
    mt.auto_variant = LESPMV_VARIANT_MERGE;
  else
    mt.auto_variant = LESPMV_VARIANT_VECTOR;
  // merge tiles cut at row starts when every row fits a quarter tile: the
  // same tile kernel without the carry fix-up pass, never slower than the
  // merge path where both apply (27-point stencil 160^3 FP64 278 against
  // 280 us, banded rows of 17 on 2M rows 72 against 74, power-law rows of up
  // to 400 entries 28.5 against 33.4)
  if (mt.auto_variant == LESPMV_VARIANT_MERGE && mt.max_row + 1 <= kMergeItems / 4)
    mt.auto_variant = LESPMV_VARIANT_ROW_TILES;
  const int64_t items = A.m + A.nnz;
  mt.merge_tiles = ceil_div(items, kMergeItems);
  mt.merge_coord.alloc(sizeof(int) * 2 * (mt.merge_tiles + 1));
  mt.carry_row.alloc(sizeof(int) * std::max<int64_t>(mt.merge_tiles, 1));
  mt.carry_val.alloc(vsize(A.dtype) * std::max<int64_t>(mt.merge_tiles, 1));
  mt.head_val.alloc(vsize(A.dtype) * std::max<int64_t>(mt.merge_tiles, 1));
  hipLaunchKernelGGL(merge_partition_kernel, grid1d(mt.merge_tiles + 1, kBlock), dim3(kBlock), 0, s, mt.merge_tiles,
                     A.m, A.nnz, A.rp.as<int>(), mt.merge_coord.as<int>(), (int64_t)kMergeItems);
  LSP_LAUNCHED();
  // the partition is complete before any call (on any stream) reads it
  LSP_HIP(hipStreamSynchronize(s));
  mt.analysed = true;
}

bool row_tiles_fit(const Matrix& A);
void row_tiles_prepare(Matrix& A, hipStream_t s);

void csr_prepare(Matrix& A, int variant, hipStream_t s) {
  if (!A.meta.analysed) csr_analyse(A, s);
  if (variant == LESPMV_VARIANT_AUTO) variant = A.meta.auto_variant;
  if (variant == LESPMV_VARIANT_ROW_TILES && A.m && row_tiles_fit(A) && !A.meta.row_coord.p) row_tiles_prepare(A, s);
  const int P = panels_of(variant);
  if (P && A.m && A.meta.panels != P) {
    if (A.dtype == LESPMV_F64) build_panels<double>(A, P, s);
    else build_panels<float>(A, P, s);
  }
}

void csr_release_aux(Matrix& A) {
  CsrMeta& mt = A.meta;
  for (DBuf* b : {&mt.p_nzoff, &mt.p_rp, &mt.p_ci, &mt.p_val, &mt.p_tiles, &mt.p_coord, &mt.p_carry_row,
                  &mt.p_carry_val, &mt.p_head, &mt.p_partial, &mt.p_ends, &mt.p_endoff, &mt.p_ci16, &mt.p_cihi})
    b->reset();
  mt.panels = 0;
}

int64_t csr_aux_bytes(const Matrix& A) {
  const CsrMeta& mt = A.meta;
  int64_t b = 0;
  for (const DBuf* d : {&mt.merge_coord, &mt.carry_row, &mt.carry_val, &mt.head_val, &mt.row_coord,
                        &mt.row_carry_row, &mt.row_carry_val, &mt.row_head, &mt.p_nzoff, &mt.p_rp,
                        &mt.p_ci, &mt.p_val, &mt.p_tiles, &mt.p_coord, &mt.p_carry_row, &mt.p_carry_val, &mt.p_head,
                        &mt.p_partial, &mt.p_ends, &mt.p_endoff, &mt.p_ci16, &mt.p_cihi})
    b += (int64_t)d->bytes;
  return b;
}

// The row-aligned partition of variant ROW_TILES (rows of at most
// kMergeItems / 4 - 1 entries): tiles of kMergeItems - (longest row + 1)
// nominal merge-path items, cut back to row starts.
bool row_tiles_fit(const Matrix& A) { return A.meta.max_row + 1 <= kMergeItems / 4; }

void row_tiles_prepare(Matrix& A, hipStream_t s) {
  CsrMeta& mt = A.meta;
  const int64_t items = kMergeItems - (mt.max_row + 1);
  mt.row_tiles = ceil_div(A.m + A.nnz, items);
  mt.row_coord.alloc(sizeof(int) * 2 * (mt.row_tiles + 1));
  mt.row_carry_row.alloc(sizeof(int) * std::max<int64_t>(mt.row_tiles, 1));
  mt.row_carry_val.alloc(vsize(A.dtype) * std::max<int64_t>(mt.row_tiles, 1));
  mt.row_head.alloc(vsize(A.dtype) * std::max<int64_t>(mt.row_tiles, 1));
  hipLaunchKernelGGL(row_partition_kernel, grid1d(mt.row_tiles + 1, kBlock), dim3(kBlock), 0, s, mt.row_tiles, A.m,
                     A.nnz, A.rp.as<int>(), mt.row_coord.as<int>(), items);
  LSP_LAUNCHED();
  LSP_HIP(hipStreamSynchronize(s));  // complete before any call (on any stream) reads it
}

template <class V>
void row_tiles_run(Matrix& A, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  CsrMeta& mt = A.meta;
  if (!mt.row_coord.p) row_tiles_prepare(A, s);
  merge_launch<V>(A.m, A.n, mt.row_tiles, A.rp.as<int>(), A.ci.as<int>(), A.val.as<V>(), x, y, alpha, beta,
                  mt.row_coord.as<int>(), mt.row_carry_row.as<int>(), mt.row_carry_val.as<V>(), mt.row_head.as<V>(),
                  s, false);
}

void csr_spmv(Matrix& A, int variant, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  if (A.m == 0) return;
  if (!A.meta.analysed) csr_analyse(A, s);
  if (variant == LESPMV_VARIANT_AUTO) variant = A.meta.auto_variant;
  if (variant == LESPMV_VARIANT_ROW_TILES) {
    if (!row_tiles_fit(A)) variant = LESPMV_VARIANT_MERGE;
    else if (A.dtype == LESPMV_F64) return row_tiles_run<double>(A, alpha, x, beta, y, s);
    else return row_tiles_run<float>(A, alpha, x, beta, y, s);
  }
  if (variant == LESPMV_VARIANT_MERGE) {
    if (A.dtype == LESPMV_F64) merge_run<double>(A, alpha, x, beta, y, s);
    else merge_run<float>(A, alpha, x, beta, y, s);
    return;
  }
  if (const int P = panels_of(variant)) {
    if (A.dtype == LESPMV_F64) panel_run<double>(A, P, alpha, x, beta, y, s);
    else panel_run<float>(A, P, alpha, x, beta, y, s);
    return;
  }
  // vector kernel; variants >= 16 request an explicit lane count (16 + log2 L)
  int lanes = A.meta.lanes;
  if (variant >= 16 && variant <= 22) lanes = 1 << (variant - 16);
  if (lanes < 2) lanes = 2;
  if (A.dtype == LESPMV_F64) vector_dispatch<double>(A, lanes, alpha, x, beta, y, s);
  else vector_dispatch<float>(A, lanes, alpha, x, beta, y, s);
}

void csr_spmm(Matrix& A, int variant, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
              int64_t ldy, hipStream_t s) {
  if (A.m == 0 || k == 0) return;
  if (variant >= 16 && variant <= 22) {  // the row engine with 2^(variant - 16) lanes per row
    LSP_REQUIRE(A.m <= INT32_MAX, LESPMV_ERR_UNSUPPORTED, "the row engine takes M < 2^31");
    if (!A.meta.analysed) csr_analyse(A, s);
    const int lanes = 1 << (variant - 16);
    if (A.dtype == LESPMV_F64) rows_dispatch<double>(A, lanes, k, alpha, X, ldx, beta, Y, ldy, s);
    else rows_dispatch<float>(A, lanes, k, alpha, X, ldx, beta, Y, ldy, s);
    return;
  }
  if (A.dtype == LESPMV_F64) spmm_dispatch<double>(A, k, alpha, X, ldx, beta, Y, ldy, s);
  else spmm_dispatch<float>(A, k, alpha, X, ldx, beta, Y, ldy, s);
}

}  // namespace lsp
