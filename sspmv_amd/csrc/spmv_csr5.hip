// CSR5 on gfx950.
//
// Reference: csr_to_csr5 (include/sparse_conversion.h:938-1379, transpose in
// include/csr5_utils.h:10-87) and LeSpMV_csr5 (src/spmv_csr5.cpp:112-457).
// The conversion runs as one kernel per reference step and produces the
// reference's arrays bit for bit for any omega/sigma whose descriptor fits one
// 32-bit packet (reference: omega = 64 B / sizeof(V), sigma = 16; GPU-native:
// omega = 64 = one wavefront per tile).
//
// SpMV: one tile per group of omega lanes; lane l owns the sigma nonzeros
// l, l+omega, ... of the transposed tile (unit-stride loads). Using the
// descriptor's bit flags and y_offset, each lane closes the rows that start
// and end inside it, a segmented lane scan adds the heads of the following
// lanes to the row each lane leaves open, and the row open at the tile start
// is parked in a per-tile calibrator that a second kernel adds in tile order
// (deterministic). The rows after the last full tile are done by a tail kernel.
// Unlike the reference kernel (which stores completed rows without beta*y,
// src/spmv_csr5.cpp:287-288), every row gets y = alpha*A*x + beta*y.
#include <algorithm>

#include "common.hpp"

namespace lsp {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMask = 0x7FFFFFFFu;

// step 1.1: row of each tile's first nonzero (:1045-1066)
__global__ void tile_ptr_search(int64_t p, int64_t m, int64_t nnz, int64_t tile, const int* __restrict__ rp,
                                uint32_t* __restrict__ tp) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g > p) return;
  if (g == p) {  // the reference leaves tile_ptr[p] = 0
    tp[p] = 0;
    return;
  }
  int64_t boundary = g * tile;
  if (boundary > nnz) boundary = nnz;
  int64_t start = 0, stop = m;
  while (stop >= start) {
    const int64_t median = (stop + start) / 2;
    if (boundary >= rp[median]) start = median + 1; else stop = median - 1;
  }
  tp[g] = (uint32_t)(start - 1);
}

// step 1.2: tiles that contain empty rows (:1068-1094)
__global__ void tile_dirty(int64_t p, const int* __restrict__ rp, const uint32_t* __restrict__ tp,
                           int* __restrict__ dirty) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= p) return;
  const uint32_t start = tp[g] & kMask, stop = tp[g + 1] & kMask;
  int d = 0;
  if (start != stop)
    for (uint32_t row = start; row <= stop; row++)
      if (rp[row] == rp[row + 1]) { d = 1; break; }
  dirty[g] = d;
}

__global__ void tile_mark(int64_t p, const int* __restrict__ dirty, uint32_t* __restrict__ tp) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < p && dirty[g]) tp[g] = (tp[g] & kMask) | 0x80000000u;
}

// step 2.1: one bit per row start inside tiles 0..p-2 (:1102-1138)
__global__ void desc_bits(int64_t m, int64_t p, int64_t tile, int omega, int sigma, int packets, int bit_all,
                          const int* __restrict__ rp, uint32_t* __restrict__ desc) {
  const int64_t rid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (rid >= m) return;
  const int64_t ptr = rp[rid];
  const int64_t t = ptr / tile;
  if (t >= p - 1) return;
  const int lx = (int)((ptr / sigma) % omega);
  const int glid = (int)(ptr % sigma) + bit_all;
  const int ly = glid / 32, llid = glid % 32;
  atomicOr(desc + t * omega * packets + (int64_t)ly * omega + lx, 1u << (31 - llid));
}

// step 2.2: y_offset and scansum_offset in the first packet (:1140-1245), and
// the per-tile count of descriptor offsets. One group of W = omega lanes per
// tile, lane l on its own descriptor word (one packet: bit_all + sigma <= 32).
// Lane l's segments: the row starts at its positions 1..sigma-1 (a popcount),
// minus one when position 0 starts no row, plus one when any position does
// (lane 0 always counts as starting); y_offset is the exclusive scan of the
// segments (less one, 0 for lane 0, in unsigned arithmetic), scansum_offset
// the number of lanes without a row start up to the next lane with one.
template <int W>
__global__ __launch_bounds__(kBlock) void desc_offsets(int64_t p, int sigma, int bit_y, int bit_all,
                                                       const uint32_t* __restrict__ tp, uint32_t* __restrict__ desc,
                                                       int64_t* __restrict__ off_count) {
  const int l = threadIdx.x & (W - 1);
  const int64_t t = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / W;
  if (t >= p) return;  // whole W-lane groups leave together
  if (l == 0) off_count[t] = 0;
  if (t >= p - 1) return;
  const uint32_t a = tp[t], b = tp[t + 1];
  if ((a & kMask) == (b & kMask)) return;  // the tile lies inside one row
  uint32_t* dt = desc + t * W;
  const uint32_t d = dt[l];
  const uint32_t flag = (d << bit_all) | ((uint32_t)(l == 0) << 31);  // bit 31 - i <-> position i
  const uint32_t rest = (flag >> (32 - sigma)) & ((1u << (sigma - 1)) - 1u);  // positions 1..sigma-1
  const int first = (flag >> 31) & 1;
  const int pres = first | (rest != 0);
  const int cnt = __popc(rest) - !first + pres;
  const int seg = cnt > 0 ? cnt : 0;
  int inc = seg;
#pragma unroll
  for (int o = 1; o < W; o <<= 1) {
    const int v = __shfl_up(inc, o, W);
    if (l >= o) inc += v;
  }
  const int total = __shfl(inc, W - 1, W);
  const uint64_t bal = __ballot(pres);
  const int gbase = (threadIdx.x & 63) & ~(W - 1);
  const uint64_t gm = (bal >> gbase) & (W == 64 ? ~0ull : ((1ull << W) - 1));
  uint32_t scan_off = 0;
  if (pres) {
    const uint64_t after = l + 1 < W ? gm >> (l + 1) : 0;
    const int next = after ? l + __ffsll((unsigned long long)after) : W;  // next lane with a row start
    scan_off = (uint32_t)(next - l - 1);
  }
  const uint32_t y_off = l ? (uint32_t)(inc - seg) - 1u : 0u;
  dt[l] = d | (y_off << (32 - bit_y)) | (scan_off << (32 - bit_all));
  if (l == 0 && ((a >> 31) & 1)) off_count[t] = total;
}

// tile_desc_offset (:1250-1361): one thread per (tile, lane) of tiles with empty rows
__global__ void desc_offset_fill(int64_t p, int64_t tile, int omega, int sigma, int packets, int bit_y, int bit_all,
                                 const int* __restrict__ rp, const uint32_t* __restrict__ tp,
                                 const uint32_t* __restrict__ desc, const int64_t* __restrict__ dop,
                                 int* __restrict__ dof) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t t = g / omega;
  const int lane = (int)(g - t * omega);
  if (t >= p - 1 || !((tp[t] >> 31) & 1)) return;
  const int rs = (int)(tp[t] & kMask), re = (int)(tp[t + 1] & kMask);
  const int64_t off_ptr = dop[t];
  auto row_of = [&](int64_t idx) {
    int start = 0, stop = re - rs - 1;
    while (stop >= start) {
      const int median = (stop + start) / 2;
      if (idx >= rp[rs + 1 + median]) start = median + 1; else stop = median - 1;
    }
    return start - 1;
  };
  const int bit_bitflag = 32 - bit_all;
  const uint32_t* dt = desc + t * omega * packets;
  int ly = 0;
  uint32_t d = dt[lane];
  int y_off = (int)(d >> (32 - bit_y));
  d <<= bit_all;
  if (!lane) d |= 0x80000000u;
  if (((d >> 31) & 1) && lane) dof[off_ptr + y_off++] = row_of(t * tile + (int64_t)lane * sigma);
  for (int i = 1; i < sigma; i++) {
    if ((!ly && i == bit_bitflag) || (ly && !(31 & (i - bit_bitflag)))) {
      ly++;
      d = dt[ly * omega + lane];
    }
    const int norm_i = 31 & (!ly ? i : i - bit_bitflag);
    if ((d >> (31 - norm_i)) & 1) dof[off_ptr + y_off++] = row_of(t * tile + (int64_t)lane * sigma + i);
  }
}

// step 3: transpose tiles 0..p-2 whose raw tile_ptr differs from the next one
template <class T>
__global__ void tile_transpose(int64_t p, int omega, int sigma, const uint32_t* __restrict__ tp,
                               const T* __restrict__ in, T* __restrict__ out) {
  const int64_t tile = (int64_t)omega * sigma;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t t = g / tile;
  if (t >= p - 1 || tp[t] == tp[t + 1]) return;
  const int idx = (int)(g - t * tile);
  const int iy = idx % sigma, ix = idx / sigma;
  out[t * tile + (int64_t)iy * omega + ix] = in[g];
}

__global__ void mark_empty_below(int64_t lim, const int* __restrict__ rp, int* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < lim) flag[i] = rp[i] == rp[i + 1];
}

__global__ void compact_list(int64_t lim, const int* __restrict__ flag, const int64_t* __restrict__ pos,
                             int* __restrict__ list) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < lim && flag[i]) list[pos[i]] = (int)i;
}

// Element i of a vector: x[i] for SpMV (S false), X[i * ld] for one column of
// a row-major multi-vector (S true: SpMM runs the kernels once per column).
template <bool S>
__device__ __forceinline__ int64_t at(int64_t i, int64_t ld) {
  return S ? i * ld : i;
}

template <class V, bool S>
__global__ void scale_list(int64_t n, const int* __restrict__ list, V* __restrict__ y, V beta, int64_t ldy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    V* yr = y + at<S>(list[i], ldy);
    *yr = beta == V(0) ? V(0) : beta * *yr;
  }
}

// ---- SpMV -------------------------------------------------------------------

// SIG > 0: sigma fixed at compile time (8 or 16), so the lane's SIG column /
// value slots and the loops over them unroll into registers. SIG = 0: sigma
// known at run time only (any other value in 1..32); the compiler then keeps
// the 32-slot arrays in scratch memory (272 bytes a lane: 1.1 GB written per
// launch on the lognormal matrix before the specialisations) and reports
// loops it leaves rolled.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wpass-failed"
template <class V, int W, bool S, int SIG>
__global__ __launch_bounds__(kBlock) void csr5_tiles(int64_t p, int sigma_rt, int bit_y, int bit_all,
                                                     const uint32_t* __restrict__ tp,
                                                     const uint32_t* __restrict__ desc,
                                                     const int* __restrict__ dop, const int* __restrict__ dof,
                                                     const int* __restrict__ col, const V* __restrict__ val,
                                                     const V* __restrict__ x, V* __restrict__ y,
                                                     V* __restrict__ calib, V alpha, V beta, int64_t ldx,
                                                     int64_t ldy) {
  const int l = threadIdx.x & (W - 1);
  const int64_t t = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / W;
  if (t >= p - 1) return;  // whole W-lane groups leave together
  constexpr int NS = SIG > 0 ? SIG : 32;  // slots of the register arrays
  const int sigma = SIG > 0 ? SIG : sigma_rt;
  const int64_t tile = (int64_t)W * sigma;
  const uint32_t raw = tp[t];
  const int rs = (int)(raw & kMask), re = (int)(tp[t + 1] & kMask);
  const uint32_t d = desc[t * W + l];
  const bool first_direct = (__shfl(d, 0, W) >> (31 - bit_all)) & 1;
  const int* tc = col + t * tile;
  const V* tv = val + t * tile;
  // all loads of the lane issued before use; the tile's columns and values
  // are read once (non-temporal loads keep x in the caches)
  int c[NS];
  V v[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) {
    if (i < sigma) {
      c[i] = __builtin_nontemporal_load(tc + (int64_t)i * W + l);
      v[i] = __builtin_nontemporal_load(tv + (int64_t)i * W + l);
    }
  }
#pragma unroll
  for (int i = 0; i < NS; i++)
    if (i < sigma) v[i] *= x[at<S>(c[i], ldx)];
  if (rs == re) {  // the whole tile lies inside one row
    V s = 0;
#pragma unroll
    for (int i = 0; i < NS; i++)
      if (i < sigma) s += v[i];
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, W);
    if (l == 0) {
      if (first_direct) {
        y[at<S>(rs, ldy)] = beta == V(0) ? alpha * s : alpha * s + beta * y[at<S>(rs, ldy)];
        calib[t] = 0;
      } else {
        calib[t] = s;
      }
    }
    return;
  }
  const bool empty_rows = (raw >> 31) & 1;
  const int off_ptr = empty_rows ? dop[t] : 0;
  const uint32_t y_off = d >> (32 - bit_y);
  uint32_t flags = d << bit_all;  // bit 31 - i <-> position i
  if (l == 0) flags |= 0x80000000u;
  auto row_of = [&](int seg) -> int64_t {
    return empty_rows ? (int64_t)rs + 1 + dof[off_ptr + seg - 1] : (int64_t)rs + seg;
  };
  // lane-local segments
  V acc = 0, head = 0;
  int seg = -1;  // open segment started in this lane (-1: the inherited head is open)
#pragma unroll
  for (int i = 0; i < NS; i++) {
    if (i >= sigma) break;
    if ((flags >> (31 - i)) & 1) {
      if (seg < 0) {
        head = acc;
        seg = (l == 0) ? 0 : (int)y_off + 1;  // lane 0 opens segment 0 at position 0
      } else if (seg == 0) {
        if (first_direct) y[at<S>(rs, ldy)] = beta == V(0) ? alpha * acc : alpha * acc + beta * y[at<S>(rs, ldy)];
        else calib[t] = acc;
        seg = 1;
      } else {
        const int64_t r = row_of(seg);
        y[at<S>(r, ldy)] = beta == V(0) ? alpha * acc : alpha * acc + beta * y[at<S>(r, ldy)];
        seg++;
      }
      acc = 0;
    }
    acc += v[i];
  }
  const bool has_flag = seg >= 0;
  const V h = has_flag ? head : acc;  // the part of this lane belonging to an earlier lane's row
  // inclusive segmented scan of h; a segment starts after every lane with a flag
  const uint64_t fmask = __ballot(has_flag);
  const int gbase = (threadIdx.x & 63) & ~(W - 1);
  const uint64_t gmask = (fmask >> gbase) & (W == 64 ? ~0ull : ((1ull << W) - 1));
  V sc = h;
#pragma unroll
  for (int o = 1; o < W; o <<= 1) {
    const V up = __shfl_up(sc, o, W);
    // lanes l-o+1 .. l must not contain a segment start (a flag at lanes l-o .. l-1)
    const uint64_t span = ((1ull << o) - 1) << (l - o);
    if (l >= o && !(gmask & span)) sc += up;
  }
  // lane l with a flag reads the scan at the next flagged lane (or the last lane)
  int target = W - 1;
  const uint64_t after = l + 1 < W ? (gmask >> (l + 1)) << (l + 1) : 0;
  if (after) target = __ffsll((unsigned long long)after) - 1;
  const V carry_all = __shfl(sc, target, W);
  if (!has_flag) return;
  const V total = acc + (l + 1 < W ? carry_all : V(0));
  if (seg == 0) {
    if (first_direct) y[at<S>(rs, ldy)] = beta == V(0) ? alpha * total : alpha * total + beta * y[at<S>(rs, ldy)];
    else calib[t] = total;
  } else {
    const int64_t r = row_of(seg);
    y[at<S>(r, ldy)] = beta == V(0) ? alpha * total : alpha * total + beta * y[at<S>(r, ldy)];
  }
  if (l == 0 && first_direct) calib[t] = 0;
}

#pragma clang diagnostic pop

// Continuation rows: the first continuation tile of a chain adds the
// calibrators of the chain in tile order (reference calibrate kernel, :367-384).
template <class V, bool S>
__global__ void csr5_calibrate(int64_t p, int W, int bit_all, const uint32_t* __restrict__ tp,
                               const uint32_t* __restrict__ desc, const V* __restrict__ calib, V* __restrict__ y,
                               V alpha, int64_t ldy) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p - 1) return;
  auto direct = [&](int64_t q) { return (desc[q * W] >> (31 - bit_all)) & 1; };
  const uint32_t r = tp[t] & kMask;
  if (direct(t)) return;
  if (t > 0 && (tp[t - 1] & kMask) == r && !direct(t - 1)) return;  // not the chain head
  V sum = 0;
  for (int64_t q = t; q < p - 1 && (tp[q] & kMask) == r && !direct(q); q++) sum += calib[q];
  y[at<S>(r, ldy)] += alpha * sum;
}

// Rows after the last full tile (reference tail kernel, :386-424).
template <class V, bool S>
__global__ void csr5_tail(int64_t m, int64_t tail_start, int64_t first_tail, const int* __restrict__ rp,
                          const int* __restrict__ col, const V* __restrict__ val, const V* __restrict__ x,
                          V* __restrict__ y, V alpha, V beta, int64_t ldx, int64_t ldy) {
  const int64_t row = tail_start + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= m) return;
  const int64_t s = row == tail_start ? first_tail : rp[row];
  V sum = 0;
  for (int64_t i = s; i < rp[row + 1]; i++) sum += val[i] * x[at<S>(col[i], ldx)];
  V* yr = y + at<S>(row, ldy);
  if (row == tail_start && rp[row] != first_tail) *yr += alpha * sum;
  else *yr = beta == V(0) ? alpha * sum : alpha * sum + beta * *yr;
}

template <class V, int W, bool S, int SIG>
void launch_tiles_sig(Csr5Data& d, double alpha, const void* x, double beta, void* y, int64_t ldx, int64_t ldy,
                      hipStream_t s) {
  const int bit_all = d.bit_y + d.bit_scan;
  hipLaunchKernelGGL((csr5_tiles<V, W, S, SIG>), grid1d((d.p - 1) * W, kBlock), dim3(kBlock), 0, s, d.p, d.sigma,
                     d.bit_y, bit_all, d.tile_ptr.as<uint32_t>(), d.tile_desc.as<uint32_t>(),
                     d.desc_off_ptr.as<int>(), d.desc_off.as<int>(), d.col.as<int>(), d.val.as<V>(),
                     static_cast<const V*>(x), static_cast<V*>(y), d.calib.as<V>(), (V)alpha, (V)beta, ldx, ldy);
  LSP_LAUNCHED();
}

// Sigma 8 or 16 (the reference's CSR5_SIGMA and the tuning grid's values)
// takes the register-resident instantiations.
template <class V, int W, bool S>
void launch_tiles(Matrix& A, Csr5Data& d, double alpha, const void* x, double beta, void* y, int64_t ldx,
                  int64_t ldy, hipStream_t s) {
  (void)A;
  if (d.sigma == 16) return launch_tiles_sig<V, W, S, 16>(d, alpha, x, beta, y, ldx, ldy, s);
  if (d.sigma == 8) return launch_tiles_sig<V, W, S, 8>(d, alpha, x, beta, y, ldx, ldy, s);
  launch_tiles_sig<V, W, S, 0>(d, alpha, x, beta, y, ldx, ldy, s);
}

// One product: y = alpha A x + beta y, with x and y strided by ldx / ldy
// when S (one column of a row-major multi-vector).
template <class V, bool S>
void run(Matrix& A, Csr5Data& d, double alpha, const void* x, double beta, void* y, int64_t ldx, int64_t ldy,
         hipStream_t s) {
  V* yv = static_cast<V*>(y);
  if (d.num_empty) {
    hipLaunchKernelGGL((scale_list<V, S>), grid1d(d.num_empty, kBlock), dim3(kBlock), 0, s, d.num_empty,
                       d.empty_rows.as<int>(), yv, (V)beta, ldy);
    LSP_LAUNCHED();
  }
  if (d.p > 1) {
    switch (d.omega) {
      case 64: launch_tiles<V, 64, S>(A, d, alpha, x, beta, y, ldx, ldy, s); break;
      case 32: launch_tiles<V, 32, S>(A, d, alpha, x, beta, y, ldx, ldy, s); break;
      case 16: launch_tiles<V, 16, S>(A, d, alpha, x, beta, y, ldx, ldy, s); break;
      case 8: launch_tiles<V, 8, S>(A, d, alpha, x, beta, y, ldx, ldy, s); break;
      case 4: launch_tiles<V, 4, S>(A, d, alpha, x, beta, y, ldx, ldy, s); break;
      default: throw Error(LESPMV_ERR_UNSUPPORTED, "CSR5 omega must be a power of two <= 64");
    }
    hipLaunchKernelGGL((csr5_calibrate<V, S>), grid1d(d.p - 1, kBlock), dim3(kBlock), 0, s, d.p, d.omega,
                       d.bit_y + d.bit_scan, d.tile_ptr.as<uint32_t>(), d.tile_desc.as<uint32_t>(), d.calib.as<V>(),
                       yv, (V)alpha, ldy);
    LSP_LAUNCHED();
  }
  const int64_t first_tail = (d.p - 1) * (int64_t)d.omega * d.sigma;
  hipLaunchKernelGGL((csr5_tail<V, S>), grid1d(A.m - d.tail_start, kBlock), dim3(kBlock), 0, s, A.m,
                     d.tail_start, first_tail, A.rp.as<int>(), d.col.as<int>(), d.val.as<V>(),
                     static_cast<const V*>(x), yv, (V)alpha, (V)beta, ldx, ldy);
  LSP_LAUNCHED();
}

}  // namespace

void convert_csr5(Matrix& A, const lespmv_params_t& prm, hipStream_t s) {
  LSP_REQUIRE(A.nnz > 0, LESPMV_ERR_UNSUPPORTED, "CSR5 needs at least one nonzero");
  const int omega = prm.omega, sigma = prm.sigma;
  LSP_REQUIRE(omega >= 4 && omega <= 64 && (omega & (omega - 1)) == 0, LESPMV_ERR_UNSUPPORTED,
              "CSR5 omega must be a power of two in [4, 64]");
  LSP_REQUIRE(sigma >= 1 && sigma <= 32, LESPMV_ERR_UNSUPPORTED, "CSR5 sigma must be in [1, 32]");
  auto d = std::make_unique<Csr5Data>();
  d->omega = omega;
  d->sigma = sigma;
  d->bit_y = 1;
  for (int base = 2; base < sigma * omega; base *= 2) d->bit_y++;
  d->bit_scan = 1;
  for (int base = 2; base < omega; base *= 2) d->bit_scan++;
  const int bit_all = d->bit_y + d->bit_scan;
  LSP_REQUIRE(bit_all <= 31, LESPMV_ERR_UNSUPPORTED, "UNSUPPORTED CSR5 OMEGA for bit saving");
  d->packets = (bit_all + sigma + 31) / 32;
  LSP_REQUIRE(d->packets == 1, LESPMV_ERR_UNSUPPORTED, "CSR5 descriptor must fit one 32-bit packet");
  const int64_t tile = (int64_t)omega * sigma;
  const int64_t p = ceil_div(A.nnz, tile);
  d->p = p;
  d->tile_ptr.alloc(sizeof(uint32_t) * (p + 1));
  d->tile_desc.alloc(sizeof(uint32_t) * p * omega * d->packets);
  d->desc_off_ptr.alloc(sizeof(int) * (p + 1));
  LSP_HIP(hipMemsetAsync(d->tile_desc.p, 0, sizeof(uint32_t) * p * omega * d->packets, s));
  hipLaunchKernelGGL(tile_ptr_search, grid1d(p + 1, kBlock), dim3(kBlock), 0, s, p, A.m, A.nnz, tile,
                     A.rp.as<int>(), d->tile_ptr.as<uint32_t>());
  LSP_LAUNCHED();
  DBuf dirty;
  dirty.alloc(sizeof(int) * p);
  hipLaunchKernelGGL(tile_dirty, grid1d(p, kBlock), dim3(kBlock), 0, s, p, A.rp.as<int>(), d->tile_ptr.as<uint32_t>(),
                     dirty.as<int>());
  LSP_LAUNCHED();
  hipLaunchKernelGGL(tile_mark, grid1d(p, kBlock), dim3(kBlock), 0, s, p, dirty.as<int>(), d->tile_ptr.as<uint32_t>());
  LSP_LAUNCHED();
  uint32_t tail_raw = 0;
  LSP_HIP(hipMemcpyAsync(&tail_raw, d->tile_ptr.as<uint32_t>() + p - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  LSP_HIP(hipStreamSynchronize(s));
  d->tail_start = tail_raw & kMask;
  hipLaunchKernelGGL(desc_bits, grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, p, tile, omega, sigma, d->packets,
                     bit_all, A.rp.as<int>(), d->tile_desc.as<uint32_t>());
  LSP_LAUNCHED();
  DBuf cnt;
  cnt.alloc(sizeof(int64_t) * (p + 1));
  {
    auto offsets = [&](auto kern) {
      hipLaunchKernelGGL(kern, grid1d(p * omega, kBlock), dim3(kBlock), 0, s, p, sigma, d->bit_y, bit_all,
                         d->tile_ptr.as<uint32_t>(), d->tile_desc.as<uint32_t>(), cnt.as<int64_t>());
    };
    switch (omega) {  // omega is a power of two in [4, 64], checked above
      case 4: offsets(desc_offsets<4>); break;
      case 8: offsets(desc_offsets<8>); break;
      case 16: offsets(desc_offsets<16>); break;
      case 32: offsets(desc_offsets<32>); break;
      default: offsets(desc_offsets<64>); break;
    }
    LSP_LAUNCHED();
  }
  d->num_offsets = scan_exclusive_i64(cnt.as<int64_t>(), p, s);
  LSP_HIP(hipMemcpyAsync(cnt.as<int64_t>() + p, &d->num_offsets, sizeof(int64_t), hipMemcpyHostToDevice, s));
  {
    auto h = to_host<int64_t>(cnt.p, p + 1, s);
    std::vector<int32_t> h32(h.begin(), h.end());
    LSP_HIP(hipMemcpyAsync(d->desc_off_ptr.p, h32.data(), sizeof(int32_t) * h32.size(), hipMemcpyHostToDevice, s));
    LSP_HIP(hipStreamSynchronize(s));
  }
  d->desc_off.alloc(sizeof(int) * std::max<int64_t>(d->num_offsets, 1));
  LSP_HIP(hipMemsetAsync(d->desc_off.p, 0, sizeof(int) * std::max<int64_t>(d->num_offsets, 1), s));
  if (d->num_offsets) {
    hipLaunchKernelGGL(desc_offset_fill, grid1d((p - 1) * omega, kBlock), dim3(kBlock), 0, s, p, tile, omega, sigma,
                       d->packets, d->bit_y, bit_all, A.rp.as<int>(), d->tile_ptr.as<uint32_t>(),
                       d->tile_desc.as<uint32_t>(), cnt.as<int64_t>(), d->desc_off.as<int>());
    LSP_LAUNCHED();
  }
  // transposed copies of col/val (tiles 0..p-2), tail in CSR order
  const size_t vs = vsize(A.dtype);
  d->col.alloc(sizeof(int) * A.nnz);
  d->val.alloc(vs * A.nnz);
  LSP_HIP(hipMemcpyAsync(d->col.p, A.ci.p, sizeof(int) * A.nnz, hipMemcpyDeviceToDevice, s));
  LSP_HIP(hipMemcpyAsync(d->val.p, A.val.p, vs * A.nnz, hipMemcpyDeviceToDevice, s));
  if (p > 1) {
    hipLaunchKernelGGL((tile_transpose<int>), grid1d((p - 1) * tile, kBlock), dim3(kBlock), 0, s, p, omega, sigma,
                       d->tile_ptr.as<uint32_t>(), A.ci.as<int>(), d->col.as<int>());
    LSP_LAUNCHED();
    if (A.dtype == LESPMV_F64)
      hipLaunchKernelGGL((tile_transpose<double>), grid1d((p - 1) * tile, kBlock), dim3(kBlock), 0, s, p, omega,
                         sigma, d->tile_ptr.as<uint32_t>(), A.val.as<double>(), d->val.as<double>());
    else
      hipLaunchKernelGGL((tile_transpose<float>), grid1d((p - 1) * tile, kBlock), dim3(kBlock), 0, s, p, omega, sigma,
                         d->tile_ptr.as<uint32_t>(), A.val.as<float>(), d->val.as<float>());
    LSP_LAUNCHED();
  }
  // empty rows before the tail get y = beta*y from a list
  if (d->tail_start > 0) {
    DBuf flag, pos;
    flag.alloc(sizeof(int) * d->tail_start);
    pos.alloc(sizeof(int64_t) * (d->tail_start + 1));
    hipLaunchKernelGGL(mark_empty_below, grid1d(d->tail_start, kBlock), dim3(kBlock), 0, s, d->tail_start,
                       A.rp.as<int>(), flag.as<int>());
    LSP_LAUNCHED();
    d->num_empty = scan_exclusive_i32_to_i64(flag.as<int32_t>(), pos.as<int64_t>(), d->tail_start, s);
    d->empty_rows.alloc(sizeof(int) * std::max<int64_t>(d->num_empty, 1));
    if (d->num_empty) {
      hipLaunchKernelGGL(compact_list, grid1d(d->tail_start, kBlock), dim3(kBlock), 0, s, d->tail_start,
                         flag.as<int>(), pos.as<int64_t>(), d->empty_rows.as<int>());
      LSP_LAUNCHED();
    }
  }
  d->calib.alloc(vs * std::max<int64_t>(p, 1));
  LSP_HIP(hipStreamSynchronize(s));
  A.fmt[LESPMV_CSR5] = std::move(d);
}

void csr5_spmv(Matrix& A, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  auto& d = *static_cast<Csr5Data*>(A.fmt[LESPMV_CSR5].get());
  if (A.m == 0) return;
  if (A.dtype == LESPMV_F64) run<double, false>(A, d, alpha, x, beta, y, 1, 1, s);
  else run<float, false>(A, d, alpha, x, beta, y, 1, 1, s);
}

// SpMM column by column: the tile, calibration and tail kernels once per
// right-hand side, reading column j of X and writing column j of Y in place
// (the per-tile calibrators are reused, in stream order). Each column is
// therefore the SpMV of that column, bit for bit.
void csr5_spmm(Matrix& A, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y, int64_t ldy,
               hipStream_t s) {
  auto& d = *static_cast<Csr5Data*>(A.fmt[LESPMV_CSR5].get());
  if (A.m == 0) return;
  const size_t sv = vsize(A.dtype);
  for (int j = 0; j < k; j++) {
    const void* xj = static_cast<const char*>(X) + j * sv;
    void* yj = static_cast<char*>(Y) + j * sv;
    if (A.dtype == LESPMV_F64) run<double, true>(A, d, alpha, xj, beta, yj, ldx, ldy, s);
    else run<float, true>(A, d, alpha, xj, beta, yj, ldx, ldy, s);
  }
}

void csr5_export(Matrix& A, const std::string& name, std::vector<unsigned char>& out) {
  auto& d = *static_cast<Csr5Data*>(A.fmt[LESPMV_CSR5].get());
  hipStream_t s = A.stream;
  if (name == "row_offset") put_bytes(out, to_host<int32_t>(A.rp.p, A.m + 1, s));
  else if (name == "col_index") put_bytes(out, to_host<int32_t>(d.col.p, A.nnz, s));
  else if (name == "values") {
    out.resize(A.nnz * vsize(A.dtype));
    LSP_HIP(hipMemcpyAsync(out.data(), d.val.p, out.size(), hipMemcpyDeviceToHost, s));
    LSP_HIP(hipStreamSynchronize(s));
  } else if (name == "tile_ptr") put_bytes(out, to_host<uint32_t>(d.tile_ptr.p, d.p + 1, s));
  else if (name == "tile_desc") put_bytes(out, to_host<uint32_t>(d.tile_desc.p, d.p * d.omega * d.packets, s));
  else if (name == "tile_desc_offset_ptr") put_bytes(out, to_host<int32_t>(d.desc_off_ptr.p, d.p + 1, s));
  else if (name == "tile_desc_offset") put_bytes(out, to_host<int32_t>(d.desc_off.p, d.num_offsets, s));
  else if (name == "omega") put_scalar(out, d.omega);
  else if (name == "sigma") put_scalar(out, d.sigma);
  else if (name == "bit_y_offset") put_scalar(out, d.bit_y);
  else if (name == "bit_scansum_offset") put_scalar(out, d.bit_scan);
  else if (name == "num_packets") put_scalar(out, d.packets);
  else if (name == "num_offsets") put_scalar(out, d.num_offsets);
  else if (name == "p") put_scalar(out, d.p);
  else if (name == "tail_partition_start") put_scalar(out, d.tail_start);
  else throw Error(LESPMV_ERR_ARG, "unknown CSR5 array " + name);
}

}  // namespace lsp
