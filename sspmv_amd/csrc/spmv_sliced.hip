// Sliced-ELL family on gfx950: ELL, S-ELL, SELL-c-sigma and SELL-c-R.
//
// Reference: csr_to_ell / csr_to_sell / csr_to_sell_c_sigma / csr_to_sell_c_R
// (include/sparse_conversion.h:186-598) and the kernels of src/spmv_ell.cpp,
// src/spmv_s_ell.cpp, src/spmv_sell_c_sigma.cpp, src/spmv_sell_c_R.cpp.
//
// Logical content is the reference's: rows (optionally reordered by a per-slice
// sort on (nnz desc, row desc)) are grouped in chunks of C rows, each chunk
// padded to chunk_len = max row length rounded up to `alignment`, pads are
// col -1 / value 0. Device layout differs: inside a chunk slots are
// column-major, slot (r, j) at chunk_off[c] + j*C + r, so with C = 64 one
// wavefront reads position j of 64 rows with unit stride. ELL is one chunk of
// C = M rows (the reference's ColMajor layout). Export transposes back to the
// reference's row-major blocks.
#include <algorithm>

#include "common.hpp"
#include "sliced_engine.hpp"

namespace lsp {
namespace {

constexpr int kBlock = 256;

__global__ void row_lengths(int64_t m, const int* __restrict__ rp, int* __restrict__ len) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) len[i] = rp[i + 1] - rp[i];
}

// Composite key: slice (asc), maxlen - len (asc = len desc), M-1-row (asc = row desc).
__global__ void sigma_keys(int64_t m, const int* __restrict__ len, int64_t sigma, int lenbits, int rowbits,
                           int maxlen, uint64_t* __restrict__ keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint64_t slice = (uint64_t)(i / sigma);
  const uint64_t hi = (lenbits + rowbits) >= 64 ? 0 : (slice << (lenbits + rowbits));
  keys[i] = hi | ((uint64_t)(maxlen - len[i]) << rowbits) | (uint64_t)(m - 1 - i);
}

__global__ void keys_to_order(int64_t m, const uint64_t* __restrict__ keys, int rowbits, int* __restrict__ order) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint64_t mask = rowbits >= 64 ? ~0ull : ((1ull << rowbits) - 1);
  order[i] = (int)(m - 1 - (int64_t)(keys[i] & mask));
}

__global__ void chunk_widths(int64_t m, int64_t nchunks, int C, int alignment, const int* __restrict__ len,
                             const int* __restrict__ order, int* __restrict__ clen, int64_t* __restrict__ cslots) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  int w = 0;
  const int64_t r0 = c * C, r1 = min(m, r0 + C);
  for (int64_t r = r0; r < r1; r++) w = max(w, len[order ? order[r] : r]);
  w = (int)(((int64_t)w + alignment - 1) / alignment * alignment);
  clen[c] = w;
  cslots[c] = (int64_t)w * C;
}

template <class V>
__global__ void sliced_fill(int64_t m, int C, const int* __restrict__ rp, const int* __restrict__ ci,
                            const V* __restrict__ val, const int* __restrict__ order,
                            const int64_t* __restrict__ coff, int* __restrict__ col, V* __restrict__ sv) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int64_t real = order ? order[t] : t;
  const int64_t c = t / C, r = t % C;
  const int64_t base = coff[c] + r;
  const int s = rp[real], e = rp[real + 1];
  for (int j = s; j < e; j++) {
    col[base + (int64_t)(j - s) * C] = ci[j];
    sv[base + (int64_t)(j - s) * C] = val[j];
  }
}

// Work unit = (chunk, slot range [j0, j1), partial slot). Long chunks are cut
// into units of at most kUnitSlots positions so no wavefront walks a long row
// alone; a unit that covers a whole chunk writes y, the others write partials
// that sliced_fixup adds in unit order. One thread per (unit, row of chunk);
// padding sits at the end of each row, so a col < 0 ends the row.
// Column and value slots are read once, with non-temporal loads, so they do
// not push x out of the caches (27-point stencil 160^3 FP64 278 -> 263 us,
// FP32 170 -> 154 us; lognormal rows 905 -> 887 us; 2-D Laplacian 1024^2
// 24.1 -> 21.6 us). Slots per pipelined step (kUnroll) follow the layout's
// slot-weighted mean unit width w (sum of w^2 over sum of w, computed by
// plan_units): 6 up to 8 positions (2-D Laplacian, 5: 16.9 us against 21.6
// with 12; the registers of a short step leave room for 8 waves per SIMD),
// else 12 (27-point stencil, 27: 263 us against 265 with 6 and 267 with 16;
// on lognormal rows, about 38, 24 gained 1-2 % for SELL-C-sigma and lost 6 %
// for S-ELL, whose unsorted chunks leave most lanes idle in wide steps).
inline int unroll_for(double w) { return w <= 8.0 ? 6 : 12; }

template <class V, int kUnroll>
__global__ __launch_bounds__(kBlock) void sliced_spmv_kernel(int64_t m, int64_t nunits, int C,
                                                             const int4* __restrict__ units,
                                                             const int64_t* __restrict__ coff,
                                                             const int* __restrict__ order,
                                                             const int* __restrict__ col, const V* __restrict__ sv,
                                                             const V* __restrict__ x, V* __restrict__ y,
                                                             V* __restrict__ partial, V alpha, V beta,
                                                             int remap) {
  const int64_t blk = remap ? xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
  const int64_t t = blk * kBlock + threadIdx.x;
  const int64_t u = t / C;
  if (u >= nunits) return;
  const int r = (int)(t - u * C);
  const int4 un = units[u];  // chunk, j0, j1, partial slot (-1: whole chunk)
  const int64_t row = (int64_t)un.x * C + r;
  if (row >= m) return;
  const V sum = sliced_row_sum<V, kUnroll>(col + coff[un.x] + r, sv + coff[un.x] + r, C, un.y, un.z, x);
  if (un.w < 0) {
    const int64_t out = order ? order[row] : row;
    y[out] = beta == V(0) ? alpha * sum : alpha * sum + beta * y[out];
  } else {
    partial[(int64_t)un.w * C + r] = sum;
  }
}

// Split chunks: (chunk, first partial slot, unit count).
template <class V>
__global__ void sliced_fixup(int64_t m, int64_t nsplit, int C, const int4* __restrict__ split,
                             const int* __restrict__ order, const V* __restrict__ partial, V* __restrict__ y, V alpha,
                             V beta) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t e = t / C;
  if (e >= nsplit) return;
  const int r = (int)(t - e * C);
  const int4 sp = split[e];
  const int64_t row = (int64_t)sp.x * C + r;
  if (row >= m) return;
  V total = 0;
  for (int q = 0; q < sp.z; q++) total += partial[(int64_t)(sp.y + q) * C + r];
  const int64_t out = order ? order[row] : row;
  y[out] = beta == V(0) ? alpha * total : alpha * total + beta * y[out];
}

// SpMM: KP lanes per row, lane owns column c of the row-major X / Y.
template <class V, int KP>
__global__ __launch_bounds__(kBlock) void sliced_spmm_kernel(int64_t m, int k, int C, const int* __restrict__ clen,
                                                             const int64_t* __restrict__ coff,
                                                             const int* __restrict__ order,
                                                             const int* __restrict__ col, const V* __restrict__ sv,
                                                             const V* __restrict__ X, int64_t ldx,
                                                             V* __restrict__ Y, int64_t ldy, V alpha, V beta) {
  const int cc = threadIdx.x & (KP - 1);
  const int64_t t = (int64_t)blockIdx.x * (kBlock / KP) + threadIdx.x / KP;
  if (t >= m || cc >= k) return;
  const int64_t c = t / C, r = t - c * C;
  const int L = clen[c];
  const int* cp = col + coff[c] + r;
  const V* vp = sv + coff[c] + r;
  V sum = 0;
  for (int j = 0; j < L; j++) {
    const int cj = cp[(int64_t)j * C];
    if (cj < 0) break;
    sum += vp[(int64_t)j * C] * X[(int64_t)cj * ldx + cc];
  }
  const int64_t out = order ? order[t] : t;
  V* yp = Y + out * ldy + cc;
  *yp = beta == V(0) ? alpha * sum : alpha * sum + beta * *yp;
}

}  // namespace

struct SlicedUnits : SlicedData {
  int64_t nunits = 0, nsplit = 0;
  int unroll = 12;  // slots per pipelined step of the SpMV kernel
  DBuf units;    // int4[nunits]: chunk, j0, j1, partial slot or -1
  DBuf split;    // int4[nsplit]: chunk, first partial slot, unit count
  DBuf partial;  // V[partial slots * C]
  int64_t device_bytes() const override {
    return SlicedData::device_bytes() + units.bytes + split.bytes + partial.bytes;
  }
};

namespace {

// Cut every chunk into units of at most kUnitSlots positions.
void plan_units(SlicedUnits& d, const std::vector<int>& clen, size_t vbytes, hipStream_t s) {
  std::vector<int4> units, split;
  int pslots = 0;
  for (int64_t c = 0; c < (int64_t)clen.size(); c++) {
    const int L = clen[c];
    if (L <= kUnitSlots) {
      units.push_back(make_int4((int)c, 0, L, -1));
      continue;
    }
    const int cnt = (L + kUnitSlots - 1) / kUnitSlots;
    split.push_back(make_int4((int)c, pslots, cnt, 0));
    for (int q = 0; q < cnt; q++) units.push_back(make_int4((int)c, q * kUnitSlots, std::min(L, (q + 1) * kUnitSlots),
                                                            pslots + q));
    pslots += cnt;
  }
  double sw = 0, sw2 = 0;
  for (const int4& u : units) {
    sw += (double)(u.z - u.y);
    sw2 += (double)(u.z - u.y) * (double)(u.z - u.y);
  }
  d.unroll = unroll_for(sw > 0 ? sw2 / sw : 0.0);
  d.nunits = (int64_t)units.size();
  d.nsplit = (int64_t)split.size();
  d.units.alloc(sizeof(int4) * std::max<size_t>(units.size(), 1));
  d.split.alloc(sizeof(int4) * std::max<size_t>(split.size(), 1));
  d.partial.alloc(vbytes * (size_t)std::max(pslots, 1) * d.C);
  if (!units.empty())
    LSP_HIP(hipMemcpyAsync(d.units.p, units.data(), sizeof(int4) * units.size(), hipMemcpyHostToDevice, s));
  if (!split.empty())
    LSP_HIP(hipMemcpyAsync(d.split.p, split.data(), sizeof(int4) * split.size(), hipMemcpyHostToDevice, s));
  LSP_HIP(hipStreamSynchronize(s));
}

int bits_for(uint64_t v) {  // number of bits to represent values 0..v
  int b = 0;
  while (b < 64 && (v >> b)) b++;
  return b;
}

template <class V>
void fill_dispatch(Matrix& A, SlicedData& d, const int* order, hipStream_t s) {
  hipLaunchKernelGGL((sliced_fill<V>), grid1d(A.m, kBlock), dim3(kBlock), 0, s, A.m, d.C, A.rp.as<int>(),
                     A.ci.as<int>(), A.val.as<V>(), order, d.chunk_off.as<int64_t>(), d.col.as<int>(),
                     d.val.as<V>());
  LSP_LAUNCHED();
}

template <class V, int KP>
void spmm_launch(Matrix& A, SlicedData& d, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
                 int64_t ldy, hipStream_t s) {
  hipLaunchKernelGGL((sliced_spmm_kernel<V, KP>), grid1d(A.m, kBlock / KP), dim3(kBlock), 0, s, A.m, k, d.C,
                     d.chunk_len.as<int>(), d.chunk_off.as<int64_t>(), d.reorder.as<int>(), d.col.as<int>(),
                     d.val.as<V>(), static_cast<const V*>(X), ldx, static_cast<V*>(Y), ldy, (V)alpha, (V)beta);
  LSP_LAUNCHED();
}

template <class V>
void spmm_dispatch(Matrix& A, SlicedData& d, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
                   int64_t ldy, hipStream_t s) {
  if (k <= 1) spmm_launch<V, 1>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 2) spmm_launch<V, 2>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 4) spmm_launch<V, 4>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 8) spmm_launch<V, 8>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 16) spmm_launch<V, 16>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else if (k <= 32) spmm_launch<V, 32>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else spmm_launch<V, 64>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
}

// Device column-major chunk -> reference row-major chunk (host side).
template <class T>
std::vector<T> to_row_major(const std::vector<T>& dev, const std::vector<int>& clen, const std::vector<int64_t>& off,
                            int C) {
  std::vector<T> out(dev.size());
  for (size_t c = 0; c < clen.size(); c++) {
    const int64_t L = clen[c];
    for (int64_t r = 0; r < C; r++)
      for (int64_t j = 0; j < L; j++) out[off[c] + r * L + j] = dev[off[c] + j * C + r];
  }
  return out;
}

}  // namespace

void convert_sliced(Matrix& A, int fmt, const lespmv_params_t& p, hipStream_t s) {
  auto d = std::make_unique<SlicedUnits>();
  d->fmt = fmt;
  const int64_t m = A.m;
  const bool ell = fmt == LESPMV_ELL;
  const bool sorted = fmt == LESPMV_SELL_C_SIGMA || fmt == LESPMV_SELL_C_R;
  if (ell) {
    d->C = (int32_t)std::max<int64_t>(m, 1);
    d->alignment = 1;
    d->sigma = 1;
  } else {
    LSP_REQUIRE(p.chunk > 0, LESPMV_ERR_ARG, "chunk (C) must be positive");
    LSP_REQUIRE(p.alignment > 0, LESPMV_ERR_ARG, "alignment must be positive");
    d->C = p.chunk;
    d->alignment = p.alignment;
    d->sigma = fmt == LESPMV_SELL_C_SIGMA ? p.sigma : (fmt == LESPMV_SELL_C_R ? (int32_t)std::max<int64_t>(m, 1) : 1);
    if (fmt == LESPMV_SELL_C_SIGMA) {
      LSP_REQUIRE(p.sigma > 0, LESPMV_ERR_ARG, "sigma must be positive");
      // include/sparse_conversion.h:357-361
      LSP_REQUIRE(p.sigma % p.chunk == 0, LESPMV_ERR_UNSUPPORTED, "Sigma must be multiple divided by c");
    }
  }
  const int C = d->C;
  d->nchunks = ceil_div(m, C);
  DBuf len;
  len.alloc(sizeof(int) * std::max<int64_t>(m, 1));
  if (m) {
    hipLaunchKernelGGL(row_lengths, grid1d(m, kBlock), dim3(kBlock), 0, s, m, A.rp.as<int>(), len.as<int>());
    LSP_LAUNCHED();
  }
  if (sorted && m > 1) {
    const int maxlen = (int)reduce_max_rowlen(A.rp.as<int32_t>(), m, s);
    const int64_t slices = ceil_div(m, d->sigma);
    const int lenbits = bits_for((uint64_t)maxlen), rowbits = bits_for((uint64_t)(m - 1));
    const int slicebits = bits_for((uint64_t)(slices - 1));
    LSP_REQUIRE(lenbits + rowbits + slicebits <= 64, LESPMV_ERR_UNSUPPORTED, "sigma sort key exceeds 64 bits");
    DBuf keys;
    keys.alloc(sizeof(uint64_t) * m);
    hipLaunchKernelGGL(sigma_keys, grid1d(m, kBlock), dim3(kBlock), 0, s, m, len.as<int>(), (int64_t)d->sigma,
                       lenbits, rowbits, maxlen, keys.as<uint64_t>());
    LSP_LAUNCHED();
    radix_sort_u64(keys.as<uint64_t>(), m, lenbits + rowbits + slicebits, s);
    d->reorder.alloc(sizeof(int) * m);
    hipLaunchKernelGGL(keys_to_order, grid1d(m, kBlock), dim3(kBlock), 0, s, m, keys.as<uint64_t>(), rowbits,
                       d->reorder.as<int>());
    LSP_LAUNCHED();
  } else if (sorted) {
    d->reorder.alloc(sizeof(int) * std::max<int64_t>(m, 1));
    LSP_HIP(hipMemsetAsync(d->reorder.p, 0, sizeof(int) * std::max<int64_t>(m, 1), s));
  }
  const int* order = sorted ? d->reorder.as<int>() : nullptr;
  d->chunk_len.alloc(sizeof(int) * std::max<int64_t>(d->nchunks, 1));
  d->chunk_off.alloc(sizeof(int64_t) * (d->nchunks + 1));
  if (d->nchunks) {
    hipLaunchKernelGGL(chunk_widths, grid1d(d->nchunks, kBlock), dim3(kBlock), 0, s, m, d->nchunks, C, d->alignment,
                       len.as<int>(), order, d->chunk_len.as<int>(), d->chunk_off.as<int64_t>());
    LSP_LAUNCHED();
  }
  d->slots = scan_exclusive_i64(d->chunk_off.as<int64_t>(), d->nchunks, s);
  LSP_HIP(hipMemcpyAsync(d->chunk_off.as<int64_t>() + d->nchunks, &d->slots, sizeof(int64_t), hipMemcpyHostToDevice,
                         s));
  LSP_HIP(hipStreamSynchronize(s));
  d->max_width = ell ? (int32_t)(d->nchunks ? reduce_max_rowlen(A.rp.as<int32_t>(), m, s) : 0) : 0;
  check_fill(p, d->slots, A.nnz, m, "sliced layout");
  d->col.alloc(sizeof(int) * d->slots);
  d->val.alloc(vsize(A.dtype) * d->slots);
  LSP_HIP(hipMemsetAsync(d->col.p, 0xFF, sizeof(int) * d->slots, s));  // pads: col -1
  LSP_HIP(hipMemsetAsync(d->val.p, 0, vsize(A.dtype) * d->slots, s));  // pads: value 0
  if (m) {
    if (A.dtype == LESPMV_F64) fill_dispatch<double>(A, *d, order, s);
    else fill_dispatch<float>(A, *d, order, s);
  }
  LSP_HIP(hipStreamSynchronize(s));
  plan_units(*d, to_host<int>(d->chunk_len.p, d->nchunks, s), vsize(A.dtype), s);
  A.fmt[fmt] = std::move(d);
}

template <class V>
void sliced_run(Matrix& A, SlicedUnits& d, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  auto kern = d.unroll == 6 ? sliced_spmv_kernel<V, 6> : sliced_spmv_kernel<V, 12>;
  hipLaunchKernelGGL(kern, grid1d(d.nunits * d.C, kBlock), dim3(kBlock), 0, s, A.m, d.nunits, d.C,
                     d.units.as<int4>(), d.chunk_off.as<int64_t>(), d.reorder.as<int>(), d.col.as<int>(),
                     d.val.as<V>(), static_cast<const V*>(x), static_cast<V*>(y), d.partial.as<V>(), (V)alpha,
                     (V)beta, (int)((int64_t)d.sigma * kXcds * 8 <= A.m));  // XCD-contiguous rows unless sorted globally
  LSP_LAUNCHED();
  if (d.nsplit) {
    hipLaunchKernelGGL((sliced_fixup<V>), grid1d(d.nsplit * d.C, kBlock), dim3(kBlock), 0, s, A.m, d.nsplit, d.C,
                       d.split.as<int4>(), d.reorder.as<int>(), d.partial.as<V>(), static_cast<V*>(y), (V)alpha,
                       (V)beta);
    LSP_LAUNCHED();
  }
}

void sliced_spmv(Matrix& A, int fmt, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  auto& d = *static_cast<SlicedUnits*>(A.fmt[fmt].get());
  if (A.m == 0) return;
  if (A.dtype == LESPMV_F64) sliced_run<double>(A, d, alpha, x, beta, y, s);
  else sliced_run<float>(A, d, alpha, x, beta, y, s);
}

void sliced_spmm(Matrix& A, int fmt, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
                 int64_t ldy, hipStream_t s) {
  auto& d = *static_cast<SlicedData*>(A.fmt[fmt].get());
  if (A.m == 0 || k == 0) return;
  if (A.dtype == LESPMV_F64) spmm_dispatch<double>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
  else spmm_dispatch<float>(A, d, k, alpha, X, ldx, beta, Y, ldy, s);
}

void sliced_export(Matrix& A, int fmt, const std::string& name, std::vector<unsigned char>& out) {
  auto& d = *static_cast<SlicedData*>(A.fmt[fmt].get());
  hipStream_t s = A.stream;
  const auto clen = to_host<int>(d.chunk_len.p, d.nchunks, s);
  auto off = to_host<int64_t>(d.chunk_off.p, d.nchunks + 1, s);
  if (fmt == LESPMV_ELL) {
    const int64_t W = d.nchunks ? clen[0] : 0, m = A.m;
    const int ld = d.params.ld;
    if (name == "col_index" || name == "values") {
      auto transpose = [&](auto tag) {
        using T = decltype(tag);
        auto dev = to_host<T>(name == "col_index" ? d.col.p : d.val.p, d.slots, s);
        if (ld == 1) { put_bytes(out, dev); return; }
        std::vector<T> rm(dev.size());
        for (int64_t r = 0; r < m; r++)
          for (int64_t j = 0; j < W; j++) rm[r * W + j] = dev[j * m + r];
        put_bytes(out, rm);
      };
      if (name == "col_index") transpose(int32_t{});
      else if (A.dtype == LESPMV_F64) transpose(double{});
      else transpose(float{});
    } else if (name == "max_row_width") put_scalar(out, W);
    else if (name == "ld") put_scalar(out, ld);
    else throw Error(LESPMV_ERR_ARG, "unknown ELL array " + name);
    return;
  }
  auto rowmajor = [&](const DBuf& b, auto tag) {
    using T = decltype(tag);
    put_bytes(out, to_row_major(to_host<T>(b.p, d.slots, s), clen, off, d.C));
  };
  if (name == "col_index") rowmajor(d.col, int32_t{});
  else if (name == "values") {
    if (A.dtype == LESPMV_F64) rowmajor(d.val, double{});
    else rowmajor(d.val, float{});
  } else if (name == "chunk_offset") put_bytes(out, off);
  else if (name == "alignment") put_scalar(out, d.alignment);
  else if (fmt == LESPMV_SELL) {
    if (name == "row_width") put_bytes(out, clen);
    else if (name == "chunk_num") put_scalar(out, d.nchunks);
    else if (name == "sliceWidth") put_scalar(out, d.C);
    else throw Error(LESPMV_ERR_ARG, "unknown S-ELL array " + name);
  } else {
    const int64_t slices = ceil_div(A.m, d.sigma);
    if (name == "reorder") put_bytes(out, to_host<int32_t>(d.reorder.p, A.m, s));
    else if (name == "chunk_len") put_bytes(out, clen);
    else if (name == "chunkWidth_C") put_scalar(out, d.C);
    else if (name == "validchunkNum") put_scalar(out, d.nchunks);
    else if (fmt == LESPMV_SELL_C_SIGMA && name == "sliceWidth_Sigma") put_scalar(out, d.sigma);
    else if (fmt == LESPMV_SELL_C_SIGMA && name == "sliceNum") put_scalar(out, slices);
    else if (fmt == LESPMV_SELL_C_SIGMA && name == "chunk_num_per_slice") put_scalar(out, d.sigma / d.C);
    else if (fmt == LESPMV_SELL_C_SIGMA && name == "chunkNum") put_scalar(out, slices * (d.sigma / d.C));
    else throw Error(LESPMV_ERR_ARG, "unknown SELL array " + name);
  }
}

}  // namespace lsp
