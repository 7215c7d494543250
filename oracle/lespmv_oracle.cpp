// CPU oracle for the LeSpMV SpMV hot path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, the smoke
// check and bench.py's cpu_baseline leg load it; the product library never
// links, loads or calls it.
//
// It restates, in plain scalar C++ (+ OpenMP only for the timed baseline),
// LeSpMV's format conversions and serial SpMV kernels. Every function cites
// the reference file:line it follows (paths relative to the reference's
// LeSpMV/ directory). Index type is int32, value type float or double.
//
// Pinning (how each restatement is checked, see tests/test_oracle_*.py):
//  * every conversion and every SpMV is compared bit-exactly / to 1e-15
//    against fixtures written by the compiled reference (tests/golden/);
//  * conversions are additionally checked against invariants (round trip to
//    the COO multiset, pads, alignment, permutation validity);
//  * SpMV against closed forms (Laplacian times constant vector, A*e_j equals
//    column j) and a brute-force dense product on tiny matrices.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "named.hpp"

NAMED_ARRAYS_CABI(oracle)

namespace {

template <class V>
struct Csr {
  int m, n, nnz;
  const int* rp;
  const int* ci;
  const V* v;
};

inline int64_t align_up(int64_t x, int64_t a) { return ((x + a - 1) / a) * a; }

// ---------------------------------------------------------------------------
// coo_to_csr: counting sort by row keeping file order inside a row
// (include/sparse_conversion.h:29-91); with compact, duplicates are summed and
// zero sums dropped by sum_csr_duplicates (include/sparse_operation.h:21-73).
// ---------------------------------------------------------------------------
template <class V>
void sum_csr_duplicates(int num_rows, int num_cols, int* row_ptr, int* col_index, V* values) {
  std::vector<int> next(num_cols, -1);
  std::vector<V> sums(num_cols, V(0));
  int nnz = 0, row_start = 0, row_end = 0;
  for (int i = 0; i < num_rows; i++) {
    int head = -2;
    row_start = row_end;
    row_end = row_ptr[i + 1];
    for (int jj = row_start; jj < row_end; jj++) {
      int j = col_index[jj];
      sums[j] += values[jj];
      if (next[j] == -1) { next[j] = head; head = j; }
    }
    while (head != -2) {
      int curr = head;
      head = next[curr];
      if (sums[curr] != 0) { col_index[nnz] = curr; values[nnz] = sums[curr]; nnz++; }
      next[curr] = -1;
      sums[curr] = 0;
    }
    row_ptr[i + 1] = nnz;
  }
}

template <class V>
void coo_to_csr(NamedArrays& r, int m, int n, int nnz, const int* ri, const int* ci, const V* v, bool compact) {
  std::vector<int> rp(m + 1, 0), col(nnz);
  std::vector<V> val(nnz);
  for (int i = 0; i < nnz; i++) rp[ri[i]]++;
  for (int i = 0, cumsum = 0; i < m; i++) { int t = rp[i]; rp[i] = cumsum; cumsum += t; }
  rp[m] = nnz;
  for (int i = 0; i < nnz; i++) {
    int row = ri[i];
    int dest = rp[row];
    col[dest] = ci[i];
    val[dest] = v[i];
    rp[row]++;
  }
  for (int i = 0, last = 0; i <= m; i++) { int t = rp[i]; rp[i] = last; last = t; }
  int out_nnz = nnz;
  if (compact) {
    sum_csr_duplicates(m, n, rp.data(), col.data(), val.data());
    out_nnz = rp[m];
  }
  r.put("row_offset", rp);
  r.put("col_index", col.data(), out_nnz);
  r.put("values", val.data(), out_nnz);
  r.scalar("num_nnzs", out_nnz);
}

// csr_to_coo: expand rows in row order (include/sparse_conversion.h:144-168).
template <class V>
void csr_to_coo(NamedArrays& r, const Csr<V>& a) {
  std::vector<int> ri(a.nnz);
  for (int row = 0; row < a.m; ++row)
    for (int i = a.rp[row]; i < a.rp[row + 1]; ++i) ri[i] = row;
  r.put("row_index", ri);
  r.put("col_index", a.ci, a.nnz);
  r.put("values", a.v, a.nnz);
}

// csr_to_ell: pad every row to max_row_width with col -1 / value 0, RowMajor
// (ld=0) or ColMajor (ld=1) (include/sparse_conversion.h:186-248). The fill
// count is computed in 64 bits here; the reference computes it as int*int
// (:212-213), which only differs once M*W exceeds 2^31 slots.
template <class V>
void csr_to_ell(NamedArrays& r, const Csr<V>& a, int ld) {
  int w = 0;
  for (int i = 0; i < a.m; i++) w = std::max(w, a.rp[i + 1] - a.rp[i]);
  size_t slots = (size_t)a.m * w;
  std::vector<int> col(slots, -1);
  std::vector<V> val(slots, V(0));
  for (int row = 0; row < a.m; ++row) {
    size_t k = ld == 1 ? (size_t)row : (size_t)row * w;
    for (int j = a.rp[row]; j < a.rp[row + 1]; ++j) {
      col[k] = a.ci[j];
      val[k] = a.v[j];
      k += ld == 1 ? (size_t)a.m : 1;
    }
  }
  r.put("col_index", col);
  r.put("values", val);
  r.scalar("max_row_width", w);
  r.scalar("ld", ld);
}

// Sliced-ELL family. All three reference conversions produce, per chunk of C
// consecutive (possibly reordered) rows, a row-major C x chunk_len block padded
// with col -1 / value 0, chunk_len rounded up to `alignment`.
// The chunk blocks are returned concatenated in chunk order, with "chunk_offset"
// giving each block's first slot.
template <class V>
void fill_sliced(NamedArrays& r, const Csr<V>& a, int C, const std::vector<int>& order,
                 std::vector<int>& chunk_len, int alignment) {
  const int nchunks = (int)chunk_len.size();
  for (int row = 0; row < a.m; ++row) {
    int real = order.empty() ? row : order[row];
    int c = row / C;
    chunk_len[c] = std::max(chunk_len[c], a.rp[real + 1] - a.rp[real]);
  }
  for (int c = 0; c < nchunks; c++) chunk_len[c] = (int)align_up(chunk_len[c], alignment);
  std::vector<int64_t> off(nchunks + 1, 0);
  for (int c = 0; c < nchunks; c++) off[c + 1] = off[c] + (int64_t)chunk_len[c] * C;
  std::vector<int> col(off[nchunks], -1);
  std::vector<V> val(off[nchunks], V(0));
  for (int row = 0; row < a.m; ++row) {
    int real = order.empty() ? row : order[row];
    int c = row / C, rin = row % C;
    int64_t base = off[c] + (int64_t)rin * chunk_len[c];
    for (int j = a.rp[real]; j < a.rp[real + 1]; ++j) {
      col[base + (j - a.rp[real])] = a.ci[j];
      val[base + (j - a.rp[real])] = a.v[j];
    }
  }
  r.put("chunk_offset", off);
  r.put("col_index", col);
  r.put("values", val);
}

// csr_to_sell: S-ELL, chunks of `chunkwidth` rows in natural order
// (include/sparse_conversion.h:257-334).
template <class V>
void csr_to_sell(NamedArrays& r, const Csr<V>& a, int C, int alignment) {
  int nchunks = (a.m + C - 1) / C;
  std::vector<int> width(nchunks, 0);
  fill_sliced(r, a, C, {}, width, alignment);
  r.put("row_width", width);
  r.scalar("chunk_num", nchunks);
  r.scalar("sliceWidth", C);
  r.scalar("alignment", alignment);
}

// Row order of SELL-c-sigma: inside each slice of sigma rows, rows sorted by
// (nnz, row) with std::greater, i.e. nnz descending then row index descending
// (include/sparse_conversion.h:369-401). SELL-c-R is the same with one slice
// covering all rows (:496-518).
std::vector<int> sigma_order(const int* rp, int m, int64_t sigma) {
  std::vector<std::pair<int, int>> cnt(m);
  for (int i = 0; i < m; ++i) cnt[i] = {rp[i + 1] - rp[i], i};
  for (int64_t s = 0; s < m; s += sigma) {
    int64_t e = std::min<int64_t>(s + sigma, m);
    std::sort(cnt.begin() + s, cnt.begin() + e, std::greater<std::pair<int, int>>());
  }
  std::vector<int> order(m);
  for (int i = 0; i < m; ++i) order[i] = cnt[i].second;
  return order;
}

// csr_to_sell_c_sigma (include/sparse_conversion.h:336-479); sigma % C != 0 is
// rejected (the reference exits at :357-361).
template <class V>
int csr_to_sell_c_sigma(NamedArrays& r, const Csr<V>& a, int sigma, int C, int alignment) {
  if (sigma % C) { r.error = "Sigma must be multiple divided by c"; return 3; }
  int slices = (a.m + sigma - 1) / sigma;
  int valid = (a.m + C - 1) / C;
  std::vector<int> order = sigma_order(a.rp, a.m, sigma);
  std::vector<int> len(valid, 0);
  fill_sliced(r, a, C, order, len, alignment);
  r.put("reorder", order);
  r.put("chunk_len", len);
  r.scalar("chunkWidth_C", C);
  r.scalar("sliceWidth_Sigma", sigma);
  r.scalar("sliceNum", slices);
  r.scalar("chunk_num_per_slice", sigma / C);
  r.scalar("chunkNum", (int64_t)slices * (sigma / C));
  r.scalar("validchunkNum", valid);
  r.scalar("alignment", alignment);
  return 0;
}

// csr_to_sell_c_R (include/sparse_conversion.h:481-598).
template <class V>
void csr_to_sell_c_R(NamedArrays& r, const Csr<V>& a, int C, int alignment) {
  int valid = (a.m + C - 1) / C;
  std::vector<int> order = sigma_order(a.rp, a.m, std::max(a.m, 1));
  std::vector<int> len(valid, 0);
  fill_sliced(r, a, C, order, len, alignment);
  r.put("reorder", order);
  r.put("chunk_len", len);
  r.scalar("chunkWidth_C", C);
  r.scalar("validchunkNum", valid);
  r.scalar("alignment", alignment);
}

// csr_to_dia: diagonals numbered in order of first appearance in a row-major
// scan, offsets col-row, data[d*stride + row] with stride = M rounded up to
// `alignment`; later duplicates overwrite (include/sparse_conversion.h:611-766).
// More than max_diags diagonals is an error (the reference exits, :720-735).
template <class V>
int csr_to_dia(NamedArrays& r, const Csr<V>& a, int max_diags, int alignment) {
  std::vector<int> map((size_t)a.m + a.n, -1);
  int nd = 0;
  for (int i = 0; i < a.m; i++)
    for (int jj = a.rp[i]; jj < a.rp[i + 1]; jj++) {
      size_t k = (size_t)(a.m - i) + a.ci[jj];
      if (map[k] == -1) map[k] = nd++;
    }
  r.scalar("complete_ndiags", nd);
  if (nd > max_diags) { r.error = "Number of diagonals excedes limit"; return 4; }
  int64_t stride = align_up(a.m, alignment);
  std::vector<int64_t> offs(nd);
  std::vector<V> data((size_t)nd * stride, V(0));
  for (size_t k = 0; k < (size_t)a.m + a.n; k++)
    if (map[k] != -1) offs[map[k]] = (int64_t)k - a.m;
  for (int i = 0; i < a.m; i++)
    for (int jj = a.rp[i]; jj < a.rp[i + 1]; jj++) {
      size_t k = (size_t)(a.m - i) + a.ci[jj];
      data[(size_t)map[k] * stride + i] = a.v[jj];
    }
  r.put("diag_offsets", offs);
  r.put("diag_data", data);
  r.scalar("stride", stride);
  return 0;
}

// csr_to_bsr (include/sparse_conversion.h:768-926): block rows of br rows,
// block columns in ascending order, blocks stored row-major, later duplicates
// overwrite. The 1x1 case copies CSR unchanged, keeping its column order
// (:784-815). The reference finds the occupied block columns with an
// nb-long flag vector per block row; sorting the distinct block columns gives
// the same ascending list in O(nnz log nnz).
template <class V>
void csr_to_bsr(NamedArrays& r, const Csr<V>& a, int br, int bc) {
  int mb = (a.m + br - 1) / br, nb = (a.n + bc - 1) / bc;
  r.scalar("mb", mb);
  r.scalar("nb", nb);
  r.scalar("blockDim_r", br);
  r.scalar("blockDim_c", bc);
  if (br == 1 && bc == 1) {
    r.put("row_ptr", a.rp, a.m + 1);
    r.put("block_colindex", a.ci, a.nnz);
    r.put("block_data", a.v, a.nnz);
    r.scalar("nnzb", a.nnz);
    return;
  }
  std::vector<int> row_ptr(mb + 1, 0);
  std::vector<std::vector<int>> cols(mb);
  for (int b = 0; b < mb; b++) {
    int s = a.rp[b * br], e = a.rp[std::min(a.m, (b + 1) * br)];
    auto& c = cols[b];
    for (int j = s; j < e; j++) c.push_back(a.ci[j] / bc);
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    row_ptr[b + 1] = row_ptr[b] + (int)c.size();
  }
  int nnzb = row_ptr[mb];
  std::vector<int> bcol(nnzb);
  for (int b = 0; b < mb; b++) std::copy(cols[b].begin(), cols[b].end(), bcol.begin() + row_ptr[b]);
  size_t bs = (size_t)br * bc;
  std::vector<V> data((size_t)nnzb * bs, V(0));
  for (int i = 0; i < a.m; i++) {
    int b = i / br;
    for (int j = a.rp[i]; j < a.rp[i + 1]; j++) {
      int blk = a.ci[j] / bc;
      int k = (int)(std::lower_bound(bcol.begin() + row_ptr[b], bcol.begin() + row_ptr[b + 1], blk) -
                    (bcol.begin() + row_ptr[b]));
      size_t in = (size_t)(a.ci[j] % bc) + (size_t)(i % br) * bc;
      data[(size_t)row_ptr[b] * bs + (size_t)k * bs + in] = a.v[j];
    }
  }
  r.put("row_ptr", row_ptr);
  r.put("block_colindex", bcol);
  r.put("block_data", data);
  r.scalar("nnzb", nnzb);
}

// csr_to_csr5 (include/sparse_conversion.h:938-1379, transpose in
// include/csr5_utils.h:10-87), generalised over omega/sigma. The reference uses
// omega = 64 bytes / sizeof(V) and sigma = 16 with uint32 descriptors; its
// quirks are kept: tile_ptr[p] stays 0, bit 31 of tile_ptr marks tiles with
// empty rows (never for the last tile), the last tile is left untransposed and
// tiles whose raw tile_ptr equals the next one are not transposed either.
template <class V>
int csr_to_csr5(NamedArrays& r, const Csr<V>& a, int omega, int sigma) {
  if (a.nnz == 0) { r.error = "CSR5 needs at least one nonzero"; return 1; }
  int bit_y = 1;
  for (int base = 2; base < sigma * omega; base *= 2) ++bit_y;
  int bit_scan = 1;
  for (int base = 2; base < omega; base *= 2) ++bit_scan;
  if (bit_y + bit_scan > 31) { r.error = "UNSUPPORTED CSR5 OMEGA for bit saving"; return 1; }
  const int bit_all_offset = bit_y + bit_scan;
  const int bit_all = bit_all_offset + sigma;
  const int packets = (bit_all + 31) / 32;
  const int64_t tile = (int64_t)omega * sigma;
  const int p = (int)((a.nnz + tile - 1) / tile);
  const int m = a.m;
  const int* rp = a.rp;

  std::vector<uint32_t> tile_ptr(p + 1, 0);
  std::vector<uint32_t> desc((size_t)p * omega * packets, 0);
  std::vector<int> desc_off_ptr(p + 1, 0);

  // step 1.1: binary search of each tile's first nonzero (:1045-1066)
  for (int g = 0; g < p; g++) {
    int64_t boundary = std::min<int64_t>((int64_t)g * tile, a.nnz);
    int start = 0, stop = m, median;
    while (stop >= start) {
      median = (stop + start) / 2;
      if (boundary >= rp[median]) start = median + 1; else stop = median - 1;
    }
    tile_ptr[g] = (uint32_t)(start - 1);
  }
  // step 1.2: mark tiles containing empty rows (:1068-1094)
  for (int g = 0; g < p; g++) {
    uint32_t start = tile_ptr[g] & 0x7FFFFFFFu, stop = tile_ptr[g + 1] & 0x7FFFFFFFu;
    if (start == stop) continue;
    bool dirty = false;
    for (uint32_t row = start; row <= stop; row++)
      if (rp[row] == rp[row + 1]) { dirty = true; break; }
    if (dirty) tile_ptr[g] = start | 0x80000000u;
  }
  const int tail_start = (int)(tile_ptr[p - 1] & 0x7FFFFFFFu);

  // step 2.1: one bit per row start inside each full tile (:1102-1138)
  for (int t = 0; t < p - 1; t++) {
    int rs = (int)(tile_ptr[t] & 0x7FFFFFFFu), re = (int)(tile_ptr[t + 1] & 0x7FFFFFFFu);
    for (int rid = rs; rid <= re; rid++) {
      int64_t ptr = rp[rid];
      if (ptr / tile != t) continue;
      int lx = (int)((ptr / sigma) % omega);
      int glid = (int)(ptr % sigma) + bit_all_offset;
      int ly = glid / 32, llid = glid % 32;
      desc[(size_t)t * omega * packets + (size_t)ly * omega + lx] |= 1u << (31 - llid);
    }
  }
  // step 2.2: y_offset and scansum_offset in the first packet (:1140-1245)
  std::vector<int> segn(omega + 1), present(omega + 1);
  for (int t = 0; t < p - 1; t++) {
    bool empty_rows = (tile_ptr[t] >> 31) & 1;
    int rs = (int)(tile_ptr[t] & 0x7FFFFFFFu), re = (int)(tile_ptr[t + 1] & 0x7FFFFFFFu);
    if (rs == re) continue;
    std::fill(segn.begin(), segn.end(), 0);
    std::fill(present.begin(), present.end(), 0);
    present[omega] = 1;
    for (int lane = 0; lane < omega; lane++) {
      int start = 0, stop = 0, pres = !lane;
      int ly = 0;
      uint32_t first = desc[(size_t)t * omega * packets + lane];
      uint32_t bitflag = (first << bit_all_offset) | ((uint32_t)pres << 31);
      start = !((bitflag >> 31) & 1);
      pres |= (bitflag >> 31) & 1;
      for (int i = 1; i < sigma; i++) {
        if ((!ly && i == 32 - bit_all_offset) || (ly && (i - (32 - bit_all_offset)) % 32 == 0)) {
          ly++;
          bitflag = desc[(size_t)t * omega * packets + (size_t)ly * omega + lane];
        }
        int norm_i = !ly ? i : i - (32 - bit_all_offset);
        stop += (bitflag >> (31 - norm_i % 32)) & 1;
        pres |= (bitflag >> (31 - norm_i % 32)) & 1;
      }
      int s = stop - start + pres;
      segn[lane] = s > 0 ? s : 0;
      present[lane] = pres;
    }
    // exclusive scan over omega+1 entries (segn[omega] = 0)
    int old_val = segn[0];
    segn[0] = 0;
    for (int i = 1; i < omega + 1; i++) {
      int nv = segn[i];
      segn[i] = old_val + segn[i - 1];
      old_val = nv;
    }
    if (empty_rows) {
      desc_off_ptr[t] = segn[omega];
      desc_off_ptr[p] += segn[omega];
    }
    for (int lane = 0; lane < omega; lane++) {
      uint32_t y_off = (uint32_t)segn[lane];
      uint32_t scan_off = 0;
      int next1 = lane + 1;
      if (present[lane])
        while (!present[next1] && next1 < omega) { scan_off++; next1++; }
      y_off = lane ? y_off - 1 : 0;
      uint32_t& first = desc[(size_t)t * omega * packets + lane];
      first |= y_off << (32 - bit_y);
      first |= scan_off << (32 - bit_all_offset);
    }
  }
  if (desc_off_ptr[p]) {
    int old_val = desc_off_ptr[0];
    desc_off_ptr[0] = 0;
    for (int i = 1; i < p + 1; i++) {
      int nv = desc_off_ptr[i];
      desc_off_ptr[i] = old_val + desc_off_ptr[i - 1];
      old_val = nv;
    }
  }
  const int num_offsets = desc_off_ptr[p];
  std::vector<int> desc_off(num_offsets, 0);
  // tile_desc_offset: row of each row start, relative to row_start+1, for tiles
  // with empty rows (:1250-1361)
  if (num_offsets) {
    const int bit_bitflag = 32 - bit_all_offset;
    for (int t = 0; t < p - 1; t++) {
      if (!((tile_ptr[t] >> 31) & 1)) continue;
      int rs = (int)(tile_ptr[t] & 0x7FFFFFFFu), re = (int)(tile_ptr[t + 1] & 0x7FFFFFFFu);
      int off_ptr = desc_off_ptr[t];
      auto row_of = [&](int64_t idx) {
        int start = 0, stop = re - rs - 1, median;
        while (stop >= start) {
          median = (stop + start) / 2;
          if (idx >= rp[rs + 1 + median]) start = median + 1; else stop = median - 1;
        }
        return start - 1;
      };
      for (int lane = 0; lane < omega; lane++) {
        int ly = 0;
        uint32_t d = desc[(size_t)t * omega * packets + lane];
        int y_off = (int)(d >> (32 - bit_y));
        d = d << bit_all_offset;
        d = lane ? d : (d | 0x80000000u);
        bool bit = (d >> 31) & 1;
        if (bit && lane) {
          desc_off[off_ptr + y_off] = row_of((int64_t)t * tile + (int64_t)lane * sigma);
          y_off++;
        }
        for (int i = 1; i < sigma; i++) {
          if ((!ly && i == bit_bitflag) || (ly && !(31 & (i - bit_bitflag)))) {
            ly++;
            d = desc[(size_t)t * omega * packets + (size_t)ly * omega + lane];
          }
          int norm_i = 31 & (!ly ? i : i - bit_bitflag);
          bit = (d >> (31 - norm_i)) & 1;
          if (bit) {
            desc_off[off_ptr + y_off] = row_of((int64_t)t * tile + (int64_t)lane * sigma + i);
            y_off++;
          }
        }
      }
    }
  }
  // step 3: transpose the first p-1 tiles to (sigma-step, omega-lane) order
  std::vector<int> col(a.ci, a.ci + a.nnz);
  std::vector<V> val(a.v, a.v + a.nnz);
  for (int t = 0; t < p - 1; t++) {
    if (tile_ptr[t] == tile_ptr[t + 1]) continue;
    size_t base = (size_t)t * tile;
    std::vector<int> sc(tile);
    std::vector<V> sv(tile);
    for (int idx = 0; idx < tile; idx++) {
      int iy = idx % sigma, ix = idx / sigma;
      sc[(size_t)iy * omega + ix] = col[base + idx];
      sv[(size_t)iy * omega + ix] = val[base + idx];
    }
    std::copy(sc.begin(), sc.end(), col.begin() + base);
    std::copy(sv.begin(), sv.end(), val.begin() + base);
  }
  r.put("row_offset", a.rp, m + 1);
  r.put("col_index", col);
  r.put("values", val);
  r.put("tile_ptr", tile_ptr);
  r.put("tile_desc", desc);
  r.put("tile_desc_offset_ptr", desc_off_ptr);
  r.put("tile_desc_offset", desc_off);
  r.scalar("omega", omega);
  r.scalar("sigma", sigma);
  r.scalar("bit_y_offset", bit_y);
  r.scalar("bit_scansum_offset", bit_scan);
  r.scalar("num_packets", packets);
  r.scalar("num_offsets", num_offsets);
  r.scalar("p", p);
  r.scalar("tail_partition_start", tail_start);
  return 0;
}

// ---------------------------------------------------------------------------
// Serial SpMV kernels, y = alpha*A*x + beta*y
// ---------------------------------------------------------------------------

// __spmv_csr_serial_simple (src/spmv_csr.cpp:64-90)
template <class V>
void spmv_csr_serial(const Csr<V>& a, V alpha, const V* x, V beta, V* y) {
  for (int row = 0; row < a.m; ++row) {
    V sum = 0;
    for (int jj = a.rp[row]; jj < a.rp[row + 1]; ++jj) sum += a.v[jj] * x[a.ci[jj]];
    if (alpha == 1 && beta == 0) y[row] = sum;
    else y[row] = alpha * sum + beta * y[row];
  }
}

// balanced_partition_row_by_nnz + lower_bound_int (utils/partition.cpp:27-75)
void balanced_partition_row_by_nnz(const int* row_ptr, int rows, int threads, int* part) {
  const int* acc = row_ptr + 1;
  int nnz = acc[rows - 1];
  int ave = nnz / threads;
  if (ave == 0) ave = 1;
  part[0] = 0;
  for (int i = 1; i < threads; i++) {
    int l = 0, rr = rows - 1, value = ave * i;
    while (rr > l) {
      int mid = (l + rr) / 2;
      if (acc[mid] < value) l = mid + 1; else rr = mid;
    }
    part[i] = l;
  }
  part[threads] = rows;
}

// __spmv_csr_omp_lb + __spmv_csr_perthread (src/spmv_csr.cpp:33-62,137-162):
// the reference's kernel_flag=2 path, the timed CPU baseline.
template <class V>
void spmv_csr_omp_lb(const Csr<V>& a, V alpha, const V* x, V beta, V* y, const int* part, int threads) {
#pragma omp parallel num_threads(threads)
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    for (int row = part[tid]; row < part[tid + 1]; row++) {
      V sum = 0;
      const int s = a.rp[row], e = a.rp[row + 1];
#pragma omp simd reduction(+ : sum)
      for (int c = s; c < e; ++c) sum += a.v[c] * x[a.ci[c]];
      if (alpha == 1 && beta == 0) y[row] = sum;
      else y[row] = alpha * sum + beta * y[row];
    }
  }
}

// __spmv_coo_serial_simple (src/spmv_coo.cpp:18-34)
template <class V>
void spmv_coo_serial(int m, int nnz, const int* ri, const int* ci, const V* v, V alpha, const V* x, V beta, V* y) {
  for (int i = 0; i < m; ++i) y[i] *= beta;
  for (int i = 0; i < nnz; ++i) y[ri[i]] += alpha * v[i] * x[ci[i]];
}

// __spmv_ell_serial_simple (src/spmv_ell.cpp:16-67)
template <class V>
void spmv_ell_serial(int m, int w, const int* col, const V* val, int ld, V alpha, const V* x, V beta, V* y) {
  for (int row = 0; row < m; ++row) {
    V sum = 0;
    for (int k = 0; k < w; k++) {
      size_t pos = ld == 1 ? (size_t)row + (size_t)k * m : (size_t)row * w + k;
      int c = col[pos];
      if (c >= 0) sum += val[pos] * x[c];
    }
    y[row] = alpha * sum + beta * y[row];
  }
}

// __spmv_sell_serial_simple / __spmv_sell_cs_serial_simple
// (src/spmv_s_ell.cpp:17-55, src/spmv_sell_c_sigma.cpp:17-54,
// src/spmv_sell_c_R.cpp:17-54). Padded slots (col -1) are skipped; the
// reference tests `size_t col >= 0`, which is always true, and so multiplies
// the zero pad by x[(size_t)-1].
template <class V>
void spmv_sliced_serial(int m, int C, int nchunks, const int64_t* chunk_off, const int* width,
                        const int* col, const V* val, const int* reorder, V alpha, const V* x, V beta, V* y) {
  for (int c = 0; c < nchunks; ++c) {
    for (int rin = 0; rin < C; ++rin) {
      int64_t row = (int64_t)c * C + rin;
      if (row >= m) break;
      int out = reorder ? reorder[row] : (int)row;
      V sum = 0;
      for (int k = 0; k < width[c]; k++) {
        int64_t pos = chunk_off[c] + (int64_t)rin * width[c] + k;
        if (col[pos] >= 0) sum += val[pos] * x[col[pos]];
      }
      y[out] = alpha * sum + beta * y[out];
    }
  }
}

// __spmv_dia_serial_simple (src/spmv_dia.cpp:15-43), with the column bound of
// __spmv_dia_alpha (:100-174): a diagonal with offset d covers rows
// max(0,-d) .. min(M, N-d). The serial kernel bounds the row range by
// num_rows on both sides, which equals this for square matrices.
template <class V>
void spmv_dia_serial(int m, int n, int nd, int64_t stride, const int64_t* offs, const V* data, V alpha, const V* x,
                     V beta, V* y) {
  for (int i = 0; i < m; ++i) y[i] = beta * y[i];
  for (int d = 0; d < nd; ++d) {
    int64_t off = offs[d];
    int64_t s = std::max<int64_t>(0, -off), e = std::min<int64_t>(m, (int64_t)n - off);
    for (int64_t i = s; i < e; ++i) y[i] += alpha * data[(size_t)d * stride + i] * x[i + off];
  }
}

// __spmv_bsr_serial_simple (src/spmv_bsr.cpp:17-66). Columns past N inside
// the last block column hold zero padding and are skipped; the reference
// multiplies them with x[j >= N], reading past the end of x.
template <class V>
void spmv_bsr_serial(int m, int n, int br, int bc, int mb, const int* rptr, const int* bcol, const V* data, V alpha,
                     const V* x, V beta, V* y) {
  std::vector<V> tmp(br);
  for (int i = 0; i < mb; i++) {
    std::fill(tmp.begin(), tmp.end(), V(0));
    for (int j = rptr[i]; j < rptr[i + 1]; j++)
      for (int r = 0; r < br; ++r)
        for (int c = 0; c < bc; ++c) {
          const size_t xc = (size_t)bcol[j] * bc + c;
          if (xc < (size_t)n) tmp[r] += data[(size_t)j * br * bc + (size_t)r * bc + c] * x[xc];
        }
    for (int r = 0; r < br; r++) {
      int64_t yi = (int64_t)i * br + r;
      if (yi < m) y[yi] = alpha * tmp[r] + beta * y[yi];
    }
  }
}

// LeSpMV_csr5 (src/spmv_csr5.cpp:112-457) restated lane by lane for one thread:
// the compute kernel over tiles 0..p-2, then the tail kernel. With one thread
// the calibrator path reduces to direct accumulation. Like the reference,
// rows whose segment is completed inside a tile are stored without beta*y
// (:287-288), so the result equals BLAS semantics only for y0 = 0 or beta = 0.
// Requires the descriptor to fit one 32-bit packet (sigma + bits <= 32), as
// the reference's compiled path (CSR5_SIGMA <= 20) does.
template <class V>
int spmv_csr5_serial(const NamedArrays& h, int m, V alpha, const V* x, V beta, V* y) {
  const int omega = (int)h.get_scalar("omega"), sigma = (int)h.get_scalar("sigma");
  const int bit_y = (int)h.get_scalar("bit_y_offset"), bit_scan = (int)h.get_scalar("bit_scansum_offset");
  const int packets = (int)h.get_scalar("num_packets"), p = (int)h.get_scalar("p");
  const int tail_start = (int)h.get_scalar("tail_partition_start");
  if (packets != 1 || sigma + bit_y + bit_scan > 32) return 1;
  const int* rp = h.get<int>("row_offset");
  const int* col = h.get<int>("col_index");
  const V* val = h.get<V>("values");
  const uint32_t* tp = h.get<uint32_t>("tile_ptr");
  const uint32_t* desc = h.get<uint32_t>("tile_desc");
  const int* dop = h.get<int>("tile_desc_offset_ptr");
  const int* dof = h.get<int>("tile_desc_offset");
  const int64_t tile = (int64_t)omega * sigma;
  std::vector<V> sum(omega), first_sum(omega), last_sum(omega), sh(omega), scan(omega);
  std::vector<uint32_t> d(omega), y_off(omega), s_off(omega);
  std::vector<int> direct(omega), start(omega), stop(omega);
  for (int t = 0; t < p - 1; t++) {
    const V* tv = val + (size_t)t * tile;
    const int* tc = col + (size_t)t * tile;
    uint32_t row_start = tp[t];
    const uint32_t row_stop = tp[t + 1] & 0x7FFFFFFFu;
    if (row_start == row_stop) {  // partition_fast_track (:57-109)
      bool fast_direct = (desc[(size_t)t * omega * packets] >> (31 - (bit_y + bit_scan))) & 1;
      V s = 0;
      for (int i = 0; i < tile; i++) s += tv[i] * x[tc[i]];
      s = s * alpha;
      if (fast_direct) y[row_start] = beta * y[row_start] + s; else y[row_start] += s;
      continue;
    }
    const bool empty_rows = (row_start >> 31) & 1;
    row_start &= 0x7FFFFFFFu;
    V* y_local = y + row_start + 1;
    const int off_ptr = empty_rows ? dop[t] : 0;
    for (int l = 0; l < omega; l++) {
      uint32_t dd = desc[(size_t)t * omega * packets + l];
      y_off[l] = dd >> (32 - bit_y);
      s_off[l] = (dd << bit_y) >> (32 - bit_scan);
      d[l] = dd << (bit_y + bit_scan);
    }
    const bool first_direct = d[0] >> 31;
    for (int l = 0; l < omega; l++) {
      int bit = l == 0 ? 1 : (int)(d[l] >> 31);
      start[l] = bit ? 0 : 1;
      direct[l] = bit && l != 0;
      stop[l] = 0;
      first_sum[l] = 0;
      sum[l] = tv[l] * (x[tc[l]] * alpha);
    }
    for (int i = 1; i < sigma; i++) {
      for (int l = 0; l < omega; l++) {
        int bit = (d[l] >> (31 - i)) & 1;
        if (bit) {
          if (direct[l]) {
            int yi = empty_rows ? dof[off_ptr + y_off[l]] : (int)y_off[l];
            y_local[yi] = sum[l];
            y_off[l]++;
          } else {
            first_sum[l] = sum[l];
          }
          sum[l] = 0;
          direct[l] = 1;
        }
        if (direct[l]) stop[l]++;
        sum[l] += tv[(size_t)i * omega + l] * (x[tc[(size_t)i * omega + l]] * alpha);
      }
    }
    for (int l = 0; l < omega; l++) {
      if (!direct[l]) first_sum[l] = sum[l];
      last_sum[l] = sum[l];
    }
    // carry of each lane: first segments of the next s_off+1 lanes (:310-331)
    for (int l = 0; l < omega; l++) sh[l] = (l + 1 < omega && start[l + 1] == 1) ? first_sum[l + 1] : V(0);
    V acc = 0;
    for (int l = 0; l < omega; l++) { acc += sh[l]; scan[l] = acc; }
    for (int l = 0; l < omega; l++) {
      int hi = l + (int)s_off[l];
      V carry = scan[hi] - scan[l] + sh[l];
      if (start[l] <= stop[l]) last_sum[l] += carry;
    }
    for (int l = 0; l < omega; l++)
      if (direct[l]) {
        int yi = empty_rows ? dof[off_ptr + y_off[l]] : (int)y_off[l];
        y_local[yi] = last_sum[l];
      }
    V s0 = direct[0] ? first_sum[0] : last_sum[0];
    if (first_direct) y[row_start] = beta * y[row_start] + s0; else y[row_start] += s0;
  }
  // spmv_csr5_tail_partition_kernel (:386-424)
  const int64_t first_tail = (int64_t)(p - 1) * tile;
  for (int row = tail_start; row < m; row++) {
    int64_t s = row == tail_start ? first_tail : rp[row];
    V acc = 0;
    for (int64_t i = s; i < rp[row + 1]; i++) acc += val[i] * x[col[i]] * alpha;
    if (row == tail_start && rp[row] != first_tail) y[row] = y[row] + acc;
    else y[row] = y[row] * beta + acc;
  }
  return 0;
}

template <class V>
Csr<V> mk(int m, int n, int nnz, const int* rp, const int* ci, const void* v) {
  return Csr<V>{m, n, nnz, rp, ci, static_cast<const V*>(v)};
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI (ctypes). dtype: 0 = float, 1 = double. Return 0 on success.
// ---------------------------------------------------------------------------
#define DISPATCH(dtype, CALL) ((dtype) == 1 ? CALL(double) : CALL(float))

extern "C" {

int oracle_coo_to_csr(void* res, int dtype, int m, int n, int nnz, const int* ri, const int* ci, const void* v,
                      int compact) {
  auto& r = *static_cast<NamedArrays*>(res);
#define C(V) (coo_to_csr<V>(r, m, n, nnz, ri, ci, static_cast<const V*>(v), compact != 0), 0)
  return DISPATCH(dtype, C);
#undef C
}

int oracle_convert(void* res, int dtype, int fmt, int m, int n, int nnz, const int* rp, const int* ci, const void* v,
                   const int64_t* params) {
  auto& r = *static_cast<NamedArrays*>(res);
  // fmt ids follow LeSpMV/script/write_label.py:63-75
  // params: [0]=C or ld or br or omega, [1]=sigma or bc, [2]=alignment, [3]=max_diags
  auto convert_impl_d = [&](auto tag) -> int {
    using V = decltype(tag);
    auto a = mk<V>(m, n, nnz, rp, ci, v);
    switch (fmt) {
      case 0: csr_to_coo<V>(r, a); return 0;
      case 2: return csr_to_dia<V>(r, a, (int)params[3], (int)params[2]);
      case 3: csr_to_ell<V>(r, a, (int)params[0]); return 0;
      case 4: csr_to_sell<V>(r, a, (int)params[0], (int)params[2]); return 0;
      case 5: return csr_to_sell_c_sigma<V>(r, a, (int)params[1], (int)params[0], (int)params[2]);
      case 6: csr_to_sell_c_R<V>(r, a, (int)params[0], (int)params[2]); return 0;
      case 7: return csr_to_csr5<V>(r, a, (int)params[0], (int)params[1]);
      case 8: csr_to_bsr<V>(r, a, (int)params[0], (int)params[1]); return 0;
      default: r.error = "unknown format"; return 2;
    }
  };
  return dtype == 1 ? convert_impl_d(double{}) : convert_impl_d(float{});
}

int oracle_spmv_csr(int dtype, int flag, int m, int n, int nnz, const int* rp, const int* ci, const void* v,
                    double alpha, const void* x, double beta, void* y, int threads) {
  auto run = [&](auto tag) -> int {
    using V = decltype(tag);
    auto a = mk<V>(m, n, nnz, rp, ci, v);
    if (flag == 0 || threads <= 1 || m == 0) {
      spmv_csr_serial<V>(a, (V)alpha, (const V*)x, (V)beta, (V*)y);
    } else {
      std::vector<int> part(threads + 1);
      balanced_partition_row_by_nnz(rp, m, threads, part.data());
      spmv_csr_omp_lb<V>(a, (V)alpha, (const V*)x, (V)beta, (V*)y, part.data(), threads);
    }
    return 0;
  };
  return dtype == 1 ? run(double{}) : run(float{});
}

// Time `iters` back-to-back kernel_flag=2 CSR SpMVs (alpha=1, beta=0) on
// `threads` OpenMP threads with a steady clock; returns ms per SpMV.
double oracle_time_csr_lb(int dtype, int m, int n, int nnz, const int* rp, const int* ci, const void* v,
                          const void* x, void* y, int threads, int iters) {
  double ms = -1;
  auto run = [&](auto tag) {
    using V = decltype(tag);
    auto a = mk<V>(m, n, nnz, rp, ci, v);
    std::vector<int> part(threads + 1);
    balanced_partition_row_by_nnz(rp, m, threads, part.data());
    double t0 = omp_get_wtime();
    for (int it = 0; it < iters; it++) spmv_csr_omp_lb<V>(a, (V)1, (const V*)x, (V)0, (V*)y, part.data(), threads);
    ms = (omp_get_wtime() - t0) * 1e3 / iters;
  };
  if (dtype == 1) run(double{}); else run(float{});
  return ms;
}

// SpMV of a format held in a result bag produced by oracle_convert.
int oracle_spmv_fmt(void* res, int dtype, int fmt, int m, int n, int nnz, double alpha, const void* x, double beta,
                    void* y) {
  const auto& h = *static_cast<NamedArrays*>(res);
  auto run = [&](auto tag) -> int {
    using V = decltype(tag);
    const V* xv = (const V*)x;
    V* yv = (V*)y;
    switch (fmt) {
      case 0:
        spmv_coo_serial<V>(m, nnz, h.get<int>("row_index"), h.get<int>("col_index"), h.get<V>("values"), (V)alpha,
                           xv, (V)beta, yv);
        return 0;
      case 2:
        spmv_dia_serial<V>(m, n, (int)h.get_scalar("complete_ndiags"), h.get_scalar("stride"),
                           h.get<int64_t>("diag_offsets"), h.get<V>("diag_data"), (V)alpha, xv, (V)beta, yv);
        return 0;
      case 3:
        spmv_ell_serial<V>(m, (int)h.get_scalar("max_row_width"), h.get<int>("col_index"), h.get<V>("values"),
                           (int)h.get_scalar("ld"), (V)alpha, xv, (V)beta, yv);
        return 0;
      case 4:
        spmv_sliced_serial<V>(m, (int)h.get_scalar("sliceWidth"), (int)h.get_scalar("chunk_num"),
                              h.get<int64_t>("chunk_offset"), h.get<int>("row_width"), h.get<int>("col_index"),
                              h.get<V>("values"), nullptr, (V)alpha, xv, (V)beta, yv);
        return 0;
      case 5:
      case 6: {
        int valid = (int)h.get_scalar("validchunkNum");
        int C = (int)h.get_scalar("chunkWidth_C");
        spmv_sliced_serial<V>(m, C, valid, h.get<int64_t>("chunk_offset"), h.get<int>("chunk_len"),
                              h.get<int>("col_index"), h.get<V>("values"),
                              h.get<int>("reorder"), (V)alpha, xv, (V)beta, yv);
        return 0;
      }
      case 7:
        return spmv_csr5_serial<V>(h, m, (V)alpha, xv, (V)beta, yv);
      case 8:
        spmv_bsr_serial<V>(m, n, (int)h.get_scalar("blockDim_r"), (int)h.get_scalar("blockDim_c"),
                           (int)h.get_scalar("mb"), h.get<int>("row_ptr"), h.get<int>("block_colindex"),
                           h.get<V>("block_data"), (V)alpha, xv, (V)beta, yv);
        return 0;
      default:
        return 2;
    }
  };
  return dtype == 1 ? run(double{}) : run(float{});
}

// maximum_relative_error (include/sparse_operation.h:75-94)
double oracle_max_rel_error(int dtype, const void* a, const void* b, int64_t n) {
  auto run = [&](auto tag) -> double {
    using T = decltype(tag);
    const T* A = (const T*)a;
    const T* B = (const T*)b;
    T max_error = 0, eps = std::sqrt(std::numeric_limits<T>::epsilon());
    for (int64_t i = 0; i < n; i++) {
      T err = std::abs(A[i] - B[i]);
      if (err != 0) max_error = std::max(max_error, err / (std::abs(A[i]) + std::abs(B[i]) + eps));
    }
    return (double)max_error;
  };
  return dtype == 1 ? run(double{}) : run(float{});
}

int oracle_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
