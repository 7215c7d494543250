// Row sums of the sliced layouts (ELL, S-ELL, SELL-C-sigma, SELL-C-R), shared
// by the per-call SELL SpMV (spmv_sliced.hip) and the packed serving kernel
// (serve.hip), so a request served in a batch and the same product launched
// on its own give bitwise equal results.
//
// A chunk of C rows stores slot (r, j) at chunk offset + j*C + r (column-major
// inside the chunk, so the lanes of a wave read consecutive addresses). Pads
// sit at the end of each row with column -1 and value 0.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lsp {

// Slots per chunk that one thread walks alone: the per-call kernel cuts wider
// chunks into units of this many positions and adds their partial sums
// afterwards; the serving kernel takes only layouts within it.
constexpr int kUnitSlots = 64;

// Sum over slots [j0, j1) of one row, cp / vp pointing at the row's slot 0
// (chunk offset + r) and consecutive slots C elements apart; x[j] is read at
// x + j * ldx (a column of a row-major block of vectors). kUnroll slots per
// step; the next step's column and value loads are issued before this step's
// gathers are used, so they overlap. The slots are read with non-temporal
// loads (read once; they should not push x out of the caches). The products
// are added in slot order whatever kUnroll is, and the row ends at its first
// pad.
template <class V, int kUnroll>
__device__ __forceinline__ V sliced_row_sum(const int* __restrict__ cp, const V* __restrict__ vp, int64_t C,
                                            int j0, int j1, const V* __restrict__ x, int64_t ldx = 1) {
  V sum = 0;
  int c[kUnroll];
  V v[kUnroll];
#pragma unroll
  for (int q = 0; q < kUnroll; q++) {
    const bool ok = j0 + q < j1;
    c[q] = ok ? __builtin_nontemporal_load(cp + (int64_t)(j0 + q) * C) : -1;
    v[q] = ok ? __builtin_nontemporal_load(vp + (int64_t)(j0 + q) * C) : V(0);
  }
  for (int j = j0; j < j1; j += kUnroll) {
    V xv[kUnroll];
#pragma unroll
    for (int q = 0; q < kUnroll; q++) xv[q] = c[q] >= 0 ? x[(int64_t)c[q] * ldx] : V(0);
    const bool more = c[kUnroll - 1] >= 0 && j + kUnroll < j1;
    int cn[kUnroll];
    V vn[kUnroll];
#pragma unroll
    for (int q = 0; q < kUnroll; q++) {
      const bool ok = more && j + kUnroll + q < j1;
      cn[q] = ok ? __builtin_nontemporal_load(cp + (int64_t)(j + kUnroll + q) * C) : -1;
      vn[q] = ok ? __builtin_nontemporal_load(vp + (int64_t)(j + kUnroll + q) * C) : V(0);
    }
#pragma unroll
    for (int q = 0; q < kUnroll; q++)
      if (c[q] >= 0) sum += v[q] * xv[q];
    if (!more) break;
#pragma unroll
    for (int q = 0; q < kUnroll; q++) {
      c[q] = cn[q];
      v[q] = vn[q];
    }
  }
  return sum;
}

}  // namespace lsp
