// Driver around the compiled reference (LeSpMV) used to write golden fixtures
// and to cross-check the oracle in this container. TEST INFRASTRUCTURE ONLY.
//
// It is compiled together with the reference's own source files, read where
// they lie (see build_ref.sh); nothing of the reference is copied. It calls the
// reference's conversions (include/sparse_conversion.h) and LeSpMV_<fmt> entry
// points (src/spmv_*.cpp) and returns their arrays in the same named layout as
// the oracle (oracle/lespmv_oracle.cpp), so the two can be compared directly,
// with the reference's own bytes_per_spmv (include/spmv_benchmark.h:26-150)
// of each converted struct as the scalar "bytes_per_spmv".
#include "LeSpMV.h"

#include <set>

#include "named.hpp"

NAMED_ARRAYS_CABI(ref)

namespace {

template <class V>
CSR_Matrix<int, V> wrap(int m, int n, int nnz, const int* rp, const int* ci, const V* v) {
  CSR_Matrix<int, V> a;
  a.num_rows = m;
  a.num_cols = n;
  a.num_nnzs = nnz;
  a.row_offset = copy_array(rp, (size_t)m + 1);
  a.col_index = copy_array(ci, (size_t)nnz);
  a.values = copy_array(v, (size_t)nnz);
  a.kernel_flag = 0;
  a.partition = nullptr;
  return a;
}

template <class M>
void set_partition_csr(M& a) {
  const int t = Le_get_thread_num();
  a.partition = new_array<int>(t + 1);
  balanced_partition_row_by_nnz(a.row_offset, a.num_rows, t, a.partition);
}

template <class V, class Mat>
void put_sliced(NamedArrays& r, int C, int nchunks, const int* len, int* const* col, V* const* val) {
  std::vector<int64_t> off(nchunks + 1, 0);
  for (int c = 0; c < nchunks; c++) off[c + 1] = off[c] + (int64_t)len[c] * C;
  std::vector<int> ci(off[nchunks]);
  std::vector<V> vv(off[nchunks]);
  for (int c = 0; c < nchunks; c++) {
    std::copy(col[c], col[c] + (size_t)len[c] * C, ci.begin() + off[c]);
    std::copy(val[c], val[c] + (size_t)len[c] * C, vv.begin() + off[c]);
  }
  r.put("chunk_offset", off);
  r.put("col_index", ci);
  r.put("values", vv);
}

template <class V>
int run(NamedArrays& r, int fmt, int m, int n, int nnz, const int* rp, const int* ci, const V* v,
        const int64_t* prm, int flag, V alpha, const V* x, V beta, V* y) {
  CSR_Matrix<int, V> csr = wrap(m, n, nnz, rp, ci, v);
  int rc = 0;
  switch (fmt) {
    case 1: {
      r.scalar("bytes_per_spmv", (int64_t)bytes_per_spmv(csr));
      csr.kernel_flag = flag;
      if (flag == 2) set_partition_csr(csr);
      if (y) LeSpMV_csr(alpha, csr, x, beta, y);
      break;
    }
    case 0: {
      auto coo = csr_to_coo(csr);
      r.put("row_index", coo.row_index, nnz);
      r.put("col_index", coo.col_index, nnz);
      r.put("values", coo.values, nnz);
      r.scalar("bytes_per_spmv", (int64_t)bytes_per_spmv(coo));
      coo.kernel_flag = flag;
      if (y) LeSpMV_coo(alpha, coo, x, beta, y);
      delete_coo_matrix(coo);
      break;
    }
    case 2: {
      // csr_to_dia exits when there are too many diagonals; count first.
      std::set<long> diags;
      for (int i = 0; i < m; i++)
        for (int j = rp[i]; j < rp[i + 1]; j++) diags.insert((long)ci[j] - i);
      r.scalar("complete_ndiags", (int64_t)diags.size());
      if ((int64_t)diags.size() > prm[3]) { rc = 4; break; }
      auto dia = csr_to_dia(csr, (int)prm[3], (FILE*)nullptr, (int)prm[2]);
      r.put("diag_offsets", dia.diag_offsets, dia.complete_ndiags);
      r.put("diag_data", dia.diag_data, (size_t)dia.complete_ndiags * dia.stride);
      r.scalar("stride", dia.stride);
      r.scalar("bytes_per_spmv", (int64_t)bytes_per_spmv(dia));
      dia.kernel_flag = flag;
      if (y) LeSpMV_dia(alpha, dia, x, beta, y);
      delete_dia_matrix(dia);
      break;
    }
    case 3: {
      auto ell = csr_to_ell(csr, prm[0] == 1 ? ColMajor : RowMajor);
      r.put("col_index", ell.col_index, (size_t)m * ell.max_row_width);
      r.put("values", ell.values, (size_t)m * ell.max_row_width);
      r.scalar("max_row_width", ell.max_row_width);
      r.scalar("ld", (int)ell.ld);
      ell.kernel_flag = flag;
      ell.num_nnzs = nnz;
      r.scalar("bytes_per_spmv", (int64_t)bytes_per_spmv(ell));
      if (y) LeSpMV_ell(alpha, ell, x, beta, y);
      delete_ell_matrix(ell);
      break;
    }
    case 4: {
      auto s = csr_to_sell(csr, (FILE*)nullptr, (int)prm[0], (int)prm[2]);
      put_sliced<V, int>(r, s.sliceWidth, s.chunk_num, s.row_width, s.col_index, s.values);
      r.put("row_width", s.row_width, s.chunk_num);
      r.scalar("chunk_num", s.chunk_num);
      r.scalar("sliceWidth", s.sliceWidth);
      r.scalar("alignment", s.alignment);
      r.scalar("bytes_per_spmv", (int64_t)bytes_per_spmv(s));
      s.kernel_flag = flag;
      if (y) LeSpMV_sell(alpha, s, x, beta, y);
      delete_s_ell_matrix(s);
      break;
    }
    case 5: {
      if (prm[1] % prm[0]) { rc = 3; break; }
      auto s = csr_to_sell_c_sigma(csr, (FILE*)nullptr, (int)prm[1], (int)prm[0], (int)prm[2]);
      put_sliced<V, int>(r, s.chunkWidth_C, s.validchunkNum, s.chunk_len, s.col_index, s.values);
      r.put("reorder", s.reorder, m);
      r.put("chunk_len", s.chunk_len, s.validchunkNum);
      r.scalar("chunkWidth_C", s.chunkWidth_C);
      r.scalar("sliceWidth_Sigma", s.sliceWidth_Sigma);
      r.scalar("sliceNum", s.sliceNum);
      r.scalar("chunk_num_per_slice", s.chunk_num_per_slice);
      r.scalar("chunkNum", s.chunkNum);
      r.scalar("validchunkNum", s.validchunkNum);
      r.scalar("alignment", s.alignment);
      r.scalar("bytes_per_spmv", (int64_t)bytes_per_spmv(s));
      s.kernel_flag = flag;
      if (y) LeSpMV_sell_c_sigma(alpha, s, x, beta, y);
      delete_s_ell_c_sigma_matrix(s);
      break;
    }
    case 6: {
      auto s = csr_to_sell_c_R(csr, (FILE*)nullptr, (int)prm[0], (int)prm[2]);
      put_sliced<V, int>(r, s.chunkWidth_C, s.validchunkNum, s.chunk_len, s.col_index, s.values);
      r.put("reorder", s.reorder, m);
      r.put("chunk_len", s.chunk_len, s.validchunkNum);
      r.scalar("chunkWidth_C", s.chunkWidth_C);
      r.scalar("validchunkNum", s.validchunkNum);
      r.scalar("alignment", s.alignment);
      r.scalar("bytes_per_spmv", (int64_t)bytes_per_spmv(s));
      s.kernel_flag = flag;
      if (y) LeSpMV_sell_c_R(alpha, s, x, beta, y);
      delete_s_ell_c_R_matrix(s);
      break;
    }
    case 7: {
      if constexpr (std::is_same<V, double>::value) {
        auto c = csr_to_csr5<int, uint32_t, double>(csr, (FILE*)nullptr);
        r.put("row_offset", c.row_offset, m + 1);
        r.put("col_index", c.col_index, nnz);
        r.put("values", c.values, nnz);
        r.put("tile_ptr", c.tile_ptr, c._p + 1);
        r.put("tile_desc", c.tile_desc, (size_t)c._p * c.omega * c.num_packets);
        r.put("tile_desc_offset_ptr", c.tile_desc_offset_ptr, c._p + 1);
        r.put("tile_desc_offset", c.tile_desc_offset, c.num_offsets);
        r.scalar("omega", c.omega);
        r.scalar("sigma", c.sigma);
        r.scalar("bit_y_offset", c.bit_y_offset);
        r.scalar("bit_scansum_offset", c.bit_scansum_offset);
        r.scalar("num_packets", c.num_packets);
        r.scalar("num_offsets", c.num_offsets);
        r.scalar("p", c._p);
        r.scalar("tail_partition_start", c.tail_partition_start);
        r.scalar("bytes_per_spmv", (int64_t)bytes_per_spmv(c));
        if (y) LeSpMV_csr5(alpha, c, x, beta, y);
        delete_csr5_matrix(c);
      } else {
        rc = 5;
      }
      break;
    }
    case 8: {
      auto b = csr_to_bsr(csr, (int)prm[0], (int)prm[1]);
      r.put("row_ptr", b.row_ptr, b.mb + 1);
      r.put("block_colindex", b.block_colindex, b.nnzb);
      r.put("block_data", b.block_data, (size_t)b.nnzb * b.blockDim_r * b.blockDim_c);
      r.scalar("mb", b.mb);
      r.scalar("nb", b.nb);
      r.scalar("nnzb", b.nnzb);
      r.scalar("blockDim_r", b.blockDim_r);
      r.scalar("blockDim_c", b.blockDim_c);
      r.scalar("bytes_per_spmv", (int64_t)bytes_per_spmv(b));
      b.kernel_flag = flag;
      if (flag == 2) {
        const int t = Le_get_thread_num();
        b.partition = new_array<int>(t + 1);
        balanced_partition_row_by_nnz(b.row_ptr, b.mb, t, b.partition);
      }
      if (y) LeSpMV_bsr(alpha, b, x, beta, y);
      delete_bsr_matrix(b);
      break;
    }
    default:
      rc = 2;
  }
  delete_csr_matrix(csr);
  return rc;
}

}  // namespace

extern "C" {

// The reference's CSR5 kernel opens `#pragma omp parallel` without a thread
// count but sizes its calibrator by Le_get_thread_num(); both must agree or
// the kernel writes past the calibrator (src/spmv_csr5.cpp:134-140,
// include/sparse_conversion.h:1024-1030).
void ref_set_threads(int t) {
  Le_set_thread_num(t);
  omp_set_num_threads(t);
}

// fmt ids follow LeSpMV/script/write_label.py:63-75 (1 = CSR).
// params: [0]=C / ld / br, [1]=sigma / bc, [2]=alignment, [3]=max_diags.
// y may be null (conversion only); otherwise y = alpha*A*x + beta*y through
// LeSpMV_<fmt> with the given kernel_flag.
int ref_run(void* res, int dtype, int fmt, int m, int n, int nnz, const int* rp, const int* ci, const void* v,
            const int64_t* params, int flag, double alpha, const void* x, double beta, void* y) {
  auto& r = *static_cast<NamedArrays*>(res);
  if (dtype == 1)
    return run<double>(r, fmt, m, n, nnz, rp, ci, (const double*)v, params, flag, alpha, (const double*)x, beta,
                       (double*)y);
  return run<float>(r, fmt, m, n, nnz, rp, ci, (const float*)v, params, flag, (float)alpha, (const float*)x,
                    (float)beta, (float*)y);
}

// coo_to_csr with the compact flag (include/sparse_conversion.h:29-91).
int ref_coo_to_csr(void* res, int dtype, int m, int n, int nnz, const int* ri, const int* ci, const void* v,
                   int compact) {
  auto& r = *static_cast<NamedArrays*>(res);
  auto go = [&](auto tag) {
    using V = decltype(tag);
    COO_Matrix<int, V> coo;
    coo.num_rows = m;
    coo.num_cols = n;
    coo.num_nnzs = nnz;
    coo.row_index = const_cast<int*>(ri);
    coo.col_index = const_cast<int*>(ci);
    coo.values = const_cast<V*>(static_cast<const V*>(v));
    auto csr = coo_to_csr(coo, compact != 0);
    r.put("row_offset", csr.row_offset, m + 1);
    r.put("col_index", csr.col_index, csr.num_nnzs);
    r.put("values", csr.values, csr.num_nnzs);
    r.scalar("num_nnzs", csr.num_nnzs);
    delete_csr_matrix(csr);
  };
  if (dtype == 1) go(double{}); else go(float{});
  return 0;
}

// read_csr_matrix (utils/sparse_io.cpp:179-204) on a MatrixMarket file.
int ref_read_mtx(void* res, int dtype, const char* path, int compact) {
  auto& r = *static_cast<NamedArrays*>(res);
  auto go = [&](auto tag) {
    using V = decltype(tag);
    auto csr = read_csr_matrix<int, V>(path, compact != 0);
    r.scalar("num_rows", csr.num_rows);
    r.scalar("num_cols", csr.num_cols);
    r.scalar("num_nnzs", csr.num_nnzs);
    r.put("row_offset", csr.row_offset, csr.num_rows + 1);
    r.put("col_index", csr.col_index, csr.num_nnzs);
    r.put("values", csr.values, csr.num_nnzs);
    delete_csr_matrix(csr);
  };
  if (dtype == 1) go(double{}); else go(float{});
  return 0;
}

}  // extern "C"
