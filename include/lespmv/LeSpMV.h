// LeSpMV.h - C++ drop-in for the reference's SpMV interface, on the MI355X
// library (include/lespmv_hip.h).
//
// Source-compatible with the reference's umbrella header (LeSpMV/include/
// LeSpMV.h:1-30) for the hot path: the same struct names and fields
// (include/sparse_format.h:49-295), the same conversion functions and
// parameters (include/sparse_conversion.h:29-1379), the same entry points
// LeSpMV_<fmt>(alpha, A, x, beta, y) (include/spmv_*.h), read_csr_matrix /
// read_coo_matrix (include/sparse_io.h) and the test routines
// test_<fmt>_matrix_kernels (include/spmv_testroutine.h:7-50).
//
// Differences a caller sees:
//  * every conversion and every SpMV runs on the GPU; host struct fields are
//    filled from the device arrays in the reference's layout, so code reading
//    them keeps working;
//  * LeSpMV_<fmt> with host x/y stages them through device buffers; the
//    *_device overloads take device pointers and do no copies;
//  * kernel_flag / kernel_tag select the GPU kernel variant
//    (0 auto, 1 vector, 2 merge / LDS window, 3 COO atomics, 4 / 5 / 6 CSR
//    column panels: 8 / 16 / 4);
//  * errors throw lespmv::Error instead of calling exit();
//  * IndexType may be int or long long; long long matrices keep 32-bit row
//    and column indices on the device, and beyond 2^31 - 1 nonzeros 64-bit
//    row offsets, which run the CSR kernels only.
#pragma once

#include <malloc.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../lespmv_hip.h"

namespace lespmv {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const char* what) {
  if (rc != LESPMV_OK) throw Error(rc, std::string(what) + ": " + lespmv_last_error());
}

template <class V>
constexpr int dtype_of() {
  static_assert(std::is_same<V, double>::value || std::is_same<V, float>::value, "ValueType must be float/double");
  return std::is_same<V, double>::value ? LESPMV_F64 : LESPMV_F32;
}

// Owner of one device matrix handle.
struct Handle {
  lespmv_matrix_t h = nullptr;
  Handle() = default;
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  ~Handle() {
    if (h) lespmv_destroy(h);
  }
};

// Reusable device staging buffer for host vectors.
struct DeviceBuffer {
  void* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t b) {
    if (b <= bytes) return;
    if (p) lespmv_free(p);
    p = nullptr;
    check(lespmv_malloc(&p, (int64_t)b), "lespmv_malloc");
    bytes = b;
  }
  ~DeviceBuffer() {
    if (p) lespmv_free(p);
  }
};

inline DeviceBuffer& staging(int which) {
  static thread_local DeviceBuffer buf[2];
  return buf[which];
}

template <class T>
T* new_array(size_t n) {
  return static_cast<T*>(memalign(64, std::max<size_t>(n, 1) * sizeof(T)));  // memopt.h:38-45
}
template <class T>
void delete_array(T* p) {
  free(p);
}

}  // namespace lespmv

using lespmv::delete_array;
using lespmv::new_array;

typedef enum { RowMajor = 0, ColMajor = 1 } LeadingDimension;

// ---- format structs (include/sparse_format.h:49-295) -------------------------

template <typename IndexType>
struct Matrix_Features {
  typedef IndexType index_type;
  IndexType num_rows = 0, num_cols = 0, num_nnzs = 0;
  double sparsity = 0;
  IndexType* partition = nullptr;
  double time = 0, gflops = 0, gbytes = 0;
  int tag = 0;
  int kernel_flag = 0;
  // device side: the handle holding this matrix (and its converted format)
  std::shared_ptr<lespmv::Handle> gpu;
  int gpu_format = LESPMV_CSR;
};

template <typename IndexType, typename ValueType>
struct COO_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType* row_index = nullptr;
  IndexType* col_index = nullptr;
  ValueType* values = nullptr;
};

template <typename IndexType, typename ValueType>
struct CSR_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType* row_offset = nullptr;
  IndexType* col_index = nullptr;
  ValueType* values = nullptr;
};

template <typename IndexType, typename ValueType>
struct BSR_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType blockDim_r = 0, blockDim_c = 0, blockNNZ = 0;
  IndexType mb = 0, nb = 0, nnzb = 0;
  IndexType* row_ptr = nullptr;
  IndexType* block_colindex = nullptr;
  ValueType* block_data = nullptr;
};

template <typename IndexType, typename UIndexType, typename ValueType>
struct CSR5_Matrix : public CSR_Matrix<IndexType, ValueType> {
  IndexType sigma = 0, omega = 0;
  IndexType bit_y_offset = 0, bit_scansum_offset = 0, num_packets = 0, num_offsets = 0, _p = 0;
  IndexType tail_partition_start = 0;
  UIndexType* tile_ptr = nullptr;
  UIndexType* tile_desc = nullptr;
  IndexType* tile_desc_offset_ptr = nullptr;
  IndexType* tile_desc_offset = nullptr;
  ValueType* calibrator = nullptr;
};

template <typename IndexType, typename ValueType>
struct DIA_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType stride = 0;
  IndexType complete_ndiags = 0;
  long int* diag_offsets = nullptr;
  ValueType* diag_data = nullptr;
};

template <typename IndexType, typename ValueType>
struct ELL_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  LeadingDimension ld = RowMajor;
  IndexType max_row_width = 0, min_row_width = 0;
  IndexType* col_index = nullptr;
  ValueType* values = nullptr;
};

template <typename IndexType, typename ValueType>
struct S_ELL_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType sliceWidth = 0, chunk_num = 0, alignment = 0;
  IndexType* row_width = nullptr;
  IndexType** col_index = nullptr;
  ValueType** values = nullptr;
};

template <typename IndexType, typename ValueType>
struct SELL_C_Sigma_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType sliceWidth_Sigma = 0, chunkWidth_C = 0, sliceNum = 0, chunkNum = 0, validchunkNum = 0;
  IndexType chunk_num_per_slice = 0, alignment = 0;
  IndexType* reorder = nullptr;
  IndexType* chunk_len = nullptr;
  IndexType** col_index = nullptr;
  ValueType** values = nullptr;
};

template <typename IndexType, typename ValueType>
struct SELL_C_R_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType chunkWidth_C = 0, validchunkNum = 0, alignment = 0;
  IndexType* reorder = nullptr;
  IndexType* chunk_len = nullptr;
  IndexType** col_index = nullptr;
  ValueType** values = nullptr;
};

namespace lespmv {

// Host CSR -> a new device handle (32-bit indices on the device).
template <class I, class V>
std::shared_ptr<Handle> upload(const CSR_Matrix<I, V>& csr) {
  auto h = std::make_shared<Handle>();
  const int64_t m = csr.num_rows, n = csr.num_cols, nnz = csr.num_nnzs;
  if (std::is_same<I, int>::value) {
    check(lespmv_create_csr(&h->h, dtype_of<V>(), m, n, nnz, reinterpret_cast<const int32_t*>(csr.row_offset),
                            reinterpret_cast<const int32_t*>(csr.col_index), csr.values, 0),
          "lespmv_create_csr");
  } else {
    if (m >= INT32_MAX || n >= INT32_MAX)
      throw Error(LESPMV_ERR_UNSUPPORTED, "rows and columns must fit 32-bit device indices");
    std::vector<int32_t> ci(csr.col_index, csr.col_index + nnz);
    if (nnz >= INT32_MAX) {  // 64-bit row offsets on the device: CSR SpMV / SpMM only
      if (sizeof(I) != sizeof(int64_t)) throw Error(LESPMV_ERR_UNSUPPORTED, "64-bit row offsets need a 64-bit IndexType");
      check(lespmv_create_csr64(&h->h, dtype_of<V>(), m, n, nnz, reinterpret_cast<const int64_t*>(csr.row_offset),
                                ci.data(), csr.values, 0),
            "lespmv_create_csr64");
    } else {
      std::vector<int32_t> rp(csr.row_offset, csr.row_offset + m + 1);
      check(lespmv_create_csr(&h->h, dtype_of<V>(), m, n, nnz, rp.data(), ci.data(), csr.values, 0),
            "lespmv_create_csr");
    }
  }
  return h;
}

template <class I, class V>
lespmv_matrix_t csr_handle(const CSR_Matrix<I, V>& csr) {
  auto& self = const_cast<CSR_Matrix<I, V>&>(csr);
  if (!self.gpu) self.gpu = upload(csr);
  return self.gpu->h;
}

inline std::vector<unsigned char> fetch(lespmv_matrix_t h, int fmt, const char* name) {
  int64_t bytes = 0;
  check(lespmv_export_size(h, fmt, name, &bytes), name);
  std::vector<unsigned char> buf(bytes);
  check(lespmv_export(h, fmt, name, buf.data(), bytes), name);
  return buf;
}

template <class T, class S>
T* fetch_array(lespmv_matrix_t h, int fmt, const char* name, size_t* count = nullptr) {
  auto buf = fetch(h, fmt, name);
  const size_t n = buf.size() / sizeof(S);
  T* out = new_array<T>(n);
  const S* src = reinterpret_cast<const S*>(buf.data());
  for (size_t i = 0; i < n; i++) out[i] = (T)src[i];
  if (count) *count = n;
  return out;
}

inline int64_t fetch_scalar(lespmv_matrix_t h, int fmt, const char* name) {
  auto buf = fetch(h, fmt, name);
  int64_t v;
  std::memcpy(&v, buf.data(), sizeof v);
  return v;
}

// Convert `csr` to `fmt` on a fresh device handle and record it in `out`.
template <class I, class V, class Out>
lespmv_matrix_t convert_on_device(const CSR_Matrix<I, V>& csr, int fmt, const lespmv_params_t& p, Out& out) {
  auto h = upload(csr);
  check(lespmv_convert(h->h, fmt, &p), "lespmv_convert");
  out.num_rows = csr.num_rows;
  out.num_cols = csr.num_cols;
  out.num_nnzs = csr.num_nnzs;
  out.tag = 0;
  out.kernel_flag = csr.kernel_flag;
  out.gpu = h;
  out.gpu_format = fmt;
  return h->h;
}

inline lespmv_params_t params(int fmt, int dtype) {
  lespmv_params_t p;
  check(lespmv_default_params(fmt, dtype, 1, &p), "lespmv_default_params");
  return p;
}

// Split the reference-layout concatenated chunk blocks into per-chunk arrays.
template <class T, class S, class I>
T** split_chunks(lespmv_matrix_t h, int fmt, const char* name, const int64_t* off, I nchunks) {
  auto buf = fetch(h, fmt, name);
  const S* src = reinterpret_cast<const S*>(buf.data());
  T** out = new T*[nchunks];
  for (I c = 0; c < nchunks; c++) {
    const int64_t n = off[c + 1] - off[c];
    out[c] = new_array<T>(n);
    for (int64_t k = 0; k < n; k++) out[c][k] = (T)src[off[c] + k];
  }
  return out;
}

template <class V>
double sqrt_eps() {
  return std::sqrt((double)std::numeric_limits<V>::epsilon());
}

}  // namespace lespmv

// ---- conversions (include/sparse_conversion.h) -------------------------------

template <class IndexType, class ValueType>
COO_Matrix<IndexType, ValueType> csr_to_coo(const CSR_Matrix<IndexType, ValueType>& csr) {
  COO_Matrix<IndexType, ValueType> coo;
  auto p = lespmv::params(LESPMV_COO, lespmv::dtype_of<ValueType>());
  auto h = lespmv::convert_on_device(csr, LESPMV_COO, p, coo);
  coo.row_index = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_COO, "row_index");
  coo.col_index = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_COO, "col_index");
  coo.values = lespmv::fetch_array<ValueType, ValueType>(h, LESPMV_COO, "values");
  return coo;
}

template <class IndexType, class ValueType>
ELL_Matrix<IndexType, ValueType> csr_to_ell(const CSR_Matrix<IndexType, ValueType>& csr,
                                            const LeadingDimension ld = RowMajor) {
  ELL_Matrix<IndexType, ValueType> ell;
  auto p = lespmv::params(LESPMV_ELL, lespmv::dtype_of<ValueType>());
  p.ld = (int32_t)ld;
  auto h = lespmv::convert_on_device(csr, LESPMV_ELL, p, ell);
  ell.ld = ld;
  ell.max_row_width = (IndexType)lespmv::fetch_scalar(h, LESPMV_ELL, "max_row_width");
  ell.col_index = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_ELL, "col_index");
  ell.values = lespmv::fetch_array<ValueType, ValueType>(h, LESPMV_ELL, "values");
  return ell;
}

template <class IndexType, class ValueType>
S_ELL_Matrix<IndexType, ValueType> csr_to_sell(const CSR_Matrix<IndexType, ValueType>& csr, FILE* fp_feature,
                                               const int chunkwidth = 4,
                                               const IndexType alignment = (512 / 8 / sizeof(ValueType))) {
  (void)fp_feature;
  S_ELL_Matrix<IndexType, ValueType> s;
  auto p = lespmv::params(LESPMV_SELL, lespmv::dtype_of<ValueType>());
  p.chunk = chunkwidth;
  p.alignment = (int32_t)alignment;
  auto h = lespmv::convert_on_device(csr, LESPMV_SELL, p, s);
  s.sliceWidth = chunkwidth;
  s.alignment = alignment;
  s.chunk_num = (IndexType)lespmv::fetch_scalar(h, LESPMV_SELL, "chunk_num");
  s.row_width = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_SELL, "row_width");
  auto offb = lespmv::fetch(h, LESPMV_SELL, "chunk_offset");
  const int64_t* off = reinterpret_cast<const int64_t*>(offb.data());
  s.col_index = lespmv::split_chunks<IndexType, int32_t>(h, LESPMV_SELL, "col_index", off, s.chunk_num);
  s.values = lespmv::split_chunks<ValueType, ValueType>(h, LESPMV_SELL, "values", off, s.chunk_num);
  return s;
}

template <class IndexType, class ValueType>
SELL_C_Sigma_Matrix<IndexType, ValueType> csr_to_sell_c_sigma(const CSR_Matrix<IndexType, ValueType>& csr,
                                                              FILE* fp_feature, const int slicewidth = 16384,
                                                              const int chunkwidth = 4,
                                                              const IndexType alignment = (512 / 8 /
                                                                                           sizeof(ValueType))) {
  (void)fp_feature;
  SELL_C_Sigma_Matrix<IndexType, ValueType> s;
  auto p = lespmv::params(LESPMV_SELL_C_SIGMA, lespmv::dtype_of<ValueType>());
  p.chunk = chunkwidth;
  p.sigma = slicewidth;
  p.alignment = (int32_t)alignment;
  auto h = lespmv::convert_on_device(csr, LESPMV_SELL_C_SIGMA, p, s);
  s.sliceWidth_Sigma = slicewidth;
  s.chunkWidth_C = chunkwidth;
  s.alignment = alignment;
  s.sliceNum = (IndexType)lespmv::fetch_scalar(h, LESPMV_SELL_C_SIGMA, "sliceNum");
  s.chunkNum = (IndexType)lespmv::fetch_scalar(h, LESPMV_SELL_C_SIGMA, "chunkNum");
  s.validchunkNum = (IndexType)lespmv::fetch_scalar(h, LESPMV_SELL_C_SIGMA, "validchunkNum");
  s.chunk_num_per_slice = (IndexType)lespmv::fetch_scalar(h, LESPMV_SELL_C_SIGMA, "chunk_num_per_slice");
  s.reorder = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_SELL_C_SIGMA, "reorder");
  s.chunk_len = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_SELL_C_SIGMA, "chunk_len");
  auto offb = lespmv::fetch(h, LESPMV_SELL_C_SIGMA, "chunk_offset");
  const int64_t* off = reinterpret_cast<const int64_t*>(offb.data());
  s.col_index = lespmv::split_chunks<IndexType, int32_t>(h, LESPMV_SELL_C_SIGMA, "col_index", off, s.validchunkNum);
  s.values = lespmv::split_chunks<ValueType, ValueType>(h, LESPMV_SELL_C_SIGMA, "values", off, s.validchunkNum);
  return s;
}

template <class IndexType, class ValueType>
SELL_C_R_Matrix<IndexType, ValueType> csr_to_sell_c_R(const CSR_Matrix<IndexType, ValueType>& csr, FILE* fp_feature,
                                                      const int chunkwidth = 4,
                                                      const IndexType alignment = (512 / 8 / sizeof(ValueType))) {
  (void)fp_feature;
  SELL_C_R_Matrix<IndexType, ValueType> s;
  auto p = lespmv::params(LESPMV_SELL_C_R, lespmv::dtype_of<ValueType>());
  p.chunk = chunkwidth;
  p.alignment = (int32_t)alignment;
  auto h = lespmv::convert_on_device(csr, LESPMV_SELL_C_R, p, s);
  s.chunkWidth_C = chunkwidth;
  s.alignment = alignment;
  s.validchunkNum = (IndexType)lespmv::fetch_scalar(h, LESPMV_SELL_C_R, "validchunkNum");
  s.reorder = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_SELL_C_R, "reorder");
  s.chunk_len = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_SELL_C_R, "chunk_len");
  auto offb = lespmv::fetch(h, LESPMV_SELL_C_R, "chunk_offset");
  const int64_t* off = reinterpret_cast<const int64_t*>(offb.data());
  s.col_index = lespmv::split_chunks<IndexType, int32_t>(h, LESPMV_SELL_C_R, "col_index", off, s.validchunkNum);
  s.values = lespmv::split_chunks<ValueType, ValueType>(h, LESPMV_SELL_C_R, "values", off, s.validchunkNum);
  return s;
}

template <class IndexType, class ValueType>
DIA_Matrix<IndexType, ValueType> csr_to_dia(const CSR_Matrix<IndexType, ValueType>& csr, const IndexType max_diags,
                                            FILE* fp_feature,
                                            const IndexType alignment = (512 / 8 / sizeof(ValueType))) {
  (void)fp_feature;
  DIA_Matrix<IndexType, ValueType> dia;
  auto p = lespmv::params(LESPMV_DIA, lespmv::dtype_of<ValueType>());
  p.max_diags = (int32_t)max_diags;
  p.alignment = (int32_t)alignment;
  auto h = lespmv::convert_on_device(csr, LESPMV_DIA, p, dia);
  dia.complete_ndiags = (IndexType)lespmv::fetch_scalar(h, LESPMV_DIA, "complete_ndiags");
  dia.stride = (IndexType)lespmv::fetch_scalar(h, LESPMV_DIA, "stride");
  dia.diag_offsets = lespmv::fetch_array<long int, int64_t>(h, LESPMV_DIA, "diag_offsets");
  dia.diag_data = lespmv::fetch_array<ValueType, ValueType>(h, LESPMV_DIA, "diag_data");
  return dia;
}

template <class IndexType, class ValueType>
BSR_Matrix<IndexType, ValueType> csr_to_bsr(const CSR_Matrix<IndexType, ValueType>& csr,
                                            const IndexType blockDimRow = 16,
                                            IndexType blockDimCol = (512 / 8 / sizeof(ValueType))) {
  BSR_Matrix<IndexType, ValueType> bsr;
  auto p = lespmv::params(LESPMV_BSR, lespmv::dtype_of<ValueType>());
  p.block_rows = (int32_t)blockDimRow;
  p.block_cols = (int32_t)blockDimCol;
  auto h = lespmv::convert_on_device(csr, LESPMV_BSR, p, bsr);
  bsr.blockDim_r = blockDimRow;
  bsr.blockDim_c = blockDimCol;
  bsr.blockNNZ = blockDimRow * blockDimCol;
  bsr.mb = (IndexType)lespmv::fetch_scalar(h, LESPMV_BSR, "mb");
  bsr.nb = (IndexType)lespmv::fetch_scalar(h, LESPMV_BSR, "nb");
  bsr.nnzb = (IndexType)lespmv::fetch_scalar(h, LESPMV_BSR, "nnzb");
  bsr.row_ptr = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_BSR, "row_ptr");
  bsr.block_colindex = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_BSR, "block_colindex");
  bsr.block_data = lespmv::fetch_array<ValueType, ValueType>(h, LESPMV_BSR, "block_data");
  return bsr;
}

template <class IndexType, typename UIndexType, class ValueType>
CSR5_Matrix<IndexType, UIndexType, ValueType> csr_to_csr5(const CSR_Matrix<IndexType, ValueType>& csr,
                                                          FILE* fp_feature) {
  (void)fp_feature;
  CSR5_Matrix<IndexType, UIndexType, ValueType> c;
  auto p = lespmv::params(LESPMV_CSR5, lespmv::dtype_of<ValueType>());
  auto h = lespmv::convert_on_device(csr, LESPMV_CSR5, p, c);
  c.omega = (IndexType)lespmv::fetch_scalar(h, LESPMV_CSR5, "omega");
  c.sigma = (IndexType)lespmv::fetch_scalar(h, LESPMV_CSR5, "sigma");
  c.bit_y_offset = (IndexType)lespmv::fetch_scalar(h, LESPMV_CSR5, "bit_y_offset");
  c.bit_scansum_offset = (IndexType)lespmv::fetch_scalar(h, LESPMV_CSR5, "bit_scansum_offset");
  c.num_packets = (IndexType)lespmv::fetch_scalar(h, LESPMV_CSR5, "num_packets");
  c.num_offsets = (IndexType)lespmv::fetch_scalar(h, LESPMV_CSR5, "num_offsets");
  c._p = (IndexType)lespmv::fetch_scalar(h, LESPMV_CSR5, "p");
  c.tail_partition_start = (IndexType)lespmv::fetch_scalar(h, LESPMV_CSR5, "tail_partition_start");
  c.row_offset = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_CSR5, "row_offset");
  c.col_index = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_CSR5, "col_index");
  c.values = lespmv::fetch_array<ValueType, ValueType>(h, LESPMV_CSR5, "values");
  c.tile_ptr = lespmv::fetch_array<UIndexType, uint32_t>(h, LESPMV_CSR5, "tile_ptr");
  c.tile_desc = lespmv::fetch_array<UIndexType, uint32_t>(h, LESPMV_CSR5, "tile_desc");
  c.tile_desc_offset_ptr = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_CSR5, "tile_desc_offset_ptr");
  c.tile_desc_offset = lespmv::fetch_array<IndexType, int32_t>(h, LESPMV_CSR5, "tile_desc_offset");
  return c;
}

// ---- SpMV entry points (include/spmv_*.h) --------------------------------------

namespace lespmv {

// y = alpha*A*x + beta*y with device pointers; variant = kernel_flag.
template <class M>
void spmv_device(const M& A, typename M::value_type alpha, const typename M::value_type* dx,
                 typename M::value_type beta, typename M::value_type* dy, int default_format) {
  lespmv_matrix_t h;
  int fmt = A.gpu_format;
  if (A.gpu) {
    h = A.gpu->h;
  } else if (default_format == LESPMV_CSR) {
    h = csr_handle(reinterpret_cast<const CSR_Matrix<typename M::index_type, typename M::value_type>&>(A));
    fmt = LESPMV_CSR;
  } else {
    throw Error(LESPMV_ERR_NOT_CONVERTED, "matrix was not produced by a csr_to_<fmt> conversion");
  }
  check(lespmv_spmv(h, fmt, A.kernel_flag, alpha, dx, beta, dy, nullptr), "lespmv_spmv");
}

// Host vectors: stage x (and y when beta != 0) through device buffers.
template <class M>
void spmv_host(const M& A, typename M::value_type alpha, const typename M::value_type* x,
               typename M::value_type beta, typename M::value_type* y, int default_format) {
  using V = typename M::value_type;
  const size_t xb = sizeof(V) * (size_t)A.num_cols, yb = sizeof(V) * (size_t)A.num_rows;
  DeviceBuffer& dx = staging(0);
  DeviceBuffer& dy = staging(1);
  dx.ensure(xb);
  dy.ensure(yb);
  check(lespmv_memcpy_h2d(dx.p, x, xb), "copy x");
  if (beta != V(0)) check(lespmv_memcpy_h2d(dy.p, y, yb), "copy y");
  spmv_device(A, alpha, static_cast<const V*>(dx.p), beta, static_cast<V*>(dy.p), default_format);
  check(lespmv_memcpy_d2h(y, dy.p, yb), "copy y back");
}

}  // namespace lespmv

#define LESPMV_ENTRY(NAME, STRUCT, DEF)                                                                      \
  template <typename IndexType, typename ValueType>                                                         \
  void NAME(const ValueType alpha, const STRUCT<IndexType, ValueType>& A, const ValueType* x, const ValueType beta, \
            ValueType* y) {                                                                                 \
    lespmv::spmv_host(A, alpha, x, beta, y, DEF);                                                           \
  }                                                                                                         \
  template <typename IndexType, typename ValueType>                                                         \
  void NAME##_device(const ValueType alpha, const STRUCT<IndexType, ValueType>& A, const ValueType* dx,      \
                     const ValueType beta, ValueType* dy) {                                                 \
    lespmv::spmv_device(A, alpha, dx, beta, dy, DEF);                                                       \
  }

LESPMV_ENTRY(LeSpMV_csr, CSR_Matrix, LESPMV_CSR)
LESPMV_ENTRY(LeSpMV_coo, COO_Matrix, LESPMV_COO)
LESPMV_ENTRY(LeSpMV_ell, ELL_Matrix, LESPMV_ELL)
LESPMV_ENTRY(LeSpMV_sell, S_ELL_Matrix, LESPMV_SELL)
LESPMV_ENTRY(LeSpMV_sell_c_sigma, SELL_C_Sigma_Matrix, LESPMV_SELL_C_SIGMA)
LESPMV_ENTRY(LeSpMV_sell_c_R, SELL_C_R_Matrix, LESPMV_SELL_C_R)
LESPMV_ENTRY(LeSpMV_dia, DIA_Matrix, LESPMV_DIA)
LESPMV_ENTRY(LeSpMV_bsr, BSR_Matrix, LESPMV_BSR)
#undef LESPMV_ENTRY

template <typename IndexType, typename UIndexType, typename ValueType>
void LeSpMV_csr5(const ValueType alpha, const CSR5_Matrix<IndexType, UIndexType, ValueType>& A, const ValueType* x,
                 const ValueType beta, ValueType* y) {
  lespmv::spmv_host(A, alpha, x, beta, y, LESPMV_CSR5);
}

// ---- destructors (include/sparse_format.h:300-449) ----------------------------

template <class I, class V> void delete_coo_matrix(COO_Matrix<I, V>& a) {
  delete_array(a.row_index); delete_array(a.col_index); delete_array(a.values); a.gpu.reset();
}
template <class I, class V> void delete_csr_matrix(CSR_Matrix<I, V>& a) {
  delete_array(a.row_offset); delete_array(a.col_index); delete_array(a.values); a.gpu.reset();
}
template <class I, class V> void delete_ell_matrix(ELL_Matrix<I, V>& a) {
  delete_array(a.col_index); delete_array(a.values); a.gpu.reset();
}
template <class I, class V> void delete_dia_matrix(DIA_Matrix<I, V>& a) {
  delete_array(a.diag_offsets); delete_array(a.diag_data); a.gpu.reset();
}
template <class I, class V> void delete_bsr_matrix(BSR_Matrix<I, V>& a) {
  delete_array(a.row_ptr); delete_array(a.block_colindex); delete_array(a.block_data); a.gpu.reset();
}
template <class I, class UI, class V> void delete_csr5_matrix(CSR5_Matrix<I, UI, V>& a) {
  delete_array(a.row_offset); delete_array(a.col_index); delete_array(a.values); delete_array(a.tile_ptr);
  delete_array(a.tile_desc); delete_array(a.tile_desc_offset_ptr); delete_array(a.tile_desc_offset);
  a.gpu.reset();
}
template <class M> void delete_chunks(M& a, typename M::index_type n) {
  for (typename M::index_type c = 0; c < n; c++) { delete_array(a.col_index[c]); delete_array(a.values[c]); }
  delete[] a.col_index;
  delete[] a.values;
  a.gpu.reset();
}
template <class I, class V> void delete_s_ell_matrix(S_ELL_Matrix<I, V>& a) {
  delete_array(a.row_width); delete_chunks(a, a.chunk_num);
}
template <class I, class V> void delete_s_ell_c_sigma_matrix(SELL_C_Sigma_Matrix<I, V>& a) {
  delete_array(a.reorder); delete_array(a.chunk_len); delete_chunks(a, a.validchunkNum);
}
template <class I, class V> void delete_s_ell_c_R_matrix(SELL_C_R_Matrix<I, V>& a) {
  delete_array(a.reorder); delete_array(a.chunk_len); delete_chunks(a, a.validchunkNum);
}
template <class I, class V> void delete_host_matrix(COO_Matrix<I, V>& a) { delete_coo_matrix(a); }
template <class I, class V> void delete_host_matrix(CSR_Matrix<I, V>& a) { delete_csr_matrix(a); }
template <class I, class V> void delete_host_matrix(ELL_Matrix<I, V>& a) { delete_ell_matrix(a); }
template <class I, class V> void delete_host_matrix(DIA_Matrix<I, V>& a) { delete_dia_matrix(a); }
template <class I, class V> void delete_host_matrix(BSR_Matrix<I, V>& a) { delete_bsr_matrix(a); }
template <class I, class V> void delete_host_matrix(S_ELL_Matrix<I, V>& a) { delete_s_ell_matrix(a); }
template <class I, class V> void delete_host_matrix(SELL_C_Sigma_Matrix<I, V>& a) { delete_s_ell_c_sigma_matrix(a); }
template <class I, class V> void delete_host_matrix(SELL_C_R_Matrix<I, V>& a) { delete_s_ell_c_R_matrix(a); }
template <class I, class UI, class V> void delete_host_matrix(CSR5_Matrix<I, UI, V>& a) { delete_csr5_matrix(a); }

// ---- input (utils/sparse_io.cpp:26-204) ------------------------------------------

template <class IndexType, class ValueType>
CSR_Matrix<IndexType, ValueType> read_csr_matrix(const char* mm_filename, bool compact = false) {
  lespmv_host_csr_t h;
  lespmv::check(lespmv_read_mtx(mm_filename, lespmv::dtype_of<ValueType>(), compact ? 1 : 0, &h), "read_csr_matrix");
  CSR_Matrix<IndexType, ValueType> csr;
  csr.num_rows = (IndexType)h.m;
  csr.num_cols = (IndexType)h.n;
  csr.num_nnzs = (IndexType)h.nnz;
  csr.row_offset = new_array<IndexType>(h.m + 1);
  csr.col_index = new_array<IndexType>(h.nnz);
  csr.values = new_array<ValueType>(h.nnz);
  for (int64_t i = 0; i <= h.m; i++) csr.row_offset[i] = (IndexType)h.row_offset[i];
  for (int64_t i = 0; i < h.nnz; i++) csr.col_index[i] = (IndexType)h.col_index[i];
  std::memcpy(csr.values, h.values, sizeof(ValueType) * h.nnz);
  csr.kernel_flag = 1;  // KERNEL_FLAG, general_config.h:35
  lespmv_host_csr_free(&h);
  return csr;
}

// ---- correctness harness (include/sparse_operation.h:75-193) ---------------------

template <typename T>
T maximum_relative_error(const T* A, const T* B, const size_t N) {
  T max_error = 0;
  const T eps = std::sqrt(std::numeric_limits<T>::epsilon());
  for (size_t i = 0; i < N; i++) {
    const T error = std::abs(A[i] - B[i]);
    if (error != 0) max_error = std::max(max_error, error / (std::abs(A[i]) + std::abs(B[i]) + eps));
  }
  return max_error;
}

// compare_spmv_kernels: alpha 0.8, beta 0.7, y0 = 0, x = rand()/(RAND_MAX+1).
template <typename M1, typename F1, typename M2, typename F2>
double compare_spmv_kernels(const M1& a, F1 spmv1, const M2& b, F2 spmv2) {
  typedef typename M2::value_type V;
  const V alpha = (V)0.8, beta = (V)0.7;
  std::vector<V> x(a.num_cols), y1(a.num_rows, V(0)), y2(a.num_rows, V(0));
  for (auto& v : x) v = (V)(rand() / (RAND_MAX + 1.0));
  spmv1(alpha, a, x.data(), beta, y1.data());
  spmv2(alpha, b, x.data(), beta, y2.data());
  const V err = maximum_relative_error(y1.data(), y2.data(), (size_t)a.num_rows);
  printf(" [max error %9f]", (double)err);
  if (err > 5 * lespmv::sqrt_eps<V>() && err < 0.01) printf(" POSSIBLE small Round-Error");
  else if (err >= 0.005) printf(" POSSIBLE FAILURE");
  return (double)err;
}

template <typename M1, typename F1, typename M2, typename F2>
double test_spmv_kernel(const M1& a, F1 spmv1, const M2& b, F2 spmv2, const char* method_name) {
  printf("\ttesting %-26s[gpu]:", method_name);
  const double e = compare_spmv_kernels(a, spmv1, b, spmv2);
  printf("\n");
  return e;
}

// ---- benchmark (include/spmv_benchmark.h:166-228), timed on the device ------------

template <class M>
double benchmark_spmv_on_device(M& A, const char* method_name) {
  typedef typename M::value_type V;
  lespmv_matrix_t h = A.gpu ? A.gpu->h
                            : lespmv::csr_handle(reinterpret_cast<const CSR_Matrix<typename M::index_type, V>&>(A));
  const int fmt = A.gpu ? A.gpu_format : LESPMV_CSR;
  std::vector<V> x(A.num_cols);
  for (auto& v : x) v = (V)(rand() / (RAND_MAX + 1.0));
  lespmv::DeviceBuffer dx, dy;
  dx.ensure(sizeof(V) * x.size());
  dy.ensure(sizeof(V) * (size_t)A.num_rows);
  lespmv::check(lespmv_memcpy_h2d(dx.p, x.data(), sizeof(V) * x.size()), "copy x");
  double est = 0;
  lespmv::check(lespmv_time_spmv(h, fmt, A.kernel_flag, 1.0, dx.p, 0.0, dy.p, 1, &est), "warm-up");
  // iteration policy of the reference (:191-195): 1000 below 20 ms, else clamp(20, 20 s / t, 1000)
  int iters = est < 20.0 ? 1000 : std::min(1000, std::max(20, (int)(20.0 * 1000 / est)));
  double ms = 0;
  lespmv::check(lespmv_time_spmv(h, fmt, A.kernel_flag, 1.0, dx.p, 0.0, dy.p, iters, &ms), "benchmark");
  int64_t bytes = 0;
  lespmv::check(lespmv_reference_bytes(h, fmt, &bytes), "bytes");
  const double sec = ms / 1000.0;
  A.time = ms;
  A.gflops = sec == 0 ? 0 : 2.0 * (double)A.num_nnzs / sec / 1e9;
  A.gbytes = sec == 0 ? 0 : (double)bytes / sec / 1e9;
  printf("\tbenchmarking %-20s [gpu]: %8.4f ms ( %5.4f GFLOP/s %5.4f GB/s)\n", method_name, ms, A.gflops, A.gbytes);
  return ms;
}

// Same name and arguments as the reference; the SpMV argument is not needed
// because the matrix carries its device handle, and timing is on the device.
template <class M, class F>
double benchmark_spmv_on_host(M& A, F spmv, std::string method_name) {
  (void)spmv;
  return benchmark_spmv_on_device(A, method_name.c_str());
}

// ---- test routines (include/spmv_testroutine.h:7-50) -------------------------------

namespace lespmv {

template <class I, class V, class M, class F>
double run_test(const CSR_Matrix<I, V>& csr_ref, M& A, F spmv, int kernel_tag, const char* name, double* convert_ms) {
  A.kernel_flag = kernel_tag;
  if (convert_ms) {
    double ms = 0;
    check(lespmv_convert_ms(A.gpu->h, A.gpu_format, &ms), "convert time");
    *convert_ms = ms;
  }
  CSR_Matrix<I, V> ref = csr_ref;  // shares arrays; the device handle is its own
  ref.kernel_flag = LESPMV_VARIANT_AUTO;
  test_spmv_kernel(ref, LeSpMV_csr<I, V>, A, spmv, name);
  return benchmark_spmv_on_device(A, name);
}

}  // namespace lespmv

template <typename IndexType, typename ValueType>
double test_csr_matrix_kernels(const CSR_Matrix<IndexType, ValueType>& csr_ref, int kernel_tag, int schedule_mod) {
  (void)schedule_mod;
  printf("=====  Testing CSR Kernels  =====\n");
  CSR_Matrix<IndexType, ValueType> a = csr_ref;
  a.gpu = lespmv::upload(csr_ref);
  a.gpu_format = LESPMV_CSR;
  return lespmv::run_test(csr_ref, a, LeSpMV_csr<IndexType, ValueType>, kernel_tag, "csr_gpu", nullptr);
}

#define LESPMV_TEST_ROUTINE(NAME, CONV, SPMV, LABEL)                                                         \
  template <typename IndexType, typename ValueType>                                                         \
  double NAME(const CSR_Matrix<IndexType, ValueType>& csr_ref, int kernel_tag, int schedule_mod,             \
              double& convert_time) {                                                                       \
    (void)schedule_mod;                                                                                     \
    printf("=====  Testing " LABEL " Kernels  =====\n");                                                     \
    auto a = CONV;                                                                                          \
    const double ms = lespmv::run_test(csr_ref, a, SPMV<IndexType, ValueType>, kernel_tag, LABEL "_gpu",     \
                                       &convert_time);                                                      \
    delete_host_matrix(a);                                                                                  \
    return ms;                                                                                              \
  }

LESPMV_TEST_ROUTINE(test_coo_matrix_kernels, csr_to_coo(csr_ref), LeSpMV_coo, "COO")
LESPMV_TEST_ROUTINE(test_dia_matrix_kernels, csr_to_dia(csr_ref, (IndexType)10240, (FILE*)nullptr), LeSpMV_dia, "DIA")
LESPMV_TEST_ROUTINE(test_bsr_matrix_kernels, csr_to_bsr(csr_ref), LeSpMV_bsr, "BSR")
LESPMV_TEST_ROUTINE(test_s_ell_matrix_kernels, csr_to_sell(csr_ref, (FILE*)nullptr), LeSpMV_sell, "S_ELL")
LESPMV_TEST_ROUTINE(test_sell_c_sigma_matrix_kernels, csr_to_sell_c_sigma(csr_ref, (FILE*)nullptr),
                    LeSpMV_sell_c_sigma, "SELL-c-sigma")
#undef LESPMV_TEST_ROUTINE

template <typename IndexType, typename ValueType>
double test_ell_matrix_kernels(const CSR_Matrix<IndexType, ValueType>& csr_ref, int kernel_tag, LeadingDimension ld,
                               int schedule_mod, double& convert_time) {
  (void)schedule_mod;
  printf("=====  Testing ELL Kernels  =====\n");
  auto a = csr_to_ell(csr_ref, ld);
  const double ms = lespmv::run_test(csr_ref, a, LeSpMV_ell<IndexType, ValueType>, kernel_tag, "ELL_gpu",
                                     &convert_time);
  delete_host_matrix(a);
  return ms;
}

template <typename IndexType, typename ValueType>
double test_sell_c_R_matrix_kernels(const CSR_Matrix<IndexType, ValueType>& csr_ref, int kernel_tag,
                                    int schedule_mod) {
  (void)schedule_mod;
  printf("=====  Testing SELL-c-R Kernels  =====\n");
  auto a = csr_to_sell_c_R(csr_ref, (FILE*)nullptr);
  const double ms = lespmv::run_test(csr_ref, a, LeSpMV_sell_c_R<IndexType, ValueType>, kernel_tag, "SELL-c-R_gpu",
                                     nullptr);
  delete_host_matrix(a);
  return ms;
}

template <typename IndexType, typename UIndexType, typename ValueType>
double test_csr5_matrix_kernels(const CSR_Matrix<IndexType, ValueType>& csr_ref, int kernel_tag, int schedule_mod) {
  (void)schedule_mod;
  printf("=====  Testing CSR5 Kernels  =====\n");
  auto a = csr_to_csr5<IndexType, UIndexType, ValueType>(csr_ref, (FILE*)nullptr);
  const double ms = lespmv::run_test(csr_ref, a, LeSpMV_csr5<IndexType, UIndexType, ValueType>, kernel_tag,
                                     "CSR5_gpu", nullptr);
  delete_host_matrix(a);
  return ms;
}

// ---- feature extractor (include/sparse_features.h, src/sparse_features.cpp) -------
//
// MTX keeps the reference's method names and call order (test/Get_features.cpp:
// MtxLoad, CalculateFeatures, CalculateTilesExtraFeatures, FeaturesPrint,
// ExtraFeaturesPrint, FeaturesWrite). MtxLoad reads the file into a device CSR;
// CalculateFeatures computes every feature on the GPU in one call
// (lespmv_features), including the tile features and the row distance, so the
// two later Calculate* calls have nothing left to do. Statistics are computed in
// double whatever ValueType is. FeaturesWrite takes the output root as its
// argument instead of the reference's fixed /data/lsl/MModel-Data.

template <typename IndexType, typename ValueType>
class MTX {
 public:
  explicit MTX(IndexType matID = 0) : matrixID_(matID) { std::cout << "== MTX Features Extraction ==\n" << std::endl; }
  IndexType getRowNum() const { return (IndexType)f_.num_rows; }
  IndexType getColNum() const { return (IndexType)f_.num_cols; }
  IndexType getTileSize() const { return (IndexType)LESPMV_FEAT_TILES; }
  std::string getMatName() const { return matrixName; }
  const lespmv_features_t& features() const { return f_; }

  bool MtxLoad(const char* mat_path) {
    const std::string p(mat_path);
    const size_t slash = p.find_last_of("/\\");
    const size_t start = slash == std::string::npos ? 0 : slash + 1;
    size_t dot = p.find_last_of('.');
    if (dot == std::string::npos || dot < start) dot = p.size();
    matrixName = p.substr(start, dot - start);
    lespmv_host_csr_t h;
    lespmv::check(lespmv_read_mtx(mat_path, lespmv::dtype_of<ValueType>(), 0, &h), "MtxLoad");
    flags_ = (h.mm_flags & LESPMV_MM_SYMMETRIC) ? LESPMV_FEAT_SYMMETRIC_FILE : 0;
    handle_ = std::make_shared<lespmv::Handle>();
    const int rc = lespmv_create_csr(&handle_->h, h.dtype, h.m, h.n, h.nnz, h.row_offset, h.col_index, h.values, 0);
    lespmv_host_csr_free(&h);
    lespmv::check(rc, "MtxLoad");
    return true;
  }

  bool CalculateFeatures() {
    if (!handle_) throw lespmv::Error(LESPMV_ERR_ARG, "MtxLoad first");
    const size_t T2 = (size_t)LESPMV_FEAT_TILES * LESPMV_FEAT_TILES, T = LESPMV_FEAT_TILES;
    tile_nnz_.assign(T2, 0); tile_max_.assign(T2, 0); tile_ave_.assign(T2, 0.0); tile_std_.assign(T2, 0.0);
    rb_nnz_.assign(T, 0); rb_max_.assign(T, 0); rb_ave_.assign(T, 0.0); rb_std_.assign(T, 0.0);
    cb_nnz_.assign(T, 0); cb_max_.assign(T, 0); cb_ave_.assign(T, 0.0); cb_std_.assign(T, 0.0);
    lespmv_feature_images_t img{};
    img.tile_nnz = tile_nnz_.data(); img.tile_max = tile_max_.data(); img.tile_ave = tile_ave_.data();
    img.tile_std = tile_std_.data();
    img.rb_nnz = rb_nnz_.data(); img.rb_max = rb_max_.data(); img.rb_ave = rb_ave_.data(); img.rb_std = rb_std_.data();
    img.cb_nnz = cb_nnz_.data(); img.cb_max = cb_max_.data(); img.cb_ave = cb_ave_.data(); img.cb_std = cb_std_.data();
    lespmv::check(lespmv_features(handle_->h, flags_, &f_, &img), "CalculateFeatures");
    return true;
  }
  bool CalculateTilesFeatures() { return true; }
  bool CalculateTilesExtraFeatures(const char*) { return true; }

  void FeaturesPrint() const {
    std::cout << "[" << matrixID_ << "]  " << matrixName << "counting by " << 8 * sizeof(IndexType)
              << "-bit Index, 64-bit statistics" << std::endl;
    std::cout << "is_symmetric " << f_.is_symmetric << std::endl;
    std::cout << "pattern_symm " << f_.pattern_symm * 100 << "%" << std::endl;
    std::cout << "value_symm   " << f_.value_symm * 100 << "%" << std::endl << std::endl;
    std::cout << "num_rows   = " << f_.num_rows << std::endl;
    std::cout << "num_cols   = " << f_.num_cols << std::endl;
    std::cout << "nnz_mtx    = " << f_.nnz_mtx << std::endl;
    std::cout << "real_nnz   = " << f_.num_nnzs << std::endl;
    std::cout << "nnz_ratio  = " << f_.nnz_ratio * 100 << "%" << std::endl;
    std::cout << "num_of_diags  = " << f_.complete_ndiags << std::endl;
    std::cout << "nnz_closed2diag_ratio = " << f_.diag_close_ratio * 100 << "%" << std::endl << std::endl;
    std::cout << "nnz_lower    : " << f_.nnz_lower << std::endl;
    std::cout << "nnz_upper    : " << f_.nnz_upper << std::endl;
    std::cout << "nnz_diagonal : " << f_.nnz_diagonal << std::endl << std::endl;
    std::cout << "NoneZero_Row_Ratio      : " << f_.nz_row_ratio * 100 << "%" << std::endl;
    std::cout << "min_nnz_each_row        : " << f_.min_nnz_each_row << std::endl;
    std::cout << "max_nnz_each_row        : " << f_.max_nnz_each_row << std::endl;
    std::cout << "ave_nnz_each_row        : " << f_.ave_nnz_each_row << std::endl;
    std::cout << "var_nnz_each_row        : " << f_.var_nnz_each_row << std::endl;
    std::cout << "standard_dev_row        : " << f_.standard_dev_row << std::endl;
    std::cout << "P-ratio_row             : " << f_.P_ratio_row << std::endl;
    std::cout << "Gini_coeff_row          : " << f_.Gini_row << std::endl << std::endl;
    std::cout << "NoneZero_Col_Ratio      : " << f_.nz_col_ratio * 100 << "%" << std::endl;
    std::cout << "min_nnz_each_col        : " << f_.min_nnz_each_col << std::endl;
    std::cout << "max_nnz_each_col        : " << f_.max_nnz_each_col << std::endl;
    std::cout << "ave_nnz_each_col        : " << f_.ave_nnz_each_col << std::endl;
    std::cout << "var_nnz_each_col        : " << f_.var_nnz_each_col << std::endl;
    std::cout << "standard_dev_col        : " << f_.standard_dev_col << std::endl;
    std::cout << "P-ratio_col             : " << f_.P_ratio_col << std::endl;
    std::cout << "Gini_coeff_col          : " << f_.Gini_col << std::endl << std::endl;
    std::cout << "diagonal_dominant_ratio : " << f_.diagonal_dominant_ratio * 100 << "%" << std::endl << std::endl;
    std::cout << "max_value_offdiag  = " << f_.max_value_offdiag << std::endl;
    std::cout << "max_value_diagonal = " << f_.max_value_diagonal << std::endl;
    std::cout << "row_variability    = " << f_.row_variability << std::endl;
    std::cout << "col_variability    = " << f_.col_variability << std::endl << std::endl;
    std::cout << "==========   Tile Features   ==========" << std::endl;
    std::cout << "Number of Tiles  : " << f_.t_num_blocks << std::endl;
    std::cout << "Tile inner size  : " << f_.t_num_RB << " * " << f_.t_num_CB << std::endl;
    std::cout << "t_ave_nnz_tiles     : " << f_.t_ave_nnz_all_tiles << std::endl;
    std::cout << "t_ave_nnz_RB        : " << f_.t_ave_nnz_RB << std::endl;
    std::cout << "t_ave_nnz_CB        : " << f_.t_ave_nnz_CB << std::endl << std::endl;
    std::cout << "t_var_nnz_tiles     : " << f_.t_var_nnz_all_tiles << std::endl;
    std::cout << "t_var_nnz_RB        : " << f_.t_var_nnz_RB << std::endl;
    std::cout << "t_var_nnz_CB        : " << f_.t_var_nnz_CB << std::endl << std::endl;
    std::cout << "t_stand_dev_tiles     : " << f_.t_standard_dev_all_tiles << std::endl;
    std::cout << "t_stand_dev_RB        : " << f_.t_standard_dev_RB << std::endl;
    std::cout << "t_stand_dev_CB        : " << f_.t_standard_dev_CB << std::endl << std::endl;
    std::cout << "min_nnz_each_tiles       : " << f_.t_min_nnz_all_tiles << std::endl;
    std::cout << "max_nnz_each_tiles       : " << f_.t_max_nnz_all_tiles << std::endl;
    std::cout << "min_nnz_each_RB          : " << f_.t_min_nnz_each_RB << std::endl;
    std::cout << "max_nnz_each_RB          : " << f_.t_max_nnz_each_RB << std::endl;
    std::cout << "min_nnz_each_CB          : " << f_.t_min_nnz_each_CB << std::endl;
    std::cout << "max_nnz_each_CB          : " << f_.t_max_nnz_each_CB << std::endl << std::endl;
    std::cout << "Gini_nnz_tiles     : " << f_.t_Gini_all_tiles << std::endl;
    std::cout << "Gini_nnz_RB        : " << f_.t_Gini_RB << std::endl;
    std::cout << "Gini_nnz_CB        : " << f_.t_Gini_CB << std::endl << std::endl;
    std::cout << "P-ratio_tiles      : " << f_.t_P_ratio_all_tiles << std::endl;
    std::cout << "P-ratio_RB         : " << f_.t_P_ratio_RB << std::endl;
    std::cout << "P-ratio_CB         : " << f_.t_P_ratio_CB << std::endl << std::endl;
    std::cout << "NE_ratio_tiles_     : " << f_.t_nz_ratio_tiles << std::endl;
    std::cout << "NE_ratio_RB         : " << f_.t_nz_ratio_RB << std::endl;
    std::cout << "NE_ratio_CB         : " << f_.t_nz_ratio_CB << std::endl << std::endl;
  }

  void ExtraFeaturesPrint() const {
    std::cout << "==========   Tile Extra Features   ==========" << std::endl;
    std::cout << "uniq_R         = " << f_.uniqR << std::endl;
    std::cout << "uniq_C         = " << f_.uniqC << std::endl << std::endl;
    std::cout << "GrX_uniqR      = " << f_.GrX_uniqR << std::endl;
    std::cout << "GrX_uniqC      = " << f_.GrX_uniqC << std::endl << std::endl;
    std::cout << "potReuseR      = " << f_.potReuseR << std::endl;
    std::cout << "potReuseC      = " << f_.potReuseC << std::endl << std::endl;
    std::cout << "GrX_potReuseR  = " << f_.GrX_potReuseR << std::endl;
    std::cout << "GrX_potReuseC  = " << f_.GrX_potReuseC << std::endl << std::endl;
    std::cout << "ave_distance_= " << f_.distance_per_row << std::endl << std::endl;
  }

  // <root>/<name>/<name>.features plus the .nnz/.max/.ave/.std and RB/CB images
  // (src/sparse_features.cpp:1190-1405).
  bool FeaturesWrite(const char* root) const {
    const std::string dir = std::string(root) + "/" + matrixName;
    const std::string mk = "mkdir -p '" + dir + "'";
    if (std::system(mk.c_str()) != 0) return false;
    const std::string base = dir + "/" + matrixName;
    auto image = [&](const std::string& suffix, const auto& data) {
      std::ofstream out(base + suffix);
      for (const auto& v : data) out << v << std::endl;
      return (bool)out;
    };
    bool ok = image(".ave", tile_ave_) && image(".max", tile_max_) && image(".nnz", tile_nnz_) &&
              image(".std", tile_std_) && image(".RBmax", rb_max_) && image(".RBnnz", rb_nnz_) &&
              image(".RBave", rb_ave_) && image(".RBstd", rb_std_) && image(".CBmax", cb_max_) &&
              image(".CBnnz", cb_nnz_) && image(".CBave", cb_ave_) && image(".CBstd", cb_std_);
    double v[LESPMV_NUM_FEATURES];
    lespmv::check(lespmv_features_vector(&f_, v), "FeaturesWrite");
    std::ofstream out(base + ".features");
    out << "[" << matrixID_ << "]: " << matrixName << std::endl;
    for (int i = 0; i < LESPMV_NUM_FEATURES; i++) {
      const std::string name = lespmv_feature_name(i);
      const bool integer = name == "num_rows" || name == "num_cols" || name == "num_nnz" ||
                           name.rfind("max_nnz_each_", 0) == 0;
      out << name << " ";
      if (integer) out << (long long)v[i];
      else out << v[i];
      out << std::endl;
    }
    return ok && (bool)out;
  }

 private:
  IndexType matrixID_ = 0;
  std::string matrixName;
  int32_t flags_ = 0;
  std::shared_ptr<lespmv::Handle> handle_;
  lespmv_features_t f_{};
  std::vector<int32_t> tile_nnz_, tile_max_, rb_nnz_, rb_max_, cb_nnz_, cb_max_;
  std::vector<double> tile_ave_, tile_std_, rb_ave_, rb_std_, cb_ave_, cb_std_;
};
