// C ABI of the gfx950 LeSpMV library (declared in include/lespmv_hip.h).
// Every entry point converts exceptions into a status code and keeps the
// message for lespmv_last_error(); nothing exits the process.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>

#include "common.hpp"

namespace lsp {

thread_local std::string g_last_error;

Matrix::~Matrix() {
  for (auto& f : fmt) f.reset();
  if (ev_switch) (void)hipEventDestroy(ev_switch);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

template <class F>
int guarded(F&& f) {
  return guarded_call(std::forward<F>(f));
}

// Any handle; entry points that take 64-bit-offset handles (CSR SpMV / SpMM,
// their timing, sizes) use it.
Matrix& mat_any(lespmv_matrix_t h) {
  LSP_REQUIRE(h != nullptr, LESPMV_ERR_ARG, "null matrix handle");
  return *reinterpret_cast<Matrix*>(h);
}

// A handle with 32-bit row offsets: every other entry point.
Matrix& mat(lespmv_matrix_t h) {
  Matrix& A = mat_any(h);
  LSP_REQUIRE(!A.idx64, LESPMV_ERR_UNSUPPORTED, "a handle with 64-bit row offsets runs CSR SpMV and SpMM only");
  return A;
}

void check_fmt(int fmt) {
  LSP_REQUIRE(fmt >= 0 && fmt < LESPMV_NUM_FORMATS, LESPMV_ERR_ARG, "unknown format id " + std::to_string(fmt));
}

hipStream_t pick(Matrix& A, lespmv_stream_t s) { return s ? static_cast<hipStream_t>(s) : A.stream; }

void require_format(Matrix& A, int fmt) {
  if (fmt == LESPMV_CSR) return;
  LSP_REQUIRE(A.fmt[fmt] != nullptr, LESPMV_ERR_NOT_CONVERTED,
              "format " + std::to_string(fmt) + " has not been converted on this handle");
}

// The handle's scratch arrays serve one call at a time. A call on another
// stream than the previous call waits for that call first: after a call on a
// caller's stream an event is recorded there (the stream is alive then); the
// handle's own stream, which lives as long as the handle, is marked only when
// another stream takes over.
struct StreamOrder {
  Matrix& A;
  hipStream_t s;
  StreamOrder(Matrix& a, hipStream_t st) : A(a), s(st) {
    if (A.last_stream && A.last_stream != s) {
      if (A.last_stream == A.stream) LSP_HIP(hipEventRecord(A.ev_switch, A.stream));
      LSP_HIP(hipStreamWaitEvent(s, A.ev_switch, 0));
    }
  }
  void done() {
    if (s != A.stream) LSP_HIP(hipEventRecord(A.ev_switch, s));
    A.last_stream = s;
  }
};

void dispatch_spmv(Matrix& A, int fmt, int variant, double alpha, const void* x, double beta, void* y,
                   hipStream_t s) {
  switch (fmt) {
    case LESPMV_CSR: csr_spmv(A, variant, alpha, x, beta, y, s); break;
    case LESPMV_COO: coo_spmv(A, variant, alpha, x, beta, y, s); break;
    case LESPMV_ELL:
    case LESPMV_SELL:
    case LESPMV_SELL_C_SIGMA:
    case LESPMV_SELL_C_R: sliced_spmv(A, fmt, alpha, x, beta, y, s); break;
    case LESPMV_DIA: dia_spmv(A, variant, alpha, x, beta, y, s); break;
    case LESPMV_BSR: bsr_spmv(A, variant, alpha, x, beta, y, s); break;
    case LESPMV_CSR5: csr5_spmv(A, alpha, x, beta, y, s); break;
    default: throw Error(LESPMV_ERR_ARG, "bad format");
  }
}

void run_spmv(Matrix& A, int fmt, int variant, double alpha, const void* x, double beta, void* y, hipStream_t s) {
  require_format(A, fmt);
  StreamOrder order(A, s);
  if (A.idx64) {
    for (size_t b = 0; b < A.blocks.size(); b++)
      csr_spmv(*A.blocks[b], variant, alpha, x, beta,
               static_cast<char*>(y) + A.block_row[b] * (int64_t)vsize(A.dtype), s);
  } else {
    dispatch_spmv(A, fmt, variant, alpha, x, beta, y, s);
  }
  order.done();
}

void run_spmm(Matrix& A, int fmt, int variant, int k, double alpha, const void* X, int64_t ldx, double beta,
              void* Y, int64_t ldy, hipStream_t s) {
  require_format(A, fmt);
  LSP_REQUIRE(k >= 0 && k <= 64, LESPMV_ERR_UNSUPPORTED, "SpMM supports 0 <= k <= 64");
  LSP_REQUIRE(ldx >= k && ldy >= k, LESPMV_ERR_ARG, "leading dimensions must be >= k");
  StreamOrder order(A, s);
  if (A.idx64) {
    for (size_t b = 0; b < A.blocks.size(); b++)
      csr_spmm(*A.blocks[b], variant, k, alpha, X, ldx, beta,
               static_cast<char*>(Y) + A.block_row[b] * ldy * (int64_t)vsize(A.dtype), ldy, s);
    order.done();
    return;
  }
  switch (fmt) {
    case LESPMV_CSR: csr_spmm(A, variant, k, alpha, X, ldx, beta, Y, ldy, s); break;
    case LESPMV_COO: coo_spmm(A, variant, k, alpha, X, ldx, beta, Y, ldy, s); break;
    case LESPMV_ELL:
    case LESPMV_SELL:
    case LESPMV_SELL_C_SIGMA:
    case LESPMV_SELL_C_R: sliced_spmm(A, fmt, k, alpha, X, ldx, beta, Y, ldy, s); break;
    case LESPMV_DIA: dia_spmm(A, variant, k, alpha, X, ldx, beta, Y, ldy, s); break;
    case LESPMV_BSR: bsr_spmm(A, variant, k, alpha, X, ldx, beta, Y, ldy, s); break;
    case LESPMV_CSR5: csr5_spmm(A, k, alpha, X, ldx, beta, Y, ldy, s); break;
    default: throw Error(LESPMV_ERR_UNSUPPORTED, "SpMM is not provided for this format");
  }
  order.done();
}

// Algorithmic bytes of one SpMV/SpMM: values + indices actually describing
// the matrix + x read once + y written (+ read when beta != 0).
int64_t effective_bytes(Matrix& A, int fmt, int k, bool beta_nz) {
  const int64_t sv = (int64_t)vsize(A.dtype), si = 4;
  const int64_t vec = (int64_t)k * (A.n + A.m * (beta_nz ? 2 : 1)) * sv;
  switch (fmt) {
    case LESPMV_CSR: return A.nnz * (sv + si) + (A.m + 1) * (A.idx64 ? 8 : si) + vec;
    case LESPMV_COO: return A.nnz * (sv + 2 * si) + vec;
    case LESPMV_ELL:
    case LESPMV_SELL:
    case LESPMV_SELL_C_SIGMA:
    case LESPMV_SELL_C_R: {
      // ELL is described by one width (max_row_width); the sliced formats by one
      // length per chunk and, when sorted, the row order
      auto* d = static_cast<SlicedData*>(A.fmt[fmt].get());
      const bool sorted = d->reorder.p != nullptr;
      const int64_t chunks = fmt == LESPMV_ELL ? 0 : d->nchunks;
      return A.nnz * (sv + si) + chunks * si + (sorted ? A.m * si : 0) + vec;
    }
    case LESPMV_DIA: {
      auto* d = static_cast<DiaData*>(A.fmt[fmt].get());
      return A.nnz * sv + (int64_t)d->ndiag * 8 + vec;
    }
    case LESPMV_BSR: {
      auto* d = static_cast<BsrData*>(A.fmt[fmt].get());
      return A.nnz * sv + d->nnzb * si + (d->mb + 1) * si + vec;
    }
    case LESPMV_CSR5: {
      auto* d = static_cast<Csr5Data*>(A.fmt[fmt].get());
      return A.nnz * (sv + si) + (d->p + 1) * 4 + d->p * d->omega * d->packets * 4 + d->num_offsets * si + vec;
    }
  }
  return 0;
}

// bytes_per_spmv of the reference (include/spmv_benchmark.h:26-150).
int64_t reference_bytes(Matrix& A, int fmt) {
  const int64_t sv = (int64_t)vsize(A.dtype), si = 4;
  switch (fmt) {
    case LESPMV_CSR: return 2 * si * A.m + si * A.nnz + 2 * sv * A.nnz + 2 * sv * A.m;
    case LESPMV_COO: {
      // every occupied row counts 2*sizeof(V)
      auto rp = to_host<int32_t>(A.rp.p, A.m + 1, A.stream);
      int64_t occ = 0;
      for (int64_t i = 0; i < A.m; i++) occ += rp[i + 1] > rp[i];
      return 2 * si * A.nnz + 2 * sv * A.nnz + 2 * sv * occ;
    }
    case LESPMV_BSR: {
      auto* d = static_cast<BsrData*>(A.fmt[fmt].get());
      return 2 * si * d->mb + si * d->nnzb + 2 * sv * d->nnzb * d->br * d->bc + 2 * sv * A.m;
    }
    case LESPMV_CSR5: {
      auto* d = static_cast<Csr5Data*>(A.fmt[fmt].get());
      return 2 * si * A.m + si * A.nnz + 2 * sv * A.nnz + 2 * sv * A.m + 4 * d->p + 4;
    }
    case LESPMV_ELL: {
      auto* d = static_cast<SlicedData*>(A.fmt[fmt].get());
      return si * A.m * d->max_width + sv * A.m * d->max_width + sv * A.nnz + 2 * sv * A.m;
    }
    case LESPMV_SELL:
    case LESPMV_SELL_C_SIGMA:
    case LESPMV_SELL_C_R: {
      // values are counted with sizeof(IndexType), as the reference does (:98,116,134)
      auto* d = static_cast<SlicedData*>(A.fmt[fmt].get());
      int64_t b = si * d->nchunks + 2 * si * d->slots + sv * A.nnz + 2 * sv * A.m;
      if (fmt != LESPMV_SELL) b += si * A.m;
      return b;
    }
    case LESPMV_DIA: return 2 * sv * A.nnz + 2 * sv * A.m;
  }
  return 0;
}

void export_array(Matrix& A, int fmt, const std::string& name, std::vector<unsigned char>& out) {
  hipStream_t s = A.stream;
  if (fmt == LESPMV_CSR) {
    if (name == "row_offset") put_bytes(out, to_host<int32_t>(A.rp.p, A.m + 1, s));
    else if (name == "col_index") put_bytes(out, to_host<int32_t>(A.ci.p, A.nnz, s));
    else if (name == "values") {
      out.resize(A.nnz * vsize(A.dtype));
      if (A.nnz) {
        LSP_HIP(hipMemcpyAsync(out.data(), A.val.p, out.size(), hipMemcpyDeviceToHost, s));
        LSP_HIP(hipStreamSynchronize(s));
      }
    } else throw Error(LESPMV_ERR_ARG, "unknown CSR array " + name);
    return;
  }
  require_format(A, fmt);
  switch (fmt) {
    case LESPMV_COO: {
      if (name == "row_index") put_bytes(out, to_host<int32_t>(static_cast<CooData*>(A.fmt[fmt].get())->row.p, A.nnz, s));
      else export_array(A, LESPMV_CSR, name, out);
      return;
    }
    case LESPMV_ELL:
    case LESPMV_SELL:
    case LESPMV_SELL_C_SIGMA:
    case LESPMV_SELL_C_R: sliced_export(A, fmt, name, out); return;
    case LESPMV_CSR5: csr5_export(A, name, out); return;
    case LESPMV_DIA: {
      auto* d = static_cast<DiaData*>(A.fmt[fmt].get());
      if (name == "diag_offsets") put_bytes(out, to_host<int64_t>(d->offsets.p, d->ndiag, s));
      else if (name == "diag_data") {
        out.resize((size_t)d->ndiag * d->stride * vsize(A.dtype));
        if (!out.empty()) {
          LSP_HIP(hipMemcpyAsync(out.data(), d->data.p, out.size(), hipMemcpyDeviceToHost, s));
          LSP_HIP(hipStreamSynchronize(s));
        }
      } else if (name == "complete_ndiags") put_scalar(out, d->ndiag);
      else if (name == "stride") put_scalar(out, d->stride);
      else throw Error(LESPMV_ERR_ARG, "unknown DIA array " + name);
      return;
    }
    case LESPMV_BSR: {
      auto* d = static_cast<BsrData*>(A.fmt[fmt].get());
      if (name == "row_ptr") put_bytes(out, to_host<int32_t>(d->row_ptr.p, d->mb + 1, s));
      else if (name == "block_colindex") put_bytes(out, to_host<int32_t>(d->bcol.p, d->nnzb, s));
      else if (name == "block_data") {
        out.resize((size_t)d->nnzb * d->br * d->bc * vsize(A.dtype));
        if (!out.empty()) {
          LSP_HIP(hipMemcpyAsync(out.data(), d->data.p, out.size(), hipMemcpyDeviceToHost, s));
          LSP_HIP(hipStreamSynchronize(s));
        }
      } else if (name == "mb") put_scalar(out, d->mb);
      else if (name == "nb") put_scalar(out, d->nb);
      else if (name == "nnzb") put_scalar(out, d->nnzb);
      else if (name == "blockDim_r") put_scalar(out, d->br);
      else if (name == "blockDim_c") put_scalar(out, d->bc);
      else throw Error(LESPMV_ERR_ARG, "unknown BSR array " + name);
      return;
    }
  }
  throw Error(LESPMV_ERR_ARG, "unknown export");
}

}  // namespace lsp

using namespace lsp;

namespace {

// An empty handle: device, stream and events.
std::unique_ptr<Matrix> new_matrix(int dtype, int64_t m, int64_t n, int64_t nnz) {
  auto A = std::make_unique<Matrix>();
  A->dtype = dtype;
  A->m = m;
  A->n = n;
  A->nnz = nnz;
  LSP_HIP(hipGetDevice(&A->device));
  LSP_HIP(hipStreamCreate(&A->stream));
  LSP_HIP(hipEventCreateWithFlags(&A->ev_switch, hipEventDisableTiming));
  LSP_HIP(hipEventCreate(&A->ev0));
  LSP_HIP(hipEventCreate(&A->ev1));
  return A;
}

// A 32-bit CSR handle with its arrays copied in (kind: from host or device).
std::unique_ptr<Matrix> make_csr(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* row_offset,
                                 const int32_t* col_index, const void* values, hipMemcpyKind kind) {
  auto A = new_matrix(dtype, m, n, nnz);
  A->rp.alloc(sizeof(int32_t) * (m + 1));
  A->ci.alloc(sizeof(int32_t) * nnz);
  A->val.alloc(vsize(dtype) * nnz);
  LSP_HIP(hipMemcpyAsync(A->rp.p, row_offset, sizeof(int32_t) * (m + 1), kind, A->stream));
  if (nnz) {
    LSP_HIP(hipMemcpyAsync(A->ci.p, col_index, sizeof(int32_t) * nnz, kind, A->stream));
    LSP_HIP(hipMemcpyAsync(A->val.p, values, vsize(dtype) * nnz, kind, A->stream));
  }
  LSP_HIP(hipStreamSynchronize(A->stream));
  return A;
}

}  // namespace

extern "C" {

const char* lespmv_last_error(void) { return g_last_error.c_str(); }

int lespmv_init(int device) {
  return guarded([&] {
    LSP_HIP(hipSetDevice(device));
    (void)noalloc_scratch(device);  // the allocation-free SpMV's scratch address on this device
  });
}


int lespmv_device_info(int64_t* hbm_bytes, int32_t* num_cu, int32_t* arch_gfx) {
  return guarded([&] {
    int dev = 0;
    LSP_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    LSP_HIP(hipGetDeviceProperties(&p, dev));
    if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
    if (num_cu) *num_cu = p.multiProcessorCount;
    if (arch_gfx) *arch_gfx = (int32_t)std::strtol(p.gcnArchName + 3, nullptr, 10);  // "gfx950..." -> 950
  });
}

int lespmv_default_params(int fmt, int dtype, int reference_layout, lespmv_params_t* out) {
  return guarded([&] {
    check_fmt(fmt);
    LSP_REQUIRE(out, LESPMV_ERR_ARG, "null params");
    const int32_t align = (int32_t)(64 / vsize(dtype));  // SIMD_WIDTH/8/sizeof(V), plat_config.h:17
    lespmv_params_t p{};
    p.max_diags = 10240;  // MAX_DIAG_NUM, general_config.h:9
    if (reference_layout) {
      p.chunk = 4;        // CHUNK_SIZE, general_config.h:16
      p.sigma = fmt == LESPMV_CSR5 ? 16 : 16384;  // CSR5_SIGMA :24, SELL_SIGMA :15
      p.alignment = align;
      p.block_rows = 16;  // BSR_BlockDimRow, general_config.h:25
      p.block_cols = align;
      p.omega = align;    // SIMD_WIDTH/8/sizeof(V), sparse_conversion.h:963
    } else {
      p.chunk = 64;       // one wavefront per chunk
      p.sigma = fmt == LESPMV_CSR5 ? 16 : 4096;
      p.alignment = 1;
      p.block_rows = 4;
      p.block_cols = 4;
      p.omega = 64;       // one wavefront per CSR5 tile
    }
    p.ld = 0;
    *out = p;
  });
}

int lespmv_create_csr(lespmv_matrix_t* out, int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* row_offset,
                      const int32_t* col_index, const void* values, int on_device) {
  return guarded([&] {
    LSP_REQUIRE(out, LESPMV_ERR_ARG, "null output handle");
    LSP_REQUIRE(dtype == LESPMV_F32 || dtype == LESPMV_F64, LESPMV_ERR_ARG, "dtype must be F32 or F64");
    LSP_REQUIRE(m >= 0 && n >= 0 && nnz >= 0, LESPMV_ERR_ARG, "negative size");
    LSP_REQUIRE(m < (int64_t)INT32_MAX && n < (int64_t)INT32_MAX && nnz < (int64_t)INT32_MAX,
                LESPMV_ERR_UNSUPPORTED, "sizes must fit 32-bit indices");
    LSP_REQUIRE(row_offset && (nnz == 0 || (col_index && values)), LESPMV_ERR_ARG, "null CSR array");
    auto A = make_csr(dtype, m, n, nnz, row_offset, col_index, values,
                      on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice);
    *out = reinterpret_cast<lespmv_matrix_t>(A.release());
  });
}

int lespmv_create_csr64(lespmv_matrix_t* out, int dtype, int64_t m, int64_t n, int64_t nnz, const int64_t* row_offset,
                        const int32_t* col_index, const void* values, int on_device) {
  return guarded([&] {
    LSP_REQUIRE(out, LESPMV_ERR_ARG, "null output handle");
    LSP_REQUIRE(dtype == LESPMV_F32 || dtype == LESPMV_F64, LESPMV_ERR_ARG, "dtype must be F32 or F64");
    LSP_REQUIRE(m >= 0 && n >= 0 && nnz >= 0, LESPMV_ERR_ARG, "negative size");
    LSP_REQUIRE(m < (int64_t)INT32_MAX && n < (int64_t)INT32_MAX, LESPMV_ERR_UNSUPPORTED,
                "rows and columns must fit 32-bit indices");
    LSP_REQUIRE(row_offset && (nnz == 0 || (col_index && values)), LESPMV_ERR_ARG, "null CSR array");
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    std::vector<int64_t> rp64;
    const int64_t* rp = row_offset;
    if (on_device) {
      rp64.resize(m + 1);
      LSP_HIP(hipMemcpy(rp64.data(), row_offset, sizeof(int64_t) * (m + 1), hipMemcpyDeviceToHost));
      rp = rp64.data();
    }
    LSP_REQUIRE(rp[0] == 0 && rp[m] == nnz, LESPMV_ERR_ARG, "row offsets must run from 0 to nnz");
    auto A = new_matrix(dtype, m, n, nnz);
    A->idx64 = true;
    // row blocks of at most INT32_MAX - 1 nonzeros, each with 32-bit offsets from its first entry
    int64_t r0 = 0;
    A->block_row.push_back(0);
    std::vector<int32_t> local;
    while (r0 < m || A->blocks.empty()) {
      const int64_t limit = rp[r0] + (int64_t)INT32_MAX - 1;
      const int64_t r1 = std::max<int64_t>(std::upper_bound(rp + r0, rp + m + 1, limit) - rp - 1, r0);
      LSP_REQUIRE(r1 > r0 || m == 0, LESPMV_ERR_UNSUPPORTED, "a row holds 2^31 - 1 or more nonzeros");
      local.resize(r1 - r0 + 1);
      for (int64_t r = r0; r <= r1; r++) local[r - r0] = (int32_t)(rp[r] - rp[r0]);
      const int64_t e0 = rp[r0], cnt = rp[r1] - rp[r0];
      A->blocks.push_back(make_csr(dtype, r1 - r0, n, cnt, local.data(), col_index ? col_index + e0 : nullptr,
                                   values ? static_cast<const char*>(values) + e0 * (int64_t)vsize(dtype) : nullptr,
                                   kind));
      A->block_row.push_back(r1);
      r0 = r1;
      if (m == 0) break;
    }
    *out = reinterpret_cast<lespmv_matrix_t>(A.release());
  });
}

int lespmv_destroy(lespmv_matrix_t h) {
  return guarded([&] {
    if (!h) return;
    Matrix* A = reinterpret_cast<Matrix*>(h);
    (void)hipStreamSynchronize(A->stream);
    // the last call on a caller's stream recorded ev_switch there
    if (A->last_stream && A->last_stream != A->stream) (void)hipEventSynchronize(A->ev_switch);
    delete A;
  });
}

int lespmv_matrix_info(lespmv_matrix_t h, int64_t* m, int64_t* n, int64_t* nnz, int32_t* dtype) {
  return guarded([&] {
    Matrix& A = mat_any(h);
    if (m) *m = A.m;
    if (n) *n = A.n;
    if (nnz) *nnz = A.nnz;
    if (dtype) *dtype = A.dtype;
  });
}

int lespmv_convert(lespmv_matrix_t h, int fmt, const lespmv_params_t* prm) {
  return guarded([&] {
    Matrix& A = mat(h);
    check_fmt(fmt);
    lespmv_params_t p{};
    if (prm) p = *prm;
    else lespmv_default_params(fmt, A.dtype, 0, &p);
    hipStream_t s = A.stream;
    LSP_HIP(hipEventRecord(A.ev0, s));
    switch (fmt) {
      case LESPMV_CSR: csr_analyse(A, s); break;
      case LESPMV_COO: convert_coo(A, p, s); break;
      case LESPMV_ELL:
      case LESPMV_SELL:
      case LESPMV_SELL_C_SIGMA:
      case LESPMV_SELL_C_R: convert_sliced(A, fmt, p, s); break;
      case LESPMV_DIA: convert_dia(A, p, s); break;
      case LESPMV_BSR: convert_bsr(A, p, s); break;
      case LESPMV_CSR5: convert_csr5(A, p, s); break;
    }
    LSP_HIP(hipEventRecord(A.ev1, s));
    LSP_HIP(hipEventSynchronize(A.ev1));
    float ms = 0;
    LSP_HIP(hipEventElapsedTime(&ms, A.ev0, A.ev1));
    if (fmt != LESPMV_CSR && A.fmt[fmt]) {
      A.fmt[fmt]->params = p;
      A.fmt[fmt]->convert_ms = ms;
    }
  });
}

int lespmv_prepare(lespmv_matrix_t h, int fmt, int variant) {
  return guarded([&] {
    Matrix& A = mat(h);
    check_fmt(fmt);
    require_format(A, fmt);
    if (fmt != LESPMV_CSR) return;  // the other formats are complete after lespmv_convert
    hipStream_t s = A.stream;
    if (A.last_stream && A.last_stream != A.stream) LSP_HIP(hipEventSynchronize(A.ev_switch));
    LSP_HIP(hipEventRecord(A.ev0, s));
    csr_prepare(A, variant, s);
    LSP_HIP(hipEventRecord(A.ev1, s));
    LSP_HIP(hipEventSynchronize(A.ev1));
    float ms = 0;
    LSP_HIP(hipEventElapsedTime(&ms, A.ev0, A.ev1));
    A.prepare_ms = ms;
  });
}

int lespmv_csr_plan(lespmv_matrix_t h, int32_t* variant, int32_t* lanes, double* far_lines) {
  return guarded([&] {
    Matrix& A = mat(h);
    if (!A.meta.analysed) {
      if (A.last_stream && A.last_stream != A.stream) LSP_HIP(hipEventSynchronize(A.ev_switch));
      csr_analyse(A, A.stream);
    }
    if (variant) *variant = A.meta.auto_variant;
    if (lanes) *lanes = A.meta.lanes;
    if (far_lines) *far_lines = A.meta.far_frac;
  });
}

int lespmv_release_format(lespmv_matrix_t h, int fmt) {
  return guarded([&] {
    Matrix& A = mat(h);
    check_fmt(fmt);
    LSP_HIP(hipStreamSynchronize(A.stream));
    if (A.last_stream && A.last_stream != A.stream) LSP_HIP(hipEventSynchronize(A.ev_switch));
    if (fmt == LESPMV_CSR) {
      csr_release_aux(A);  // the CSR arrays themselves stay
      A.prepare_ms = 0;
    } else {
      A.fmt[fmt].reset();
    }
  });
}

int lespmv_has_format(lespmv_matrix_t h, int fmt) {
  if (!h || fmt < 0 || fmt >= LESPMV_NUM_FORMATS) return 0;
  Matrix& A = *reinterpret_cast<Matrix*>(h);
  return fmt == LESPMV_CSR || A.fmt[fmt] != nullptr;
}

int lespmv_convert_ms(lespmv_matrix_t h, int fmt, double* ms) {
  return guarded([&] {
    Matrix& A = mat(h);
    check_fmt(fmt);
    require_format(A, fmt);
    *ms = fmt == LESPMV_CSR ? A.prepare_ms : A.fmt[fmt]->convert_ms;
  });
}

int lespmv_spmv(lespmv_matrix_t h, int fmt, int variant, double alpha, const void* x, double beta, void* y,
                lespmv_stream_t stream) {
  return guarded([&] {
    Matrix& A = mat_any(h);
    check_fmt(fmt);
    LSP_REQUIRE((x || A.n == 0) && (y || A.m == 0), LESPMV_ERR_ARG, "null vector");
    run_spmv(A, fmt, variant, alpha, x, beta, y, pick(A, stream));
  });
}

int lespmv_spmm(lespmv_matrix_t h, int fmt, int variant, int32_t k, double alpha, const void* X, int64_t ldx,
                double beta, void* Y, int64_t ldy, lespmv_stream_t stream) {
  return guarded([&] {
    Matrix& A = mat_any(h);
    check_fmt(fmt);
    run_spmm(A, fmt, variant, k, alpha, X, ldx, beta, Y, ldy, pick(A, stream));
  });
}

int lespmv_time_spmv(lespmv_matrix_t h, int fmt, int variant, double alpha, const void* x, double beta, void* y,
                     int32_t iters, double* ms_per_call) {
  return guarded([&] {
    Matrix& A = mat_any(h);
    check_fmt(fmt);
    LSP_REQUIRE(iters > 0 && ms_per_call, LESPMV_ERR_ARG, "iters must be > 0");
    hipStream_t s = A.stream;
    LSP_HIP(hipEventRecord(A.ev0, s));
    for (int i = 0; i < iters; i++) run_spmv(A, fmt, variant, alpha, x, beta, y, s);
    LSP_HIP(hipEventRecord(A.ev1, s));
    LSP_HIP(hipEventSynchronize(A.ev1));
    float ms = 0;
    LSP_HIP(hipEventElapsedTime(&ms, A.ev0, A.ev1));
    *ms_per_call = ms / iters;
  });
}

int lespmv_time_spmv_rotating(const lespmv_matrix_t* handles, int32_t count, int fmt, int variant, double alpha,
                              const void* const* xs, double beta, void* const* ys, int32_t iters,
                              double* ms_per_call) {
  return guarded([&] {
    LSP_REQUIRE(handles && xs && ys && count > 0 && iters > 0 && ms_per_call, LESPMV_ERR_ARG,
                "bad rotating-timing arguments");
    check_fmt(fmt);
    std::vector<Matrix*> mats(count);
    for (int i = 0; i < count; i++) {
      mats[i] = &mat(handles[i]);
      require_format(*mats[i], fmt);
      if (fmt == LESPMV_CSR) csr_prepare(*mats[i], variant, mats[i]->stream);
    }
    LSP_HIP(hipDeviceSynchronize());  // every handle's earlier calls are done
    Matrix& A0 = *mats[0];
    hipStream_t s = A0.stream;
    LSP_HIP(hipEventRecord(A0.ev0, s));
    for (int it = 0; it < iters; it++)
      for (int i = 0; i < count; i++) dispatch_spmv(*mats[i], fmt, variant, alpha, xs[i], beta, ys[i], s);
    LSP_HIP(hipEventRecord(A0.ev1, s));
    LSP_HIP(hipEventSynchronize(A0.ev1));
    float ms = 0;
    LSP_HIP(hipEventElapsedTime(&ms, A0.ev0, A0.ev1));
    for (Matrix* M : mats) M->last_stream = nullptr;  // every call above has completed
    *ms_per_call = ms / ((double)iters * count);
  });
}

int lespmv_time_spmm(lespmv_matrix_t h, int fmt, int variant, int32_t k, double alpha, const void* X, int64_t ldx,
                     double beta, void* Y, int64_t ldy, int32_t iters, double* ms_per_call) {
  return guarded([&] {
    Matrix& A = mat_any(h);
    check_fmt(fmt);
    LSP_REQUIRE(iters > 0 && ms_per_call, LESPMV_ERR_ARG, "iters must be > 0");
    hipStream_t s = A.stream;
    LSP_HIP(hipEventRecord(A.ev0, s));
    for (int i = 0; i < iters; i++) run_spmm(A, fmt, variant, k, alpha, X, ldx, beta, Y, ldy, s);
    LSP_HIP(hipEventRecord(A.ev1, s));
    LSP_HIP(hipEventSynchronize(A.ev1));
    float ms = 0;
    LSP_HIP(hipEventElapsedTime(&ms, A.ev0, A.ev1));
    *ms_per_call = ms / iters;
  });
}

int lespmv_effective_bytes(lespmv_matrix_t h, int fmt, int32_t k, int beta_nonzero, int64_t* bytes) {
  return guarded([&] {
    Matrix& A = mat_any(h);
    check_fmt(fmt);
    require_format(A, fmt);
    *bytes = effective_bytes(A, fmt, k < 1 ? 1 : k, beta_nonzero != 0);
  });
}

int lespmv_reference_bytes(lespmv_matrix_t h, int fmt, int64_t* bytes) {
  return guarded([&] {
    Matrix& A = mat(h);
    check_fmt(fmt);
    require_format(A, fmt);
    *bytes = reference_bytes(A, fmt);
  });
}

int lespmv_format_bytes(lespmv_matrix_t h, int fmt, int64_t* bytes) {
  return guarded([&] {
    Matrix& A = mat_any(h);
    check_fmt(fmt);
    require_format(A, fmt);
    *bytes = fmt == LESPMV_CSR ? (int64_t)(A.rp.bytes + A.ci.bytes + A.val.bytes) + csr_aux_bytes(A)
                               : A.fmt[fmt]->device_bytes();
    for (const auto& b : A.blocks)  // the row blocks of a handle given with 64-bit offsets
      *bytes += (int64_t)(b->rp.bytes + b->ci.bytes + b->val.bytes) + csr_aux_bytes(*b);
  });
}

int lespmv_export_size(lespmv_matrix_t h, int fmt, const char* name, int64_t* bytes) {
  return guarded([&] {
    Matrix& A = mat(h);
    check_fmt(fmt);
    LSP_REQUIRE(name && bytes, LESPMV_ERR_ARG, "null argument");
    std::vector<unsigned char> buf;
    export_array(A, fmt, name, buf);
    *bytes = (int64_t)buf.size();
  });
}

int lespmv_export(lespmv_matrix_t h, int fmt, const char* name, void* host_dst, int64_t bytes) {
  return guarded([&] {
    Matrix& A = mat(h);
    check_fmt(fmt);
    LSP_REQUIRE(name, LESPMV_ERR_ARG, "null name");
    std::vector<unsigned char> buf;
    export_array(A, fmt, name, buf);
    LSP_REQUIRE(bytes >= (int64_t)buf.size(), LESPMV_ERR_ARG, "destination too small for " + std::string(name));
    if (!buf.empty()) std::memcpy(host_dst, buf.data(), buf.size());
  });
}

int lespmv_malloc(void** dptr, int64_t bytes) {
  return guarded([&] {
    LSP_REQUIRE(dptr && bytes >= 0, LESPMV_ERR_ARG, "bad malloc arguments");
    LSP_HIP(hipMalloc(dptr, bytes ? bytes : 16));
  });
}
int lespmv_free(void* dptr) {
  return guarded([&] {
    if (dptr) LSP_HIP(hipFree(dptr));
  });
}
int lespmv_memcpy_h2d(void* dst, const void* src, int64_t bytes) {
  return guarded([&] {
    if (bytes) LSP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  });
}
int lespmv_memcpy_d2h(void* dst, const void* src, int64_t bytes) {
  return guarded([&] {
    if (bytes) LSP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  });
}
int lespmv_memset(void* dptr, int value, int64_t bytes) {
  return guarded([&] {
    if (bytes) LSP_HIP(hipMemset(dptr, value, bytes));
  });
}
int lespmv_synchronize(void) {
  return guarded([&] { LSP_HIP(hipDeviceSynchronize()); });
}
int lespmv_time_copy(void* dst, const void* src, int64_t bytes, int32_t iters, double* ms_per_copy) {
  return guarded([&] {
    LSP_REQUIRE(iters > 0 && ms_per_copy, LESPMV_ERR_ARG, "iters must be > 0");
    hipStream_t s = nullptr;
    hipEvent_t e0, e1;
    LSP_HIP(hipEventCreate(&e0));
    LSP_HIP(hipEventCreate(&e1));
    LSP_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < iters; i++) device_copy(dst, src, bytes, s);
    LSP_HIP(hipEventRecord(e1, s));
    LSP_HIP(hipEventSynchronize(e1));
    float ms = 0;
    LSP_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_copy = ms / iters;
  });
}

int lespmv_stream_create(lespmv_stream_t* s) {
  return guarded([&] {
    hipStream_t st;
    LSP_HIP(hipStreamCreate(&st));
    *s = st;
  });
}
int lespmv_stream_destroy(lespmv_stream_t s) {
  return guarded([&] {
    if (s) LSP_HIP(hipStreamDestroy(static_cast<hipStream_t>(s)));
  });
}
int lespmv_stream_synchronize(lespmv_stream_t s) {
  return guarded([&] { LSP_HIP(hipStreamSynchronize(static_cast<hipStream_t>(s))); });
}

}  // extern "C"
