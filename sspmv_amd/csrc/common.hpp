// Internal declarations of the gfx950 LeSpMV library (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lespmv_hip.h"

namespace lsp {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define LSP_HIP(call)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      throw ::lsp::Error(e_ == hipErrorOutOfMemory ? LESPMV_ERR_OOM : LESPMV_ERR_DEVICE,           \
                         std::string(#call) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" + \
                             std::to_string(__LINE__) + ")");                                      \
  } while (0)

#define LSP_REQUIRE(cond, code, msg)             \
  do {                                           \
    if (!(cond)) throw ::lsp::Error(code, msg);  \
  } while (0)

// Launch-error check after a kernel launch (asynchronous faults surface later).
#define LSP_LAUNCHED() LSP_HIP(hipGetLastError())

extern thread_local std::string g_last_error;

// Runs f, turning exceptions into a status code and lespmv_last_error() text.
template <class F>
int guarded_call(F&& f) {
  try {
    f();
    return LESPMV_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return LESPMV_ERR_OOM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return LESPMV_ERR_DEVICE;
  }
}

// Library-owned device buffer.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    reset();
    p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0;
    return *this;
  }
  ~DBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t nbytes) {
    reset();
    if (nbytes == 0) nbytes = 16;  // keep a valid pointer for empty arrays
    LSP_HIP(hipMalloc(&p, nbytes));
    bytes = nbytes;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

inline size_t vsize(int dtype) { return dtype == LESPMV_F64 ? 8 : 4; }

// --- per-format device data ---------------------------------------------------

struct FormatData {
  virtual ~FormatData() = default;
  lespmv_params_t params{};
  double convert_ms = 0;
  virtual int64_t device_bytes() const = 0;
};

// Row-sorted COO: column indices and values are the CSR arrays (same order),
// only the row index array is added.
struct CooData : FormatData {
  DBuf row;
  int64_t nnz = 0;
  int64_t device_bytes() const override { return row.bytes; }
};

// Sliced ELL family (ELL = one chunk per 64 rows with a single global width;
// S-ELL = sigma 1; SELL-c-sigma; SELL-c-R = sigma >= M). Device layout is
// column-major inside a chunk: slot (row r, position j) of chunk c is at
// chunk_off[c] + j*C + r, so a wavefront (C = 64) loads one position of 64 rows
// with unit stride.
struct SlicedData : FormatData {
  int fmt = 0;
  int32_t C = 0, sigma = 1, alignment = 1;
  int64_t nchunks = 0, slots = 0;
  int32_t max_width = 0;      // ELL: max_row_width
  DBuf chunk_len;             // int32[nchunks]
  DBuf chunk_off;             // int64[nchunks+1]
  DBuf reorder;               // int32[M] (reordered row -> original row); empty when not sorted
  DBuf col, val;              // int32[slots], V[slots]
  int64_t device_bytes() const override {
    return chunk_len.bytes + chunk_off.bytes + reorder.bytes + col.bytes + val.bytes;
  }
};

struct DiaData : FormatData {
  int32_t ndiag = 0;
  int64_t stride = 0;
  int64_t min_off = 0, max_off = 0;
  DBuf offsets;  // int64[ndiag], first-appearance order
  DBuf data;     // V[ndiag*stride]
  int64_t device_bytes() const override { return offsets.bytes + data.bytes; }
};

struct BsrData : FormatData {
  int32_t br = 0, bc = 0;
  int64_t mb = 0, nb = 0, nnzb = 0;
  DBuf row_ptr;  // int32[mb+1]
  DBuf bcol;     // int32[nnzb]
  DBuf data;     // V[nnzb*br*bc], row-major inside a block
  int64_t device_bytes() const override { return row_ptr.bytes + bcol.bytes + data.bytes; }
};

struct Csr5Data : FormatData {
  int32_t omega = 0, sigma = 0, bit_y = 0, bit_scan = 0, packets = 0;
  int64_t p = 0, num_offsets = 0, tail_start = 0, num_empty = 0;
  DBuf tile_ptr;      // uint32[p+1]
  DBuf tile_desc;     // uint32[p*omega*packets]
  DBuf desc_off_ptr;  // int32[p+1]
  DBuf desc_off;      // int32[num_offsets]
  DBuf col, val;      // tiles 0..p-2 transposed (sigma-step major), tail in CSR order
  DBuf empty_rows;    // int32[num_empty]
  DBuf calib;         // V[p] per-tile carry of the row open at the tile start
  int64_t device_bytes() const override {
    return tile_ptr.bytes + tile_desc.bytes + desc_off_ptr.bytes + desc_off.bytes + col.bytes + val.bytes +
           empty_rows.bytes + calib.bytes;
  }
};

// CSR analysis used by the CSR kernels.
struct CsrMeta {
  int64_t max_row = 0;
  double avg_row = 0;
  int32_t lanes = 0;         // tuned lanes-per-row of the vector kernel
  int32_t auto_variant = 1;  // tuned CSR kernel
  double far_frac = 0;       // far x lines per entry (entries > 2 MB of x off the diagonal opening a new line)
  int64_t merge_tiles = 0;
  // tiles cut at row starts (variant ROW_TILES): no carries between tiles
  int64_t row_tiles = 0;
  DBuf row_coord, row_carry_row, row_carry_val, row_head;
  DBuf merge_coord;          // int32[2*(tiles+1)] (row, nnz) at each tile start
  DBuf carry_row;            // int32[tiles]  row open at the tile end (-1: none)
  DBuf carry_val;            // V[tiles]      its partial sum inside the tile
  DBuf head_val;             // V[tiles]      partial of the continued first row
  bool analysed = false;
  // column panels (CSR panel variants): P column ranges, each its own CSR
  int32_t panels = 0;
  int64_t p_maxT = 0;
  DBuf p_nzoff, p_rp, p_ci, p_val, p_tiles, p_coord, p_carry_row, p_carry_val, p_head, p_partial, p_ends, p_endoff;
  int64_t p_width = 0;       // columns per panel
  bool p_compact = false;    // FP64 columns kept as 16 bits + a nibble (p_width <= 2^20), p_ci freed
  DBuf p_ci16, p_cihi;
};

struct Matrix {
  int dtype = LESPMV_F64;
  int device = 0;
  int64_t m = 0, n = 0, nnz = 0;
  hipStream_t stream = nullptr;
  // The handle's device scratch (merge-path carries, panel partials, fix-up
  // partials of the other formats) serves one call at a time: a call on
  // another stream than the previous one first waits for that stream.
  hipStream_t last_stream = nullptr;
  hipEvent_t ev_switch = nullptr;
  double prepare_ms = 0;  // device time of the last CSR prepare (panel layout)
  DBuf rp, ci, val;  // CSR: int32[m+1], int32[nnz], V[nnz]
  // A matrix given with 64-bit row offsets (lespmv_create_csr64, beyond
  // 2^31 - 1 nonzeros) is held as row blocks of fewer than 2^31 nonzeros,
  // each an ordinary handle with 32-bit offsets relative to its first entry;
  // block b holds rows [block_row[b], block_row[b + 1]). Such a handle runs
  // CSR SpMV / SpMM only.
  bool idx64 = false;
  std::vector<std::unique_ptr<Matrix>> blocks;
  std::vector<int64_t> block_row;
  // int32[m]: the rows the CSR row engine (SpMM variants 16..22) takes, by
  // position; empty = natural order. Set by the serving layer for matrices
  // whose rows differ in length; results do not depend on it.
  DBuf row_order;
  CsrMeta meta;
  std::unique_ptr<FormatData> fmt[LESPMV_NUM_FORMATS];
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ~Matrix();
};

// --- launchers (one translation unit per format family) ----------------------

void csr_analyse(Matrix& A, hipStream_t s);
// Builds what a CSR kernel variant needs beyond the CSR arrays (the merge-path
// partition; the column-panel layout of a PANEL variant). Synchronous.
void csr_prepare(Matrix& A, int variant, hipStream_t s);
void csr_release_aux(Matrix& A);
// Merge-path CSR pieces on plain device pointers: the tile partition (tiles
// of kMergeItems = 2048 merge-path items, coord int32[2 (tiles + 1)]) and the
// tiles with their carry fix-up (carry_row int32[tiles], carry_val and
// head_val V[tiles]).
constexpr int kMergeItems = 2048;
constexpr int kMergeThreads = 256;  // threads of a merge-path tile's workgroup
void csr_merge_partition(int64_t tiles, int64_t m, int64_t nnz, const int* rp, int* coord, hipStream_t s);
void csr_merge_tiles(int dtype, int64_t m, int64_t n, int64_t tiles, const int* rp, const int* ci, const void* val,
                     const void* x, void* y, double alpha, double beta, const int* coord, int* carry_row,
                     void* carry_val, void* head_val, hipStream_t s);
// Device address of the allocation-free SpMV's fixed scratch on `device`
// (noalloc.hip), looked up on first use with that device current.
void* noalloc_scratch(int device);
int64_t csr_aux_bytes(const Matrix& A);
void csr_spmv(Matrix& A, int variant, double alpha, const void* x, double beta, void* y, hipStream_t s);
void csr_spmm(Matrix& A, int variant, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
              int64_t ldy, hipStream_t s);

void convert_coo(Matrix& A, const lespmv_params_t& p, hipStream_t s);
void coo_spmv(Matrix& A, int variant, double alpha, const void* x, double beta, void* y, hipStream_t s);
void coo_spmm(Matrix& A, int variant, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
              int64_t ldy, hipStream_t s);

void convert_sliced(Matrix& A, int fmt, const lespmv_params_t& p, hipStream_t s);
void sliced_spmv(Matrix& A, int fmt, double alpha, const void* x, double beta, void* y, hipStream_t s);
void sliced_spmm(Matrix& A, int fmt, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
                 int64_t ldy, hipStream_t s);
void sliced_export(Matrix& A, int fmt, const std::string& name, std::vector<unsigned char>& out);

void convert_dia(Matrix& A, const lespmv_params_t& p, hipStream_t s);
void dia_spmv(Matrix& A, int variant, double alpha, const void* x, double beta, void* y, hipStream_t s);
void dia_spmm(Matrix& A, int variant, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
              int64_t ldy, hipStream_t s);

void convert_bsr(Matrix& A, const lespmv_params_t& p, hipStream_t s);
void bsr_spmv(Matrix& A, int variant, double alpha, const void* x, double beta, void* y, hipStream_t s);
void bsr_spmm(Matrix& A, int variant, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y,
              int64_t ldy, hipStream_t s);

void convert_csr5(Matrix& A, const lespmv_params_t& p, hipStream_t s);
void csr5_spmv(Matrix& A, double alpha, const void* x, double beta, void* y, hipStream_t s);
void csr5_spmm(Matrix& A, int k, double alpha, const void* X, int64_t ldx, double beta, void* Y, int64_t ldy,
               hipStream_t s);
void csr5_export(Matrix& A, const std::string& name, std::vector<unsigned char>& out);

// Generic device primitives (primitives.hip).
// Exclusive scan of int64 values in place; returns the total.
int64_t scan_exclusive_i64(int64_t* d, int64_t n, hipStream_t s);
int64_t scan_exclusive_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, hipStream_t s);
// Stable LSD radix sort of (key, value) uint32 pairs by the low `bits` key bits.
void radix_sort_pairs(uint32_t* keys, uint32_t* vals, int64_t n, int bits, hipStream_t s);
// LSD radix sort of uint32 keys by their low `bits` bits.
void radix_sort_u32(uint32_t* keys, int64_t n, int bits, hipStream_t s);
// Stable LSD radix sort of uint64 keys by their low `bits` bits.
void radix_sort_u64(uint64_t* keys, int64_t n, int bits, hipStream_t s);
int64_t reduce_max_rowlen(const int32_t* rp, int64_t m, hipStream_t s);
void device_copy(void* dst, const void* src, int64_t bytes, hipStream_t s);
void coo_to_csr_device(int dtype, int64_t m, int64_t nnz, const int32_t* ri, const int32_t* ci, const void* v,
                       DBuf& rp_out, DBuf& ci_out, DBuf& v_out, hipStream_t s);

// Refuses a padded layout of `slots` stored entries beyond p.max_fill * nnz + m.
inline void check_fill(const lespmv_params_t& p, int64_t slots, int64_t nnz, int64_t m, const char* what) {
  if (p.max_fill > 0 && (double)slots > (double)p.max_fill * (double)nnz + (double)m)
    throw Error(LESPMV_ERR_UNSUPPORTED, std::string(what) + ": " + std::to_string(slots) +
                                            " stored slots exceed the fill limit of " + std::to_string(p.max_fill) +
                                            " x nnz");
}

// Host-side copies.
template <class T>
std::vector<T> to_host(const void* d, size_t count, hipStream_t s) {
  std::vector<T> h(count);
  if (count) {
    LSP_HIP(hipMemcpyAsync(h.data(), d, count * sizeof(T), hipMemcpyDeviceToHost, s));
    LSP_HIP(hipStreamSynchronize(s));
  }
  return h;
}
template <class T>
void put_bytes(std::vector<unsigned char>& out, const std::vector<T>& v) {
  out.resize(v.size() * sizeof(T));
  if (!v.empty()) std::memcpy(out.data(), v.data(), out.size());
}
inline void put_scalar(std::vector<unsigned char>& out, int64_t v) { put_bytes(out, std::vector<int64_t>{v}); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t align_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

// MI355X: 8 XCDs of 32 CUs each with its own L2; the dispatcher deals
// workgroups to the XCDs round-robin, so workgroup b runs on XCD b % 8.
constexpr int kXcds = 8;

// Logical block of hardware workgroup b in a grid of nb: each XCD takes one
// contiguous eighth of the logical blocks, so the rows one XCD works on (and
// the x lines they share with their neighbours) stay in that XCD's L2 instead
// of being fetched by all eight. A bijection on [0, nb). Only for layouts
// whose work per block does not trend along the rows (a layout sorted by row
// length would put all long rows on one XCD).
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t q = nb / kXcds, r = nb % kXcds, x = b % kXcds, i = b / kXcds;
  return x * q + (x < r ? x : r) + i;
}

// Grid size helper (1D, capped to the hardware limit).
inline dim3 grid1d(int64_t work, int block) {
  int64_t g = ceil_div(work, block);
  if (g < 1) g = 1;
  return dim3((unsigned)std::min<int64_t>(g, 0x7FFFFFFF));
}

}  // namespace lsp
