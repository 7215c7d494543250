// The CSR SpMV entry point that allocates nothing.
//
// No counterpart in the reference: callers that cannot call hipMalloc (code
// under HIP graph capture, tight serving loops, harnesses that forbid it) pass
// their own device CSR arrays, x and y, and the merge-path kernels of
// LESPMV_VARIANT_MERGE (spmv_csr.hip) run with their tile partition and
// carries in `scratch`, a device array of fixed size in this file.
//
// Layout of `scratch` for T = kTiles tiles of kMergeItems merge-path items:
// int32 coord[2 (T + 1)] (padded to 256 bytes), int32 carry_row[T], then
// carry_val[T] and head_val[T] as FP64 slots (FP32 uses the first half of each
// slot array). Every entry a call reads was written earlier in the same call
// (each tile writes its carry_row, and its head_val whenever the fix-up reads
// it), so nothing needs clearing between calls. Each GPU holds its own copy of
// the array until the process ends; host code reaches it only through the
// address hipGetSymbolAddress gives on that device, never through its name.
// Calls that overlap in time on one GPU share the array and race.
#include <mutex>

#include "common.hpp"

namespace lsp {
namespace {

constexpr int64_t kTiles = 131072;  // m + nnz <= 268,435,456 merge-path items
constexpr size_t kCoordBytes = ((sizeof(int) * 2 * (kTiles + 1) + 255) / 256) * 256;
constexpr size_t kCarryRowBytes = sizeof(int) * kTiles;
constexpr size_t kScratchBytes = kCoordBytes + kCarryRowBytes + 2 * sizeof(double) * kTiles;  // 3,670,272
static_assert(kScratchBytes == 3670272, "scratch size stated in include/lespmv_hip.h");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

// The device address of this GPU's `scratch`, looked up once per device with
// that device current.
void* noalloc_scratch(int device) {
  LSP_REQUIRE(device >= 0 && device < kMaxDevices, LESPMV_ERR_ARG, "device id out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    LSP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    LSP_REQUIRE(p != nullptr, LESPMV_ERR_DEVICE, "hipGetSymbolAddress gave no address for the SpMV scratch");
    g_ptr[device] = p;
  }
  return g_ptr[device];
}

}  // namespace lsp

using namespace lsp;

extern "C" {

int64_t lespmv_csr_noalloc_max_items(void) { return kTiles * kMergeItems; }

int lespmv_csr_spmv_noalloc(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* row_offset,
                            const int32_t* col_index, const void* values, double alpha, const void* x, double beta,
                            void* y, lespmv_stream_t stream) {
  return guarded_call([&] {
    LSP_REQUIRE(dtype == LESPMV_F32 || dtype == LESPMV_F64, LESPMV_ERR_ARG, "dtype must be F32 or F64");
    LSP_REQUIRE(m >= 0 && n >= 0 && nnz >= 0, LESPMV_ERR_ARG, "negative size");
    LSP_REQUIRE(m < (int64_t)INT32_MAX && n < (int64_t)INT32_MAX && nnz < (int64_t)INT32_MAX,
                LESPMV_ERR_UNSUPPORTED, "sizes must fit 32-bit indices");
    const int64_t tiles = ceil_div(m + nnz, kMergeItems);
    LSP_REQUIRE(tiles <= kTiles, LESPMV_ERR_UNSUPPORTED,
                "allocation-free SpMV: m + nnz = " + std::to_string(m + nnz) +
                    " exceeds its fixed scratch limit of " + std::to_string(kTiles * kMergeItems) +
                    " merge-path items");
    LSP_REQUIRE(row_offset && (nnz == 0 || (col_index && values)) && (x || n == 0) && (y || m == 0),
                LESPMV_ERR_ARG, "null CSR array or vector");
    if (m == 0) return;
    int dev = 0;
    LSP_HIP(hipGetDevice(&dev));
    unsigned char* base = static_cast<unsigned char*>(noalloc_scratch(dev));
    int* coord = reinterpret_cast<int*>(base);
    int* carry_row = reinterpret_cast<int*>(base + kCoordBytes);
    unsigned char* slots = base + kCoordBytes + kCarryRowBytes;
    // carry_val and head_val: kTiles FP64 slots each (kTiles * 8 bytes apart)
    void* carry_val = slots;
    void* head_val = slots + sizeof(double) * kTiles;
    hipStream_t s = static_cast<hipStream_t>(stream);
    csr_merge_partition(tiles, m, nnz, row_offset, coord, s);
    csr_merge_tiles(dtype, m, n, tiles, row_offset, col_index, values, x, y, alpha, beta, coord, carry_row, carry_val,
                    head_val, s);
  });
}

}  // extern "C"
