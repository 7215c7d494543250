/*
 * lespmv_hip.h - C ABI of the MI355X (gfx950) LeSpMV SpMV hot path.
 *
 * This is the drop-in boundary. The C++ side (include/lespmv/LeSpMV.h) keeps the
 * reference's struct and function names on top of it; Python reaches it via
 * ctypes (sspmv_amd/lib.py). Signatures use plain pointers and sizes only.
 *
 * Reference interfaces each entry point replaces (paths relative to LeSpMV/):
 *   lespmv_create_csr      CSR_Matrix<I,V> construction, include/sparse_format.h:95-104,
 *                          read_csr_matrix, utils/sparse_io.cpp:179-204 (the resulting struct)
 *   lespmv_destroy         delete_host_matrix / delete_<fmt>_matrix, include/sparse_format.h:300-449
 *   lespmv_convert         csr_to_coo / csr_to_ell / csr_to_sell / csr_to_sell_c_sigma /
 *                          csr_to_sell_c_R / csr_to_dia / csr_to_bsr / csr_to_csr5,
 *                          include/sparse_conversion.h:144-1379
 *   lespmv_spmv            LeSpMV_csr (src/spmv_csr.cpp:164-185), LeSpMV_coo (src/spmv_coo.cpp:160-182),
 *                          LeSpMV_ell (src/spmv_ell.cpp:224-247), LeSpMV_sell (src/spmv_s_ell.cpp:202-224),
 *                          LeSpMV_sell_c_sigma (src/spmv_sell_c_sigma.cpp:189-211),
 *                          LeSpMV_sell_c_R (src/spmv_sell_c_R.cpp:189-211), LeSpMV_dia
 *                          (src/spmv_dia.cpp:176-197), LeSpMV_bsr (src/spmv_bsr.cpp:214-233),
 *                          LeSpMV_csr5 (src/spmv_csr5.cpp:438-457); variant = kernel_flag
 *   lespmv_spmm            no reference counterpart (multi-vector extension of lespmv_spmv)
 *   lespmv_export*         the struct fields of include/sparse_format.h:76-295, read back to host
 *   lespmv_effective_bytes bytes model replacing bytes_per_spmv, include/spmv_benchmark.h:26-150
 *   lespmv_generate        no reference counterpart (synthetic inputs; the reference reads .mtx)
 *
 * Errors: every call returns a status (LESPMV_OK = 0); lespmv_last_error()
 * gives the message of the last failure on the calling thread. Nothing exits.
 * Threading: a matrix handle may be used from one thread at a time; different
 * handles are independent. Every compute call takes a stream (NULL = the
 * handle's own stream) and is asynchronous unless stated otherwise. A handle's
 * device scratch (merge-path and panel partials, fix-up carries) serves one
 * call at a time: a call on another stream than the handle's previous call is
 * ordered after that call (an event recorded on the caller's stream after each
 * call there, waited for by the next stream). A caller's stream must stay alive
 * until the handle's next call has been issued.
 */
#ifndef LESPMV_HIP_H
#define LESPMV_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lespmv_matrix_s* lespmv_matrix_t;
typedef void* lespmv_stream_t; /* a hipStream_t; NULL selects the handle's stream */

/* Format ids: the reference's label map, LeSpMV/script/write_label.py:63-75. */
enum {
  LESPMV_COO = 0,
  LESPMV_CSR = 1,
  LESPMV_DIA = 2,
  LESPMV_ELL = 3,
  LESPMV_SELL = 4,         /* S-ELL: sliced ELL, rows in natural order */
  LESPMV_SELL_C_SIGMA = 5,
  LESPMV_SELL_C_R = 6,
  LESPMV_CSR5 = 7,
  LESPMV_BSR = 8,
  LESPMV_NUM_FORMATS = 9
};

enum { LESPMV_F32 = 0, LESPMV_F64 = 1 };

enum {
  LESPMV_OK = 0,
  LESPMV_ERR_ARG = 1,           /* bad argument or handle */
  LESPMV_ERR_OOM = 2,           /* device allocation failed */
  LESPMV_ERR_UNSUPPORTED = 3,   /* shape/parameter not supported (e.g. sigma % C != 0) */
  LESPMV_ERR_MAX_DIAGS = 4,     /* DIA: more diagonals than max_diags */
  LESPMV_ERR_DEVICE = 5,        /* HIP runtime error */
  LESPMV_ERR_NOT_CONVERTED = 6  /* format requested before lespmv_convert */
};

/* Kernel variants (the GPU meaning of the reference's kernel_flag/kernel_tag). */
enum {
  LESPMV_VARIANT_AUTO = 0,      /* the handle's tuned choice; CSR: PANEL when x spans at least 16 MB and
                                   there are 0.3 far x lines per entry or more (entries more than 2 MB of x off
                                   the diagonal that open a new 128-byte line of x in their row); MERGE when the
                                   longest row exceeds 32 x max(mean, 4), or when the matrix exceeds 256 MB
                                   (the last-level cache) and its rows average at most 32 entries or the
                                   longest exceeds 8 x the mean; VECTOR otherwise (lespmv_csr_plan) */
  LESPMV_VARIANT_VECTOR = 1,    /* CSR: lanes-per-row vector kernel; DIA: x through caches */
  LESPMV_VARIANT_MERGE = 2,     /* CSR: merge-path tiles; DIA: LDS-staged x windows */
  LESPMV_VARIANT_ATOMIC = 3,    /* COO: atomics instead of the carry fix-up pass */
  LESPMV_VARIANT_PANEL = 4,     /* CSR: 8 column panels, one per XCD (x slice held in that XCD's L2),
                                   merge-path per panel, partial y reduced in panel order */
  LESPMV_VARIANT_PANEL16 = 5,   /* CSR: the same with 16 panels (two per XCD) */
  LESPMV_VARIANT_PANEL4 = 6,    /* CSR: 4 panels, each run by two XCDs; the 4 MB L2 of an XCD holds an
                                   x slice when x is about 16 MB (FP32 x of 4M columns), as 8 panels
                                   do for 32 MB (FP64) */
  LESPMV_VARIANT_ROW_TILES = 7  /* CSR: merge-path tiles cut at row starts, so no row spans two tiles and
                                   no carry fix-up pass runs (rows of at most 511 entries; else MERGE) */
};

/* Conversion parameters. Fields a format does not use are ignored. */
typedef struct {
  int32_t chunk;       /* C: rows per chunk (S-ELL, SELL-c-sigma, SELL-c-R) */
  int32_t sigma;       /* SELL-c-sigma sort window; CSR5 sigma (nonzeros per lane) */
  int32_t alignment;   /* chunk_len rounding (sliced formats), DIA stride rounding */
  int32_t ld;          /* ELL layout of the exported arrays: 0 RowMajor, 1 ColMajor */
  int32_t max_diags;   /* DIA limit (reference MAX_DIAG_NUM = 10240) */
  int32_t block_rows;  /* BSR br */
  int32_t block_cols;  /* BSR bc */
  int32_t omega;       /* CSR5 omega (lanes per tile) */
  int32_t max_fill;    /* > 0: refuse (LESPMV_ERR_UNSUPPORTED) an ELL/sliced/DIA/BSR layout storing
                          more than max_fill * nnz + num_rows slots; 0: no limit */
} lespmv_params_t;

/* Host-side CSR produced by lespmv_generate / lespmv_read_mtx /
 * lespmv_coo_to_csr; arrays owned by the library (lespmv_host_csr_free). */
typedef struct {
  int64_t m, n, nnz;
  int32_t dtype;
  int32_t mm_flags; /* LESPMV_MM_SYMMETRIC when read from a symmetric MatrixMarket file */
  int32_t* row_offset;
  int32_t* col_index;
  void* values;
} lespmv_host_csr_t;

/* Synthetic matrix families for lespmv_generate (params as doubles). */
enum {
  LESPMV_GEN_LAPLACE2D = 0,   /* nx, ny, diag, off: 5-point natural order */
  LESPMV_GEN_STENCIL27 = 1,   /* n, diag, off: 27-point on n^3 */
  LESPMV_GEN_LOGNORMAL = 2,   /* m, n, mean, sigma, cap, frac_local, band */
  LESPMV_GEN_BLOCKRAND = 3,   /* mb, nb, br, bc, blocks_per_row */
  LESPMV_GEN_BANDED = 4,      /* m, half_bandwidth, fill */
  LESPMV_GEN_POWERLAW = 5,    /* m, zipf_exponent, mean_nnz, max_nnz */
  LESPMV_GEN_BLOCKDIAG = 6,   /* m, block */
  LESPMV_GEN_UNIFORM = 7      /* m, n, density */
};

const char* lespmv_last_error(void);
int lespmv_init(int device);
int lespmv_device_info(int64_t* hbm_bytes, int32_t* num_cu, int32_t* arch_gfx);

/* Reference (reference_layout=1) or GPU-native (0) default parameters. */
int lespmv_default_params(int fmt, int dtype, int reference_layout, lespmv_params_t* out);

/* Matrix lifecycle. Arrays are copied into library-owned device memory.
 * on_device = 1 means the three pointers are device pointers. */
int lespmv_create_csr(lespmv_matrix_t* out, int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* row_offset,
                      const int32_t* col_index, const void* values, int on_device);
/* The same with 64-bit row offsets, for matrices beyond 2^31 - 1 nonzeros
 * (CSR_Matrix<long long, V>, src/spmv_csr.cpp:187-193 instantiates it; rows
 * and columns stay below 2^31, column indices are 32-bit). The matrix is held
 * as row blocks of fewer than 2^31 nonzeros with 32-bit offsets each. Such a
 * handle runs lespmv_spmv / lespmv_spmm of LESPMV_CSR (every CSR variant, block
 * by block), their timing, lespmv_matrix_info and the byte counts; every other
 * call on it returns LESPMV_ERR_UNSUPPORTED. */
int lespmv_create_csr64(lespmv_matrix_t* out, int dtype, int64_t m, int64_t n, int64_t nnz, const int64_t* row_offset,
                        const int32_t* col_index, const void* values, int on_device);
int lespmv_destroy(lespmv_matrix_t A);
int lespmv_matrix_info(lespmv_matrix_t A, int64_t* m, int64_t* n, int64_t* nnz, int32_t* dtype);

/* Build format `fmt` on the device from the handle's CSR (GPU kernels). */
int lespmv_convert(lespmv_matrix_t A, int fmt, const lespmv_params_t* p);
/* Frees format `fmt`; for LESPMV_CSR the prepared column-panel layout (the CSR arrays stay). */
int lespmv_release_format(lespmv_matrix_t A, int fmt);
int lespmv_has_format(lespmv_matrix_t A, int fmt);
/* Builds ahead of the first call what kernel `variant` of `fmt` needs beyond
 * the format's arrays, so that the calls themselves allocate nothing and stay
 * asynchronous. CSR (LeSpMV_csr, src/spmv_csr.cpp:164-185, keeps a partition
 * made on first use, :149-153): the row analysis and merge-path partition,
 * and for a PANEL variant (or AUTO when it picks one) the column-panel layout
 * (about nnz * (sV + 2.5) + P * M * sV bytes). Other formats: nothing.
 * Synchronous; lespmv_convert_ms(A, LESPMV_CSR) then gives its device time. A
 * PANEL call without it builds the layout on its first call (synchronously). */
int lespmv_prepare(lespmv_matrix_t A, int fmt, int variant);
/* What CSR variant AUTO runs for this handle, the vector kernel's lanes per
 * row, and the far x lines per entry the choice was made from (runs the row
 * analysis if no call has yet; synchronous). No reference counterpart: the
 * reference's LeSpMV_csr runs the kernel_flag it is given. */
int lespmv_csr_plan(lespmv_matrix_t A, int32_t* variant, int32_t* lanes, double* far_lines);
/* Device milliseconds of the last lespmv_convert of `fmt` (hipEvent timed);
 * for LESPMV_CSR the last lespmv_prepare. */
int lespmv_convert_ms(lespmv_matrix_t A, int fmt, double* ms);

/* y = alpha*A*x + beta*y; x, y are device pointers of the handle's dtype.
 * beta == 0 never reads y. */
int lespmv_spmv(lespmv_matrix_t A, int fmt, int variant, double alpha, const void* x, double beta, void* y,
                lespmv_stream_t stream);
/* Y = alpha*A*X + beta*Y for k right-hand sides, X row-major N x k (leading
 * dimension ldx >= k), Y row-major M x k (ldy >= k), device pointers.
 * CSR: variant 16 + log2(L) (L = 1..64 lanes per row) runs the row engine of
 * the serving kernel (several right-hand sides per lane, eight gathers in
 * flight per lane; bitwise the result the server gives for the same request),
 * any other variant the plain kernel (a group of k lanes per row, one
 * right-hand side per lane). */
int lespmv_spmm(lespmv_matrix_t A, int fmt, int variant, int32_t k, double alpha, const void* X, int64_t ldx,
                double beta, void* Y, int64_t ldy, lespmv_stream_t stream);

/* Average device time in ms of `iters` back-to-back lespmv_spmv calls,
 * measured with hipEvents on the stream the kernels run on. Synchronous. */
int lespmv_time_spmv(lespmv_matrix_t A, int fmt, int variant, double alpha, const void* x, double beta, void* y,
                     int32_t iters, double* ms_per_call);
/* Average device ms per call of `iters` rounds over `count` handles of one
 * format, each round one SpMV per handle (handles[i] with xs[i], ys[i]) in
 * turn, all on handles[0]'s stream: matrices that together exceed the 256 MB
 * last-level cache are then read from HBM on every call, as a single small
 * matrix repeated would not be. Synchronous (waits for the device first). */
int lespmv_time_spmv_rotating(const lespmv_matrix_t* handles, int32_t count, int fmt, int variant, double alpha,
                              const void* const* xs, double beta, void* const* ys, int32_t iters,
                              double* ms_per_call);
int lespmv_time_spmm(lespmv_matrix_t A, int fmt, int variant, int32_t k, double alpha, const void* X, int64_t ldx,
                     double beta, void* Y, int64_t ldy, int32_t iters, double* ms_per_call);

/* Allocation-free CSR SpMV. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). y = alpha*A*x + beta*y with
 * the merge-path kernels of LESPMV_VARIANT_MERGE on the caller's device CSR
 * arrays (row_offset[m+1], col_index[nnz], values[nnz]) and x, y; bitwise the
 * result of lespmv_spmv(LESPMV_CSR, LESPMV_VARIANT_MERGE) on the same matrix.
 * Calls no allocator: the tile partition and carries live in a fixed device
 * array of 3,670,272 bytes per GPU (held until the process ends), which limits
 * m + nnz to lespmv_csr_noalloc_max_items() = 268,435,456 merge-path items;
 * larger inputs return LESPMV_ERR_UNSUPPORTED. stream NULL = the null stream.
 * Every call on one GPU shares that array: calls that overlap in time on one
 * GPU (other streams, other threads) race and must be serialised by the
 * caller. lespmv_init(device) looks the array up on that device; otherwise the
 * first call on a device does. */
int lespmv_csr_spmv_noalloc(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* row_offset,
                            const int32_t* col_index, const void* values, double alpha, const void* x, double beta,
                            void* y, lespmv_stream_t stream);
int64_t lespmv_csr_noalloc_max_items(void);

/* Algorithmic bytes of one SpMV (k = 1) or SpMM with this format: values +
 * indices + x read once + y written (+ y read if beta != 0). Padding and
 * repeated x gathers excluded. */
int lespmv_effective_bytes(lespmv_matrix_t A, int fmt, int32_t k, int beta_nonzero, int64_t* bytes);
/* The reference's bytes_per_spmv formula (include/spmv_benchmark.h:26-150). */
int lespmv_reference_bytes(lespmv_matrix_t A, int fmt, int64_t* bytes);
/* Device bytes held by format `fmt` (LESPMV_CSR: the CSR arrays plus the
 * prepared partition and panel layout). */
int lespmv_format_bytes(lespmv_matrix_t A, int fmt, int64_t* bytes);

/* Read back a format's arrays in the reference's layout (named as the
 * reference's struct fields, e.g. "reorder", "chunk_len", "col_index",
 * "tile_desc"; scalars as one int64). Synchronous. */
int lespmv_export_size(lespmv_matrix_t A, int fmt, const char* name, int64_t* bytes);
int lespmv_export(lespmv_matrix_t A, int fmt, const char* name, void* host_dst, int64_t bytes);

/* Device memory and stream helpers for callers without their own runtime. */
int lespmv_malloc(void** dptr, int64_t bytes);
int lespmv_free(void* dptr);
int lespmv_memcpy_h2d(void* dst, const void* src, int64_t bytes);
int lespmv_memcpy_d2h(void* dst, const void* src, int64_t bytes);
int lespmv_memset(void* dptr, int value, int64_t bytes);
int lespmv_synchronize(void);
/* Time `iters` copies of `bytes` (multiple of 16) with the library's plain
 * streaming copy kernel (reference point for HBM counters). Synchronous. */
int lespmv_time_copy(void* dst, const void* src, int64_t bytes, int32_t iters, double* ms_per_copy);
int lespmv_stream_create(lespmv_stream_t* s);
int lespmv_stream_destroy(lespmv_stream_t s);
int lespmv_stream_synchronize(lespmv_stream_t s);

#define LESPMV_MM_SYMMETRIC 1

/* MatrixMarket input, replacing read_csr_matrix (utils/sparse_io.cpp:179-204):
 * coordinate real/integer/pattern, symmetric expansion, then coo_to_csr with
 * optional duplicate compaction. Host arrays owned by the library. */
int lespmv_read_mtx(const char* path, int dtype, int compact, lespmv_host_csr_t* out);
/* coo_to_csr (include/sparse_conversion.h:29-91): stable by row; compact = 0
 * runs as a GPU radix sort, compact = 1 sums duplicates like
 * sum_csr_duplicates (include/sparse_operation.h:21-73). Host input/output. */
int lespmv_coo_to_csr(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* row_index,
                      const int32_t* col_index, const void* values, int compact, lespmv_host_csr_t* out);

/* Synthetic inputs (host). */
int lespmv_generate(int family, const double* params, int32_t nparams, uint64_t seed, int dtype,
                    lespmv_host_csr_t* out);
int lespmv_host_csr_free(lespmv_host_csr_t* csr);
/* x[i] = lo + (hi-lo) * u(splitmix64(seed ^ i)), u = (h >> 11) * 2^-53. */
int lespmv_fill_uniform(void* host_dst, int dtype, int64_t n, uint64_t seed, double lo, double hi);

/* ---- exhaustive timing and autotuning ------------------------------------------
 * No reference counterpart (the reference times one format per driver run,
 * test/benchmark_spmv_<fmt>.cpp, and collects the results in spreadsheets).
 * A candidate is a (format, variant, parameters) triple; lespmv_tune converts
 * each one on the device, times `iters` SpMV (k = 1) or SpMM (k <= 32) calls
 * after two warm-up calls with hipEvents, and releases it again. Candidates
 * that do not apply (too many diagonals, padding beyond params.max_fill * nnz,
 * CSR5 limits) get a non-zero status and ms = +inf. */
typedef struct {
  int32_t fmt;
  int32_t variant;
  lespmv_params_t params;
  int32_t status;     /* LESPMV_OK or the conversion / timing error */
  double ms;          /* device ms per call */
  double convert_ms;  /* device ms of the conversion */
} lespmv_tune_result_t;

/* The standard grid for the formats whose bit is set in fmt_mask (bit = format
 * id): CSR lanes per row 2..64 and merge-path, COO, ELL, S-ELL / SELL-C-sigma
 * (sigma 256, 4096) / SELL-C-R with C in {32, 64}, DIA both variants, BSR
 * 2x2 4x4 8x4 8x8 16x4, CSR5 (omega, sigma) in {(64,16), (32,16), (64,8)};
 * with LESPMV_TUNE_PANELS also the CSR column panels (8 and 4 panels), after
 * all the others (their layout is built with lespmv_prepare, timed as the
 * candidate's convert_ms, and released again). With out == NULL only *count
 * is set. */
#define LESPMV_TUNE_PANELS (1 << LESPMV_NUM_FORMATS)
int lespmv_tune_candidates(int32_t fmt_mask, int32_t dtype, lespmv_tune_result_t* out, int32_t cap, int32_t* count);
int lespmv_tune(lespmv_matrix_t A, lespmv_tune_result_t* candidates, int32_t n, int32_t iters, int32_t k);
/* Best candidate of one format's grid; the winner stays converted on A. */
int lespmv_autotune(lespmv_matrix_t A, int32_t fmt, int32_t iters, int32_t k, lespmv_tune_result_t* best);

/* ---- MLP trained on the GPU (the learned selector's network) -------------------
 * No reference counterpart (the reference's selector is a Keras model,
 * SmartAdpter/models/SpMV_Adapter.py:29-154). in -> hidden1 -> hidden2 ->
 * sum(head_sizes), ReLU, one softmax cross-entropy per head (label -1 = none),
 * Adam. FP32. Inputs are standardised with the stored mean / std. Host
 * buffers throughout; X is n x in row-major, labels n x nheads. */
typedef struct lespmv_mlp_s* lespmv_mlp_t;
int lespmv_mlp_create(lespmv_mlp_t* out, int32_t in_dim, int32_t hidden1, int32_t hidden2, int32_t nheads,
                      const int32_t* head_sizes, uint64_t seed);
int lespmv_mlp_destroy(lespmv_mlp_t mlp);
int lespmv_mlp_shape(lespmv_mlp_t mlp, int32_t* in_dim, int32_t* out_dim, int64_t* nparams);
/* Flat parameters: W1 (hidden1 x in), b1, W2 (hidden2 x hidden1), b2, W3 (out x hidden2), b3. */
int lespmv_mlp_get_params(lespmv_mlp_t mlp, float* host);
int lespmv_mlp_set_params(lespmv_mlp_t mlp, const float* host);
int lespmv_mlp_set_normalizer(lespmv_mlp_t mlp, const float* mean, const float* stddev);
/* Loss (sum over heads of the mean cross-entropy of the labelled samples) and
 * its gradient with respect to the flat parameters, over the whole batch. */
int lespmv_mlp_loss_grad(lespmv_mlp_t mlp, const float* X, const int32_t* labels, int32_t n, double* loss,
                         float* grad);
/* Mini-batch Adam (beta 0.9 / 0.999, eps 1e-8, decoupled weight decay); epoch_loss[epochs] may be NULL. */
int lespmv_mlp_train(lespmv_mlp_t mlp, const float* X, const int32_t* labels, int32_t n, int32_t epochs,
                     int32_t batch, float lr, float weight_decay, uint64_t seed, float* epoch_loss);
/* Per-head softmax probabilities, n x out. */
int lespmv_mlp_forward(lespmv_mlp_t mlp, const float* X, int32_t n, float* probs);
int lespmv_mlp_save(lespmv_mlp_t mlp, const char* path);
int lespmv_mlp_load(lespmv_mlp_t* out, const char* path);


/* ---- batched serving --------------------------------------------------------------
 * No reference counterpart. A server keeps its own CSR copies of resident
 * matrices within an HBM budget (admission by footprint: adding a matrix evicts
 * idle matrices, least recently used first; LESPMV_ERR_OOM when the matrices in
 * use leave no room). Requests (y = alpha A x + beta y for k = 1..32 right-hand
 * sides, row-major device X / Y) are queued; every call to submit / poll
 * retires finished batches and packs the queued requests, up to max_batch per
 * batch, into one launch of a persistent kernel replayed from a HIP graph on a
 * free slot (up to nslots batches in flight on their own streams).
 * A request's result does not depend on the batch it ran in. Latency is
 * measured on the host from submit to the poll that sees the batch finished.
 * Requests that write the same Y must not be in flight together. */
typedef struct lespmv_server_s* lespmv_server_t;
typedef struct {
  int32_t resident, queued;
  int64_t bytes_used, bytes_budget, batches, items, completed;
} lespmv_server_stats_t;
int lespmv_server_create(lespmv_server_t* out, int32_t dtype, int64_t hbm_budget, int32_t max_batch, int32_t nslots);
int lespmv_server_destroy(lespmv_server_t S);
/* Copies a host CSR in; lanes = lanes per row (rows per wave = 64 / lanes; power of two <= 64) for every
 * request, all through the CSR row engine, or 0 = autotune: SpMV lanes from the CSR vector candidates,
 * SpMM lanes from timing the row engine at k = 2 and k = 16 (requests with k = 2..3 and k >= 4 use
 * those); then, timed per call against those, a sliced layout for k = 1 (S-ELL or SELL-C-4096, C in
 * {32, 64}, for rows of at most 64 nonzeros) and 4x4 BSR on the matrix cores (FP64, at most 1.5 stored
 * values per nonzero) for k from 5, 9 or 17 up, each kept where it is at least 3 % faster. */
int lespmv_server_add_csr(lespmv_server_t S, int64_t m, int64_t n, int64_t nnz, const int32_t* row_offset,
                          const int32_t* col_index, const void* values, int32_t lanes, int32_t* id);
int lespmv_server_remove(lespmv_server_t S, int32_t id);
/* The resident matrix as a handle (owned by the server) and its lanes setting for k = 1. */
int lespmv_server_matrix(lespmv_server_t S, int32_t id, lespmv_matrix_t* A, int32_t* lanes);
/* Lanes per row the server uses for this matrix's requests with k right-hand sides (the same product
 * through lespmv_spmm with CSR variant 16 + log2(lanes) gives bitwise the server's result). */
int lespmv_server_lanes(lespmv_server_t S, int32_t id, int32_t k, int32_t* lanes);
/* Sets the layouts a resident matrix is served from (no request of it may be in flight):
 * fmt = LESPMV_SELL or LESPMV_SELL_C_SIGMA: k = 1 requests through that sliced layout (params: chunk
 *   32 or 64, sigma, alignment; LESPMV_ERR_UNSUPPORTED when a chunk holds more than 64 slots);
 * fmt = LESPMV_BSR: requests with k >= k_from (5..32) through 4x4 BSR on the matrix cores (FP64 servers;
 *   params->max_fill bounds the fill when params is given);
 * fmt = LESPMV_CSR: every request through the row engine again (the other layouts released).
 * A layout's device bytes count against the budget (LESPMV_ERR_OOM when no idle matrix can make room).
 * The layouts live on the matrix handle (lespmv_server_matrix): lespmv_spmv of the sliced format and
 * lespmv_spmm of BSR (k > 4) on it give bitwise the server's results. */
int lespmv_server_use(lespmv_server_t S, int32_t id, int32_t fmt, const lespmv_params_t* params, int32_t k_from);
/* The format a request of this matrix with k right-hand sides runs on: LESPMV_CSR (row engine),
 * LESPMV_SELL / LESPMV_SELL_C_SIGMA or LESPMV_BSR. */
int lespmv_server_format(lespmv_server_t S, int32_t id, int32_t k, int32_t* fmt);
int lespmv_server_submit(lespmv_server_t S, int32_t id, int32_t k, double alpha, const void* X, int64_t ldx,
                         double beta, void* Y, int64_t ldy, int64_t* request_id);
/* Completed requests since the last poll (wait = 1 blocks until at least one
 * completes or nothing is in flight). */
int lespmv_server_poll(lespmv_server_t S, int32_t wait, int64_t* request_ids, double* latency_us, int32_t cap,
                       int32_t* ndone);
int lespmv_server_drain(lespmv_server_t S);
int lespmv_server_stats(lespmv_server_t S, lespmv_server_stats_t* stats);

/* ---- matrix features (Get_features) -------------------------------------------
 * Replaces the MTX feature extractor: MtxLoad's histograms
 * (src/sparse_features.cpp:121-604), P_ratioAndGini (:611-640),
 * CalculateFeatures (:642-776), CalculateTilesFeatures (:782-853) and the
 * distance of CalculateTilesExtraFeatures (:941-968), driven by
 * test/Get_features.cpp:46-54, with IndexType int and ValueType double. The
 * matrix is the handle's CSR read as a MatrixMarket entry list in row order;
 * for a matrix read from a symmetric file pass LESPMV_FEAT_SYMMETRIC_FILE and the
 * expanded CSR gives the reference's counts. Member names follow
 * include/sparse_features.h:150-328. The 256 x 256 tile grid, the -1.0 start of
 * the t_var_* and t_nz_ratio_* sums, and GrX_* = -1 (their code is compiled out
 * in the reference) are kept. Counts are exact; ratios, p-ratio and Gini are
 * computed from exact integer sums with the reference's formulas; variances
 * from exact integer moments (the reference sums squared differences in file
 * order, so these agree to rounding). */
#define LESPMV_FEAT_SYMMETRIC_FILE 1
#define LESPMV_FEAT_TILES 256 /* MAT_TILE_SIZE, include/general_config.h:8 */
#define LESPMV_NUM_FEATURES 40 /* lines of the .features file, src/sparse_features.cpp:1347-1400 */

typedef struct {
  int64_t num_rows, num_cols, num_nnzs, nnz_mtx, nnz_lower, nnz_upper, nnz_diagonal, complete_ndiags, close_nnz;
  int64_t min_nnz_each_row, max_nnz_each_row, min_nnz_each_col, max_nnz_each_col;
  int64_t nz_rows, nz_cols, diag_dominant_rows;
  int64_t t_num_blocks, t_num_RB, t_num_CB, t_mod_RB, t_mod_CB;
  int64_t t_min_nnz_all_tiles, t_max_nnz_all_tiles, t_min_nnz_each_RB, t_max_nnz_each_RB;
  int64_t t_min_nnz_each_CB, t_max_nnz_each_CB;
  int64_t uniq_rows_sum, uniq_cols_sum, symm_pattern_pairs, symm_value_pairs;
  int64_t is_symmetric;
  double pattern_symm, value_symm, nnz_ratio, diag_close_ratio, distance_per_row;
  double nz_row_ratio, ave_nnz_each_row, var_nnz_each_row, standard_dev_row, Gini_row, P_ratio_row;
  double nz_col_ratio, ave_nnz_each_col, var_nnz_each_col, standard_dev_col, Gini_col, P_ratio_col;
  double diagonal_dominant_ratio, max_value_offdiag, max_value_diagonal, row_variability, col_variability;
  double t_ave_nnz_all_tiles, t_ave_nnz_RB, t_ave_nnz_CB, t_var_nnz_all_tiles, t_var_nnz_RB, t_var_nnz_CB;
  double t_standard_dev_all_tiles, t_standard_dev_RB, t_standard_dev_CB;
  double t_Gini_all_tiles, t_Gini_RB, t_Gini_CB, t_P_ratio_all_tiles, t_P_ratio_RB, t_P_ratio_CB;
  double t_nz_ratio_tiles, t_nz_ratio_RB, t_nz_ratio_CB;
  double uniqR, uniqC, potReuseR, potReuseC, GrX_uniqR, GrX_uniqC, GrX_potReuseR, GrX_potReuseC;
  double device_ms; /* GPU time of the extraction */
} lespmv_features_t;

/* The images FeaturesWrite stores (src/sparse_features.cpp:1195-1338), plus the
 * per-tile distinct-row / distinct-column counts (uniq_RB, uniq_CB). Host
 * buffers; any pointer may be NULL. Tile arrays hold 256*256 entries in
 * row-major tile order, block arrays 256. */
typedef struct {
  int32_t* tile_nnz;       /* .nnz  nnz_by_Tiles_ */
  int32_t* tile_max;       /* .max  max_rownnz_per_tile_ */
  double* tile_ave;        /* .ave  ave_rownnz_per_tile_ */
  double* tile_std;        /* .std  std_rownnz_per_tile_ */
  int32_t* tile_uniq_rows; /* uniq_RB */
  int32_t* tile_uniq_cols; /* uniq_CB */
  int32_t* rb_nnz;         /* .RBnnz */
  int32_t* rb_max;         /* .RBmax */
  double* rb_ave;          /* .RBave */
  double* rb_std;          /* .RBstd */
  int32_t* cb_nnz;         /* .CBnnz */
  int32_t* cb_max;         /* .CBmax */
  double* cb_ave;          /* .CBave */
  double* cb_std;          /* .CBstd */
} lespmv_feature_images_t;

/* Computes every feature on the GPU (histogram and reduction kernels on the
 * handle's stream); synchronous. `images` may be NULL. */
int lespmv_features(lespmv_matrix_t A, int32_t flags, lespmv_features_t* out, lespmv_feature_images_t* images);
/* The 40 values of a .features file in file order, and the name of value i. */
int lespmv_features_vector(const lespmv_features_t* f, double out[LESPMV_NUM_FEATURES]);
const char* lespmv_feature_name(int32_t i);
/* Number of non-empty br x bc blocks for each (br, bc) pair in dims[2*n]
 * (the BSR block count without building BSR); selector input, no reference
 * counterpart. Synchronous. */
int lespmv_block_counts(lespmv_matrix_t A, const int32_t* dims, int32_t n, int64_t* nnzb);

/* ---- learned format selector ----------------------------------------------------
 * Replaces the reference's MM-Adapter prediction (SmartAdpter/models/
 * SpMV_Adapter.py:90-122) with an MLP trained on this library's own timings.
 * 128 inputs from the GPU features (the 40 .features values, 24 derived
 * scalars incl. BSR block fills, 32-bin row-block and column-block shares).
 * Seven heads (35 outputs): format class {CSR-vector, CSR-merge, COO, ELL,
 * SELL family, DIA, BSR, CSR5, CSR column panels}, CSR lanes, SELL (C, sigma),
 * DIA variant, BSR block, CSR5 (omega, sigma), panels (8, 4); a class's
 * parameter head indexes its run [class_first[c], class_first[c+1]) of the
 * standard tuning grid (all formats and LESPMV_TUNE_PANELS).
 * allowed[35] marks what the features already rule out (ELL / DIA fill,
 * diagonal count, BSR fill of the counted block sizes). */
#define LESPMV_SELECTOR_INPUTS 128
#define LESPMV_SELECTOR_OUTPUTS 35
int lespmv_selector_layout(int32_t* class_first /*[10]*/, int32_t* head_of_class /*[9]*/, int32_t* head_sizes /*[7]*/);
int lespmv_selector_inputs(lespmv_matrix_t A, int32_t flags, float* inputs /*[128]*/, uint8_t* allowed /*[35]*/,
                           lespmv_features_t* features /* may be NULL */);
int lespmv_selector_decode(const float* probs /*[35]*/, const uint8_t* allowed /*[35]*/, int32_t* cls,
                           int32_t* candidate);
/* Features, forward pass and decode: the chosen grid candidate (not converted). */
int lespmv_select(lespmv_matrix_t A, lespmv_mlp_t mlp, int32_t flags, lespmv_tune_result_t* choice, int32_t* cls);

#ifdef __cplusplus
}
#endif

#endif /* LESPMV_HIP_H */
