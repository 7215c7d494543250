// Batched SpMV / SpMM serving: a packed persistent kernel replayed from a HIP
// graph, and a continuous-batching request loop with admission of resident
// matrices by HBM footprint. No reference counterpart (the reference runs one
// matrix per process, test/benchmark_spmv_<fmt>.cpp).
//
// Resident matrices are the server's own CSR copies with a lanes-per-row value
// from the autotuner, plus, where the autotuner timed them faster, a sliced
// layout (S-ELL or SELL-C-sigma, C in {32, 64}, every chunk at most
// kUnitSlots wide) for k = 1 and a 4x4 BSR layout (FP64, matrix cores) for k
// from a timed threshold up. A batch is a table of request descriptors (matrix
// arrays, X, Y, k, alpha, beta, units per work item, first work item); a
// request's work items are consecutive ranges of rows (CSR; positions in a
// per-matrix order that sorts rows by length within windows of 64 where that
// balances the row engine's passes), chunks (sliced) or block rows (BSR),
// sized on the host from the matrix's mean row length and k so that an item
// is a few passes of ~kItemIters loads per lane. Each request
// is placed on one XCD (largest work first, least loaded XCD): its items form
// that XCD's queue, so its X is gathered into one L2.
// The persistent kernel's waves take a few items at a time (the batch's grab
// size, about eight grabs per wave) from their own XCD's counter and then help
// the other queues, find the request by a search over the first item numbers
// held in LDS, and compute the item. CSR rows run on the row engine for all k
// right-hand sides: a lane covers
// W right-hand sides (1, 2 or 4, one vector access when aligned), a row gets G
// lanes (G = max(lanes, kk), kk the lanes that cover k), lane t handles
// right-hand sides (t mod kk)*W.. and every (G/kk)-th nonzero, and the partial
// sums meet in xor shuffles. Sliced chunks take one thread per row
// (sliced_engine.hpp, 64 / C chunks per wave), BSR block rows one wave each
// (bsr_engine.hpp). Formats, lane layouts and the order of every sum depend
// only on the request, never on the batch or on which wave runs an item, so a
// request's result is bitwise the same alone and in any mix, and the same as
// the per-call kernel on the same layout.
//
// Each in-flight slot owns a stream, pinned host table, device table (per-XCD
// counters, header, descriptors) and captured graphs: copy the table in
// (resized to the batch; the copy also zeroes the counters), run the
// persistent kernel. Batches of a few requests issue the same copy and kernel
// on the slot's stream without the graph, whose launch costs more.
#include <algorithm>
#include <chrono>
#include <deque>
#include <vector>

#include "bsr_engine.hpp"
#include "common.hpp"
#include "rows_engine.hpp"
#include "sliced_engine.hpp"

namespace lsp {

namespace {

constexpr int kMaxBatch = 256;
constexpr int kServeBlock = 256;

struct alignas(16) ReqDesc {
  const int* rp;        // CSR row offsets | sliced chunk widths | BSR block-row offsets
  const int* ci;        // CSR columns | sliced slot columns | BSR block columns
  const void* val;      // CSR values | sliced slot values | BSR block values (16 per block, row-major)
  const void* X;
  void* Y;
  const int64_t* coff;  // sliced: chunk offsets
  const int* order;     // sliced: original row of each sorted row (null: natural order)
  int64_t ldx, ldy;
  double alpha, beta;
  int32_t m, n, k, lanes;
  int32_t rows_per_item, units;  // units per work item and in all: rows, chunks or block rows
  int32_t item_begin, nitems;
  int32_t mode, C, pad[2];  // mode: rows_mode (+ kBufferMode for buffer loads), kSellMode or kBsrMode; C: chunk
};
constexpr int32_t kBufferMode = 8;
constexpr int32_t kSellMode = 16;  // k = 1 through a sliced layout
constexpr int32_t kBsrMode = 17;   // 4x4 BSR on the matrix cores (FP64)
constexpr int kSellSteps = 6;      // slots per pipelined step (keeps the packed kernel's registers)

// Generic pointers read from a descriptor, told to the compiler as global
// ones, so it issues global (not flat) loads.
template <class T>
__device__ __forceinline__ T* global_ptr(T* p) {
  return (T*)(LSP_GLOBAL_AS T*)p;
}

// Chunks [c0, c1) of a sliced request (k = 1): lane t takes row t mod C of
// chunk c0 + t / C, then the chunk 64 / C further on.
template <class V>
__device__ __forceinline__ void run_sell(const ReqDesc& q, int c0, int c1, int lane) {
  const int C = q.C, per = 64 / C, r = lane % C;
  const int* __restrict__ width = global_ptr(q.rp);
  const int* __restrict__ col = global_ptr(q.ci);
  const V* __restrict__ val = global_ptr((const V*)q.val);
  const int64_t* __restrict__ coff = global_ptr(q.coff);
  const int* __restrict__ order = q.order ? global_ptr(q.order) : nullptr;
  const V* __restrict__ x = global_ptr((const V*)q.X);
  V* __restrict__ y = global_ptr((V*)q.Y);
  const V alpha = (V)q.alpha, beta = (V)q.beta;
  for (int c = c0 + lane / C; c < c1; c += per) {
    const int64_t row = (int64_t)c * C + r;
    if (row >= q.m) continue;
    const int64_t off = coff[c];
    const V sum = sliced_row_sum<V, kSellSteps>(col + off + r, val + off + r, C, 0, width[c], x, q.ldx);
    V* out = y + (order ? (int64_t)order[row] : row) * q.ldy;
    *out = beta == V(0) ? alpha * sum : alpha * sum + beta * *out;
  }
}

// Block rows [b0, b1) of a 4x4 BSR request, one after another on the whole wave.
__device__ __forceinline__ void run_bsr(const ReqDesc& q, int b0, int b1, int lane) {
  const int* rptr = global_ptr(q.rp);
  const int* bcol = global_ptr(q.ci);
  const double* data = global_ptr((const double*)q.val);
  const double* X = global_ptr((const double*)q.X);
  double* Y = global_ptr((double*)q.Y);
  for (int b = b0; b < b1; b++)
    bsr44_mfma_row(rptr, bcol, data, q.m, q.n, q.k, X, q.ldx, Y, q.ldy, q.alpha, q.beta, b, lane);
}

// Each request is placed on one XCD: its work items form that XCD's queue
// [xbegin[x], xbegin[x + 1]) with its own counter, so the request's X is
// gathered into one L2 instead of eight.
struct alignas(16) BatchHeader {
  int32_t nreq, total_items, grab, pad;  // grab: work items a wave takes per atomic
  int32_t xbegin[kXcds + 1];
  int32_t pad2[3];
};

// Device table = the per-XCD counters (one 256-byte line each), the header,
// the descriptors. The host copy keeps its counters at zero, so the table
// copy that starts every launch also resets them (no memset node, no reset
// kernel).
constexpr int kCounterStride = 64;  // ints between the per-XCD counters
constexpr size_t kCounterBytes = sizeof(int) * kCounterStride * kXcds;
constexpr size_t kTableBytes = kCounterBytes + sizeof(BatchHeader) + sizeof(ReqDesc) * kMaxBatch;

// The XCD this wave runs on (hardware register XCC_ID).
__device__ __forceinline__ int xcc_id() {
  int id;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(id));
  return id & (kXcds - 1);
}

template <class V>
__global__ __launch_bounds__(kServeBlock) void packed_kernel(const unsigned char* __restrict__ table,
                                                              int* __restrict__ counter) {
  __shared__ int s_begin[kMaxBatch + 1];
  const BatchHeader* H = reinterpret_cast<const BatchHeader*>(table);
  const ReqDesc* R = reinterpret_cast<const ReqDesc*>(table + sizeof(BatchHeader));
  const int nreq = H->nreq, total = H->total_items, grab = H->grab;
  for (int i = threadIdx.x; i < nreq; i += kServeBlock) s_begin[i] = R[i].item_begin;
  if (threadIdx.x == 0) s_begin[nreq] = total;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int home = xcc_id();
  // the queue of the wave's own XCD first, then help the others; a queue is
  // left once its counter passes its end, so every wave ends after at most
  // one failed atomic per queue
  for (int v = 0; v < kXcds; v++) {
    const int xq = (home + v) & (kXcds - 1);
    const int qbeg = H->xbegin[xq], qend = H->xbegin[xq + 1];
    for (;;) {
      if (qbeg >= qend) break;
      int first = 0;
      if (lane == 0) first = atomicAdd(counter + xq * kCounterStride, grab);
      first = __builtin_amdgcn_readfirstlane(first) + qbeg;  // lane 0's value, wave-uniform from here on
      if (first >= qend) break;
      const int last = first + grab < qend ? first + grab : qend;
    int lo = 0, hi = nreq - 1;  // last request whose first item <= first
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_begin[mid] <= first) lo = mid;
      else hi = mid - 1;
    }
    for (int item = first; item < last; item++) {
      while (lo + 1 < nreq && s_begin[lo + 1] <= item) lo++;  // items of a grab are consecutive
      lo = __builtin_amdgcn_readfirstlane(lo);
      const ReqDesc& q = R[lo];
      const int r0 = (item - s_begin[lo]) * q.rows_per_item;
      const int r1 = r0 + q.rows_per_item < q.units ? r0 + q.rows_per_item : q.units;
      switch (q.mode) {
        case 0: run_rows<V, 1, false, false>(q, r0, r1, lane); break;
        case 1: run_rows<V, 2, false, false>(q, r0, r1, lane); break;
        case 2: run_rows<V, 2, true, false>(q, r0, r1, lane); break;
        case 3: run_rows<V, 4, false, false>(q, r0, r1, lane); break;
        case 4: run_rows<V, 4, true, false>(q, r0, r1, lane); break;
        case kBufferMode + 0: run_rows<V, 1, false, true>(q, r0, r1, lane); break;
        case kBufferMode + 1: run_rows<V, 2, false, true>(q, r0, r1, lane); break;
        case kBufferMode + 2: run_rows<V, 2, true, true>(q, r0, r1, lane); break;
        case kBufferMode + 3: run_rows<V, 4, false, true>(q, r0, r1, lane); break;
        case kBufferMode + 4: run_rows<V, 4, true, true>(q, r0, r1, lane); break;
        case kSellMode: run_sell<V>(q, r0, r1, lane); break;
        default:
          if constexpr (sizeof(V) == 8) run_bsr(q, r0, r1, lane);  // kBsrMode: FP64 servers only
          break;
      }
    }
    }
  }
}

double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct ServerMat {
  std::unique_ptr<Matrix> A;
  lespmv_matrix_t h = nullptr;  // the same object, for the C API
  double avg_row = 0.0;
  int32_t lanes = 1;                   // k = 1
  int32_t lanes_w[3] = {1, 1, 1};      // per right-hand sides per lane W = 1, 2, 4
  int32_t sell_fmt = -1;               // k = 1 through this sliced layout (S-ELL or SELL-C-sigma); -1: row engine
  int32_t bsr_from_k = 0;              // k >= this through the 4x4 BSR layout on the matrix cores; 0: never
  int64_t bytes = 0, last_use = 0, pending = 0;
  bool live = false;
};

// The kernel a request runs on: 0 the CSR row engine, 1 the sliced layout, 2 BSR 4x4.
int request_kind(const ServerMat& sm, int k) {
  if (k == 1 && sm.sell_fmt >= 0) return 1;
  if (sm.bsr_from_k > 0 && k >= sm.bsr_from_k) return 2;
  return 0;
}

struct Pending {
  int64_t id;
  int32_t mat, k;
  double alpha, beta;
  const void* X;
  int64_t ldx;
  void* Y;
  int64_t ldy;
  double t_submit;
};

constexpr int kGraphs = 3;  // persistent grids of three sizes per slot
// Batches of up to this many requests skip the graph: the table copy and the
// kernel are launched on the slot's stream directly, which costs less than a
// graph launch (single-request latency p50 66 -> 59 us); larger batches
// replay the captured graph.
constexpr int kDirectBatch = 4;
// A batch takes at most ceil(queue / slots) requests, but at least this many:
// the last batches of a burst then run side by side instead of one large
// batch alone (serving mix at saturation 83.6k -> 84.6k req/s).
constexpr int kSpreadMin = 32;

struct Slot {
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  hipGraph_t graph[kGraphs] = {};
  hipGraphExec_t exec[kGraphs] = {};
  hipGraphNode_t copy[kGraphs] = {};  // the table copy of each graph, resized per launch
  unsigned char* h_table = nullptr;
  DBuf d_table;
  std::vector<Pending> reqs;
  bool busy = false;
};

struct Server {
  int dtype = LESPMV_F64;
  int64_t budget = 0, used = 0;
  int32_t max_batch = kMaxBatch;
  int grid = 0;              // the largest persistent grid (blocks)
  int grids[kGraphs] = {};   // grid sizes of a slot's graphs, smallest first
  std::vector<ServerMat> mats;
  std::deque<Pending> queue;
  std::vector<Slot> slots;
  int64_t next_id = 0, tick = 0;
  int64_t batches = 0, items = 0, done_count = 0;
  std::vector<std::pair<int64_t, double>> finished;  // (id, latency us) not yet returned

  ~Server() {
    for (auto& sl : slots) {
      if (sl.s) (void)hipStreamSynchronize(sl.s);
      for (int g = 0; g < kGraphs; g++) {
        if (sl.exec[g]) (void)hipGraphExecDestroy(sl.exec[g]);
        if (sl.graph[g]) (void)hipGraphDestroy(sl.graph[g]);
      }
      if (sl.done) (void)hipEventDestroy(sl.done);
      if (sl.h_table) (void)hipHostFree(sl.h_table);
      if (sl.s) (void)hipStreamDestroy(sl.s);
    }
  }

  void build_slot(Slot& sl) {
    LSP_HIP(hipStreamCreate(&sl.s));  // default priority
    LSP_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    LSP_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.h_table), kTableBytes, hipHostMallocDefault));
    std::memset(sl.h_table, 0, kTableBytes);
    sl.d_table.alloc(kTableBytes);
    for (int g = 0; g < kGraphs; g++) {
      LSP_HIP(hipStreamBeginCapture(sl.s, hipStreamCaptureModeThreadLocal));
      LSP_HIP(hipMemcpyAsync(sl.d_table.p, sl.h_table, kTableBytes, hipMemcpyHostToDevice, sl.s));
      if (dtype == LESPMV_F64)
        hipLaunchKernelGGL(packed_kernel<double>, dim3(grids[g]), dim3(kServeBlock), 0, sl.s,
                           sl.d_table.as<unsigned char>() + kCounterBytes, sl.d_table.as<int>());
      else
        hipLaunchKernelGGL(packed_kernel<float>, dim3(grids[g]), dim3(kServeBlock), 0, sl.s,
                           sl.d_table.as<unsigned char>() + kCounterBytes, sl.d_table.as<int>());
      LSP_HIP(hipStreamEndCapture(sl.s, &sl.graph[g]));
      size_t nn = 0;
      LSP_HIP(hipGraphGetNodes(sl.graph[g], nullptr, &nn));
      std::vector<hipGraphNode_t> nodes(nn);
      LSP_HIP(hipGraphGetNodes(sl.graph[g], nodes.data(), &nn));
      for (auto nd : nodes) {
        hipGraphNodeType ty;
        LSP_HIP(hipGraphNodeGetType(nd, &ty));
        if (ty == hipGraphNodeTypeMemcpy) sl.copy[g] = nd;
      }
      LSP_REQUIRE(sl.copy[g], LESPMV_ERR_DEVICE, "captured serving graph has no table copy");
      LSP_HIP(hipGraphInstantiate(&sl.exec[g], sl.graph[g], nullptr, nullptr, 0));
    }
  }

  // evict idle matrices, least recently used first, until `need` bytes fit
  bool make_room(int64_t need) {
    while (used + need > budget) {
      int victim = -1;
      for (int i = 0; i < (int)mats.size(); i++)
        if (mats[i].live && mats[i].pending == 0 && (victim < 0 || mats[i].last_use < mats[victim].last_use))
          victim = i;
      if (victim < 0) return false;
      evict(victim);
    }
    return true;
  }

  void evict(int i) {
    ServerMat& sm = mats[i];
    used -= sm.bytes;
    sm.A.reset();
    sm.h = nullptr;
    sm.live = false;
  }

  void launch(Slot& sl) {
    // the longest queue prefix (up to max_batch) whose work items fit in
    // int32; it leaves the queue only once its table is built (an error while
    // building leaves every request queued)
    int n = 0;
    int64_t total = 0;
    // a short queue is spread over the slots (at least kSpreadMin requests a
    // batch), so that the last batches of a burst run side by side
    const size_t cap = std::min<size_t>(
        {queue.size(), (size_t)max_batch,
         std::max<size_t>((size_t)kSpreadMin, (queue.size() + slots.size() - 1) / slots.size())});
    while (n < (int)cap) {
      const Pending& p = queue[n];
      int64_t units = 0;
      const int per = item_units(p, &units);
      const int64_t add = ceil_div(units, (int64_t)per);
      if (n > 0 && total + add >= INT32_MAX) break;
      total += add;
      n++;
    }
    std::vector<Pending> batch(queue.begin(), queue.begin() + n);
    auto* H = reinterpret_cast<BatchHeader*>(sl.h_table + kCounterBytes);
    auto* R = reinterpret_cast<ReqDesc*>(sl.h_table + kCounterBytes + sizeof(BatchHeader));
    // each request on one XCD: largest estimated work (lane steps) first, on
    // the least loaded XCD; descriptors grouped by XCD
    std::vector<int> mode(n), xcd(n), order(n);
    std::vector<double> work(n);
    for (int i = 0; i < n; i++) {
      const Pending& p = batch[i];
      const Matrix& A = *mats[p.mat].A;
      mode[i] = rhs_mode(p);
      work[i] = (double)A.m + (double)A.nnz * (double)ceil_div(p.k, kModeW[mode[i]]);  // an estimate for every kind
      order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    double load[kXcds] = {};
    for (int i : order) {
      const int x = (int)(std::min_element(load, load + kXcds) - load);
      xcd[i] = x;
      load[x] += work[i];
    }
    int32_t item = 0;
    int d = 0;
    for (int x = 0; x < kXcds; x++) {
      H->xbegin[x] = item;
      for (int i = 0; i < n; i++) {
        if (xcd[i] != x) continue;
        const Pending& p = batch[i];
        ServerMat& sm = mats[p.mat];
        Matrix& A = *sm.A;
        int64_t units = 0;
        const int per = item_units(p, &units);
        const int32_t nitems = (int32_t)ceil_div(units, per);
        ReqDesc q{};
        q.X = p.X;
        q.Y = p.Y;
        q.ldx = p.ldx;
        q.ldy = p.ldy;
        q.alpha = p.alpha;
        q.beta = p.beta;
        q.m = (int32_t)A.m;
        q.n = (int32_t)A.n;
        q.k = p.k;
        q.lanes = lanes_of(sm, kModeW[mode[i]]);
        q.rows_per_item = per;
        q.units = (int32_t)units;
        q.item_begin = item;
        q.nitems = nitems;
        const int kind = request_kind(sm, p.k);
        if (kind == 1) {
          const auto& L = *static_cast<const SlicedData*>(A.fmt[sm.sell_fmt].get());
          q.rp = L.chunk_len.as<int>();
          q.ci = L.col.as<int>();
          q.val = L.val.p;
          q.coff = L.chunk_off.as<int64_t>();
          q.order = L.reorder.p ? L.reorder.as<int>() : nullptr;
          q.mode = kSellMode;
          q.C = L.C;
        } else if (kind == 2) {
          const auto& B = *static_cast<const BsrData*>(A.fmt[LESPMV_BSR].get());
          q.rp = B.row_ptr.as<int>();
          q.ci = B.bcol.as<int>();
          q.val = B.data.p;
          q.mode = kBsrMode;
        } else {
          q.rp = A.rp.as<int>();
          q.ci = A.ci.as<int>();
          q.val = A.val.p;
          q.order = A.row_order.p ? A.row_order.as<int>() : nullptr;
          q.mode = mode[i] | (rows_buffer_ok(A.nnz, A.n, p.ldx, vsize(dtype)) ? kBufferMode : 0);
        }
        R[d++] = q;
        item += nitems;
      }
    }
    H->xbegin[kXcds] = item;
    H->nreq = n;
    H->total_items = item;
    // the smallest grid with at least one wave per item (a short batch finishes
    // in about one item's time); about eight grabs per wave: small items for
    // balance, few atomics
    int g = 0;
    while (g + 1 < kGraphs && (int64_t)grids[g] * (kServeBlock / 64) < item) g++;
    H->grab = (int32_t)std::clamp<int64_t>(item / ((int64_t)grids[g] * (kServeBlock / 64) * 8), 1, 64);
    // copy only the descriptors in use (the slot's previous launch has finished)
    const size_t tbytes = kCounterBytes + sizeof(BatchHeader) + sizeof(ReqDesc) * (size_t)n;
    const bool direct = n <= kDirectBatch;
    if (!direct)
      LSP_HIP(hipGraphExecMemcpyNodeSetParams1D(sl.exec[g], sl.copy[g], sl.d_table.p, sl.h_table, tbytes,
                                                hipMemcpyHostToDevice));
    for (const Pending& p : batch) mats[p.mat].last_use = ++tick;
    queue.erase(queue.begin(), queue.begin() + n);
    sl.reqs = std::move(batch);
    if (direct) {
      LSP_HIP(hipMemcpyAsync(sl.d_table.p, sl.h_table, tbytes, hipMemcpyHostToDevice, sl.s));
      if (dtype == LESPMV_F64)
        hipLaunchKernelGGL(packed_kernel<double>, dim3(grids[g]), dim3(kServeBlock), 0, sl.s,
                           sl.d_table.as<unsigned char>() + kCounterBytes, sl.d_table.as<int>());
      else
        hipLaunchKernelGGL(packed_kernel<float>, dim3(grids[g]), dim3(kServeBlock), 0, sl.s,
                           sl.d_table.as<unsigned char>() + kCounterBytes, sl.d_table.as<int>());
      LSP_LAUNCHED();
    } else {
      LSP_HIP(hipGraphLaunch(sl.exec[g], sl.s));
    }
    LSP_HIP(hipEventRecord(sl.done, sl.s));
    sl.busy = true;
    batches++;
    items += item;
  }

  int rhs_mode(const Pending& p) const { return rows_mode(p.k, p.ldx, p.ldy, p.X, p.Y, vsize(dtype)); }

  static int lanes_of(const ServerMat& sm, int W) { return sm.lanes_w[W == 1 ? 0 : W == 2 ? 1 : 2]; }

  static int rows_per_item(const ServerMat& sm, int k, int W) {
    return rows_item(sm.avg_row, lanes_of(sm, W), k, W);
  }

  // Units (rows, chunks or block rows) of one work item of request p; *units
  // = the request's units in all. A sliced item is passes of 64 / C chunks of
  // ~kItemIters slots per lane; a BSR item block rows of ~kItemIters loads per
  // lane (one block value and one X value per group of 16 right-hand sides).
  int item_units(const Pending& p, int64_t* units) const {
    const ServerMat& sm = mats[p.mat];
    const Matrix& A = *sm.A;
    const int kind = request_kind(sm, p.k);
    if (kind == 1) {
      const auto& L = *static_cast<const SlicedData*>(A.fmt[sm.sell_fmt].get());
      *units = L.nchunks;
      const double width = L.nchunks ? (double)L.slots / ((double)L.nchunks * L.C) : 1.0;
      const int passes = (int)std::clamp(std::ceil(kItemIters / std::max(width, 1.0)), 1.0, (double)kMaxPasses);
      return passes * (64 / L.C);
    }
    if (kind == 2) {
      const auto& B = *static_cast<const BsrData*>(A.fmt[LESPMV_BSR].get());
      *units = B.mb;
      const double per_brow = B.mb ? (double)B.nnzb / (double)B.mb * (1.0 + (double)((p.k + 15) / 16)) : 1.0;
      return (int)std::clamp(std::ceil(kItemIters / std::max(per_brow, 1.0)), 1.0, (double)kMaxPasses);
    }
    *units = A.m;
    return rows_per_item(sm, p.k, kModeW[rhs_mode(p)]);
  }

  // retire finished slots; launch queued requests on free slots
  void progress(bool wait) {
    for (;;) {
      bool retired = false;
      for (auto& sl : slots) {
        if (!sl.busy) continue;
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        LSP_HIP(q);
        const double t = now_us();
        for (const auto& p : sl.reqs) {
          finished.emplace_back(p.id, t - p.t_submit);
          mats[p.mat].pending--;
        }
        done_count += (int64_t)sl.reqs.size();
        sl.reqs.clear();
        sl.busy = false;
        retired = true;
      }
      for (auto& sl : slots)
        if (!sl.busy && !queue.empty()) launch(sl);
      const bool any_busy = std::any_of(slots.begin(), slots.end(), [](const Slot& s) { return s.busy; });
      if (!wait || retired || !finished.empty() || !any_busy) return;
    }
  }
};

}  // namespace lsp

using namespace lsp;

namespace {
Server& srv(lespmv_server_t h) {
  LSP_REQUIRE(h, LESPMV_ERR_ARG, "null server");
  return *reinterpret_cast<Server*>(h);
}

// The row engine runs the rows of a pass (64 / G rows side by side) for as
// many steps as the longest of them needs. Rows sorted by length (longest
// first, then by row) within windows of kOrderWindow rows make the rows of a
// pass alike; the order is kept when it cuts the steps of passes of
// kOrderPass rows by at least 15 % (lognormal rows). A row's sum does not
// depend on the pass that computes it, so results are unchanged. Serving mix
// at saturation: 77k req/s in row order, 84k with windows of 64 rows, 83k
// with 32 or 128, 83k with 256, 82k with 1024, 79k sorted whole (p99 at 20k
// req/s 0.49, 0.49, 0.50-0.56, 0.59 and 1.0 ms: long windows put the long
// rows of a matrix into a few work items).
constexpr int kOrderWindow = 64;
constexpr int kOrderPass = 8;

std::vector<int> row_order(int64_t m, const int32_t* rp) {
  auto len = [&](int64_t r) { return rp[r + 1] - rp[r]; };
  auto steps = [&](auto row_at) {
    int64_t total = 0;
    for (int64_t b = 0; b < m; b += kOrderPass) {
      int mx = 0;
      for (int64_t i = b; i < std::min<int64_t>(b + kOrderPass, m); i++) mx = std::max(mx, len(row_at(i)));
      total += mx;
    }
    return total;
  };
  std::vector<int> ord(m);
  for (int64_t i = 0; i < m; i++) ord[i] = (int)i;
  for (int64_t w = 0; w < m; w += kOrderWindow)
    std::stable_sort(ord.begin() + w, ord.begin() + std::min<int64_t>(w + kOrderWindow, m),
                     [&](int a, int b) { return len(a) > len(b); });
  const int64_t natural = steps([](int64_t i) { return i; });
  const int64_t sorted = steps([&](int64_t i) { return (int64_t)ord[i]; });
  if ((double)sorted > 0.85 * (double)natural) ord.clear();
  return ord;
}

// A layout of a resident matrix: released (its bytes leave the budget), or
// counted once it is built, after making room for it (the matrix itself is
// held meanwhile, so it is not the one evicted).
void release_layout(Server& S, ServerMat& sm, int fmt) {
  auto& f = sm.A->fmt[fmt];
  if (!f) return;
  const int64_t b = f->device_bytes();
  LSP_REQUIRE(lespmv_release_format(sm.h, fmt) == LESPMV_OK, LESPMV_ERR_DEVICE, "releasing a layout failed");
  sm.bytes -= b;
  S.used -= b;
}

void count_layout(Server& S, ServerMat& sm, int fmt) {
  const int64_t b = sm.A->fmt[fmt]->device_bytes();
  sm.pending++;
  const bool ok = S.make_room(b);
  sm.pending--;
  if (!ok) {
    (void)lespmv_release_format(sm.h, fmt);
    LSP_REQUIRE(false, LESPMV_ERR_OOM, "no room for the layout: every other resident matrix has requests in flight");
  }
  sm.bytes += b;
  S.used += b;
}

// k = 1 requests through a sliced layout (S-ELL or SELL-C-sigma, C 32 or 64,
// every chunk at most kUnitSlots wide), requests with k >= k_from through 4x4
// BSR (FP64), or (LESPMV_CSR) everything through the row engine.
void use_format(Server& S, ServerMat& sm, int fmt, const lespmv_params_t* prm, int k_from) {
  if (fmt == LESPMV_CSR) {
    release_layout(S, sm, LESPMV_SELL);
    release_layout(S, sm, LESPMV_SELL_C_SIGMA);
    release_layout(S, sm, LESPMV_BSR);
    sm.sell_fmt = -1;
    sm.bsr_from_k = 0;
    return;
  }
  if (fmt == LESPMV_SELL || fmt == LESPMV_SELL_C_SIGMA) {
    LSP_REQUIRE(prm && (prm->chunk == 32 || prm->chunk == 64), LESPMV_ERR_ARG,
                "the serving kernel takes sliced chunks of 32 or 64 rows");
    release_layout(S, sm, LESPMV_SELL);
    release_layout(S, sm, LESPMV_SELL_C_SIGMA);
    sm.sell_fmt = -1;
    lespmv_params_t p = *prm;
    if (p.alignment <= 0) p.alignment = 1;
    const int rc = lespmv_convert(sm.h, fmt, &p);
    LSP_REQUIRE(rc == LESPMV_OK, rc, "building the sliced layout failed");
    const auto& L = *static_cast<const SlicedData*>(sm.A->fmt[fmt].get());
    std::vector<int> width(std::max<int64_t>(L.nchunks, 1), 0);
    if (L.nchunks) LSP_HIP(hipMemcpy(width.data(), L.chunk_len.p, sizeof(int) * L.nchunks, hipMemcpyDeviceToHost));
    if (*std::max_element(width.begin(), width.end()) > kUnitSlots) {
      (void)lespmv_release_format(sm.h, fmt);
      LSP_REQUIRE(false, LESPMV_ERR_UNSUPPORTED, "a chunk is wider than the serving kernel's 64 slots");
    }
    count_layout(S, sm, fmt);
    sm.sell_fmt = fmt;
    return;
  }
  LSP_REQUIRE(fmt == LESPMV_BSR, LESPMV_ERR_ARG, "the server runs CSR, S-ELL, SELL-C-sigma or BSR");
  LSP_REQUIRE(S.dtype == LESPMV_F64, LESPMV_ERR_UNSUPPORTED, "BSR requests run on the FP64 matrix cores");
  LSP_REQUIRE(k_from >= 5 && k_from <= 32, LESPMV_ERR_ARG, "BSR serves k from 5..32 up");
  release_layout(S, sm, LESPMV_BSR);
  sm.bsr_from_k = 0;
  lespmv_params_t p{};
  p.block_rows = 4;
  p.block_cols = 4;
  p.alignment = 1;
  p.max_fill = prm ? prm->max_fill : 0;
  const int rc = lespmv_convert(sm.h, LESPMV_BSR, &p);
  LSP_REQUIRE(rc == LESPMV_OK, rc, "building the BSR layout failed");
  count_layout(S, sm, LESPMV_BSR);
  sm.bsr_from_k = k_from;
}

// Sliced layouts for k = 1: S-ELL and SELL-C-4096 with C in {32, 64}, timed
// per call against the fastest CSR vector candidate (csr_ms); the fastest is
// kept when it is at least 3 % faster. Only for matrices whose rows fit the
// serving kernel's 64-slot chunks.
void tune_sliced(Server& S, ServerMat& sm, int max_row, double csr_ms) {
  if (sm.A->nnz == 0 || max_row > kUnitSlots || !(csr_ms < 1e300)) return;
  std::vector<lespmv_tune_result_t> cand;
  for (int C : {32, 64})
    for (int fmt : {LESPMV_SELL, LESPMV_SELL_C_SIGMA}) {
      lespmv_tune_result_t r{};
      r.fmt = fmt;
      r.variant = LESPMV_VARIANT_AUTO;
      r.params.chunk = C;
      r.params.sigma = 4096;
      r.params.alignment = 1;
      r.params.max_fill = 4;
      cand.push_back(r);
    }
  if (lespmv_tune(sm.h, cand.data(), (int32_t)cand.size(), 5, 1) != LESPMV_OK) return;
  const lespmv_tune_result_t* best = nullptr;
  for (const auto& c : cand)
    if (c.status == LESPMV_OK && (!best || c.ms < best->ms)) best = &c;
  if (!best || !(best->ms < 0.97 * csr_ms)) return;
  try {
    use_format(S, sm, best->fmt, &best->params, 0);
  } catch (const Error&) {  // no room: the row engine serves k = 1
  }
}

// 4x4 BSR on the matrix cores (FP64, block fill at most 1.5 stored values
// per nonzero): timed per call at k = 8, 16 and 32 against the row engine at
// its lanes; BSR takes the requests from k = 5, 9 or 17 up when it is at
// least 3 % faster at 8, 16 and 32, at 16 and 32, or at 32.
void tune_bsr(Server& S, ServerMat& sm, int64_t n) {
  Matrix& A = *sm.A;
  if (S.dtype != LESPMV_F64 || A.nnz == 0) return;
  lespmv_params_t p{};
  p.block_rows = 4;
  p.block_cols = 4;
  p.alignment = 1;
  p.max_fill = 2;
  if (lespmv_convert(sm.h, LESPMV_BSR, &p) != LESPMV_OK) return;
  const auto& B = *static_cast<const BsrData*>(A.fmt[LESPMV_BSR].get());
  int from = 0;
  if ((double)B.nnzb * 16.0 <= 1.5 * (double)A.nnz) {
    DBuf X, Y;
    X.alloc(sizeof(double) * n * 32);
    Y.alloc(sizeof(double) * A.m * 32);
    LSP_HIP(hipMemset(X.p, 0, sizeof(double) * n * 32));
    const int ks[3] = {8, 16, 32}, thr[3] = {5, 9, 17};
    int lg = 0;
    while ((1 << lg) < sm.lanes_w[2]) lg++;
    for (int i = 2; i >= 0; i--) {
      double tb = 0, tc = 0;
      const int k = ks[i];
      int rc = lespmv_time_spmm(sm.h, LESPMV_BSR, LESPMV_VARIANT_AUTO, k, 1.0, X.p, 32, 0.0, Y.p, k, 2, &tb);
      if (rc == LESPMV_OK) rc = lespmv_time_spmm(sm.h, LESPMV_BSR, LESPMV_VARIANT_AUTO, k, 1.0, X.p, 32, 0.0, Y.p, k, 5, &tb);
      if (rc == LESPMV_OK) rc = lespmv_time_spmm(sm.h, LESPMV_CSR, 16 + lg, k, 1.0, X.p, 32, 0.0, Y.p, k, 5, &tc);
      if (rc != LESPMV_OK || !(tb < 0.97 * tc)) break;
      from = thr[i];
    }
  }
  if (from == 0) {
    (void)lespmv_release_format(sm.h, LESPMV_BSR);
    return;
  }
  try {
    count_layout(S, sm, LESPMV_BSR);
    sm.bsr_from_k = from;
  } catch (const Error&) {  // no room: the row engine serves every k
  }
}

// SpMM lanes per row: the row engine timed through lespmv_spmm at k = 2
// (W = 2) and k = 16 (W = 4) for 1..16 lanes; the fastest wins.
void tune_spmm_lanes(Server& S, ServerMat& sm, lespmv_matrix_t mh, int64_t n) {
  if (sm.A->nnz == 0) return;
  const int64_t vs = vsize(S.dtype);
  DBuf X, Y;
  X.alloc(vs * n * 16);
  Y.alloc(vs * sm.A->m * 16);
  LSP_HIP(hipMemset(X.p, 0, vs * n * 16));
  const int ks[2] = {2, 16};
  for (int c = 1; c < 3; c++) {
    const int k = ks[c - 1];
    double best = 1e300;
    for (int lg = 0; lg <= 4; lg++) {
      double ms = 0;
      int rc = lespmv_time_spmm(mh, LESPMV_CSR, 16 + lg, k, 1.0, X.p, k, 0.0, Y.p, k, 1, &ms);
      if (rc == LESPMV_OK) rc = lespmv_time_spmm(mh, LESPMV_CSR, 16 + lg, k, 1.0, X.p, k, 0.0, Y.p, k, 5, &ms);
      LSP_REQUIRE(rc == LESPMV_OK, rc, "timing the row engine failed");
      if (ms < best) {
        best = ms;
        sm.lanes_w[c] = 1 << lg;
      }
    }
  }
}
}  // namespace

extern "C" int lespmv_server_create(lespmv_server_t* out, int32_t dtype, int64_t hbm_budget, int32_t max_batch,
                                    int32_t nslots) {
  return guarded_call([&] {
    LSP_REQUIRE(out && (dtype == LESPMV_F32 || dtype == LESPMV_F64) && hbm_budget > 0 && max_batch >= 1 &&
                    max_batch <= kMaxBatch && nslots >= 1 && nslots <= 8,
                LESPMV_ERR_ARG, "bad server arguments");
    auto S = std::make_unique<Server>();
    S->dtype = dtype;
    S->budget = hbm_budget;
    S->max_batch = max_batch;
    int dev = 0;
    LSP_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    LSP_HIP(hipGetDeviceProperties(&p, dev));
    // the largest grid is what stays resident: the packed kernel runs at four
    // waves per SIMD (its registers), 16 waves = four workgroups per CU. A
    // grid twice that served 2 % more requests/s at saturation but its extra
    // workgroups delayed the next batch: p50 at 50k req/s 756 -> 478 us.
    S->grid = p.multiProcessorCount * 4;
    S->grids[0] = std::max(1, p.multiProcessorCount / 4);
    S->grids[1] = p.multiProcessorCount * 2;
    S->grids[2] = S->grid;
    S->slots.resize(nslots);
    for (auto& sl : S->slots) S->build_slot(sl);
    *out = reinterpret_cast<lespmv_server_t>(S.release());
  });
}

extern "C" int lespmv_server_destroy(lespmv_server_t h) {
  return guarded_call([&] { delete reinterpret_cast<Server*>(h); });
}

extern "C" int lespmv_server_add_csr(lespmv_server_t h, int64_t m, int64_t n, int64_t nnz, const int32_t* rp,
                                     const int32_t* ci, const void* val, int32_t lanes, int32_t* id) {
  return guarded_call([&] {
    Server& S = srv(h);
    LSP_REQUIRE(id && rp && (nnz == 0 || (ci && val)) && m > 0 && n > 0, LESPMV_ERR_ARG, "bad matrix arguments");
    LSP_REQUIRE(lanes == 0 || (lanes >= 1 && lanes <= 64 && (lanes & (lanes - 1)) == 0), LESPMV_ERR_ARG,
                "lanes must be 0 (tune) or a power of two <= 64");
    const int64_t vs = vsize(S.dtype);
    const std::vector<int> ord = row_order(m, rp);
    const int64_t bytes = (m + 1) * 4 + nnz * (4 + vs) + (int64_t)ord.size() * 4;
    LSP_REQUIRE(bytes <= S.budget, LESPMV_ERR_OOM, "matrix larger than the server's HBM budget");
    LSP_REQUIRE(S.make_room(bytes), LESPMV_ERR_OOM, "no room: every resident matrix has requests in flight");
    lespmv_matrix_t mh = nullptr;
    int rc = lespmv_create_csr(&mh, S.dtype, m, n, nnz, rp, ci, val, 0);
    LSP_REQUIRE(rc == LESPMV_OK, rc, "creating the resident matrix failed");
    ServerMat sm;
    sm.A.reset(reinterpret_cast<Matrix*>(mh));
    sm.h = mh;
    if (!ord.empty()) {  // on the handle, so that the lanes below are timed in this order too
      sm.A->row_order.alloc(sizeof(int) * ord.size());
      LSP_HIP(hipMemcpy(sm.A->row_order.p, ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice));
    }
    sm.avg_row = (double)nnz / (double)m;
    const bool tune = lanes == 0;
    double csr_ms = 1e300;
    if (lanes == 0) {  // the autotuner's lanes per row (rows per wave = 64 / lanes) for this matrix
      int32_t count = 0;
      lespmv_tune_candidates(1 << LESPMV_CSR, S.dtype, nullptr, 0, &count);
      std::vector<lespmv_tune_result_t> cand(count);
      lespmv_tune_candidates(1 << LESPMV_CSR, S.dtype, cand.data(), count, &count);
      rc = lespmv_tune(mh, cand.data(), count, 5, 1);
      lanes = 8;
      for (const auto& c : cand)
        if (rc == LESPMV_OK && c.status == LESPMV_OK && c.variant >= 17 && c.variant <= 22 && c.ms < csr_ms) {
          csr_ms = c.ms;
          lanes = 1 << (c.variant - 16);
        }
    }
    sm.lanes = lanes;
    for (int c = 0; c < 3; c++) sm.lanes_w[c] = lanes;
    if (tune) tune_spmm_lanes(S, sm, mh, n);
    sm.bytes = bytes;
    sm.last_use = ++S.tick;
    sm.live = true;
    S.used += bytes;
    // ids are never reused, so a stale id of an evicted matrix stays an error
    S.mats.push_back(std::move(sm));
    *id = (int32_t)S.mats.size() - 1;
    if (tune) {  // the other layouts, each kept only where it timed faster
      int max_row = 0;
      for (int64_t i = 0; i < m; i++) max_row = std::max(max_row, rp[i + 1] - rp[i]);
      tune_sliced(S, S.mats.back(), max_row, csr_ms);
      tune_bsr(S, S.mats.back(), n);
    }
  });
}

extern "C" int lespmv_server_remove(lespmv_server_t h, int32_t id) {
  return guarded_call([&] {
    Server& S = srv(h);
    LSP_REQUIRE(id >= 0 && id < (int)S.mats.size() && S.mats[id].live, LESPMV_ERR_ARG, "matrix not resident");
    LSP_REQUIRE(S.mats[id].pending == 0, LESPMV_ERR_ARG, "matrix has requests in flight");
    S.evict(id);
  });
}

extern "C" int lespmv_server_matrix(lespmv_server_t h, int32_t id, lespmv_matrix_t* A, int32_t* lanes) {
  return guarded_call([&] {
    Server& S = srv(h);
    LSP_REQUIRE(id >= 0 && id < (int)S.mats.size() && S.mats[id].live, LESPMV_ERR_ARG, "matrix not resident");
    if (A) *A = S.mats[id].h;
    if (lanes) *lanes = S.mats[id].lanes;
  });
}

extern "C" int lespmv_server_lanes(lespmv_server_t h, int32_t id, int32_t k, int32_t* lanes) {
  return guarded_call([&] {
    Server& S = srv(h);
    LSP_REQUIRE(id >= 0 && id < (int)S.mats.size() && S.mats[id].live, LESPMV_ERR_ARG, "matrix not resident");
    LSP_REQUIRE(lanes && k >= 1 && k <= 32, LESPMV_ERR_ARG, "bad arguments");
    *lanes = Server::lanes_of(S.mats[id], k == 1 ? 1 : k <= 3 ? 2 : 4);
  });
}

extern "C" int lespmv_server_use(lespmv_server_t h, int32_t id, int32_t fmt, const lespmv_params_t* params,
                                 int32_t k_from) {
  return guarded_call([&] {
    Server& S = srv(h);
    LSP_REQUIRE(id >= 0 && id < (int)S.mats.size() && S.mats[id].live, LESPMV_ERR_ARG, "matrix not resident");
    LSP_REQUIRE(S.mats[id].pending == 0, LESPMV_ERR_ARG, "matrix has requests in flight");
    use_format(S, S.mats[id], fmt, params, k_from);
    S.mats[id].last_use = ++S.tick;
  });
}

extern "C" int lespmv_server_format(lespmv_server_t h, int32_t id, int32_t k, int32_t* fmt) {
  return guarded_call([&] {
    Server& S = srv(h);
    LSP_REQUIRE(id >= 0 && id < (int)S.mats.size() && S.mats[id].live, LESPMV_ERR_ARG, "matrix not resident");
    LSP_REQUIRE(fmt && k >= 1 && k <= 32, LESPMV_ERR_ARG, "bad arguments");
    const ServerMat& sm = S.mats[id];
    const int kind = request_kind(sm, k);
    *fmt = kind == 1 ? sm.sell_fmt : kind == 2 ? LESPMV_BSR : LESPMV_CSR;
  });
}

extern "C" int lespmv_server_submit(lespmv_server_t h, int32_t id, int32_t k, double alpha, const void* X,
                                    int64_t ldx, double beta, void* Y, int64_t ldy, int64_t* req) {
  return guarded_call([&] {
    Server& S = srv(h);
    LSP_REQUIRE(id >= 0 && id < (int)S.mats.size() && S.mats[id].live, LESPMV_ERR_ARG, "matrix not resident");
    LSP_REQUIRE(k >= 1 && k <= 32 && ldx >= k && ldy >= k && X && Y, LESPMV_ERR_ARG, "bad request");
    Pending p{S.next_id++, id, k, alpha, beta, X, ldx, Y, ldy, now_us()};
    S.mats[id].pending++;
    S.queue.push_back(p);
    if (req) *req = p.id;
    S.progress(false);
  });
}

extern "C" int lespmv_server_poll(lespmv_server_t h, int32_t wait, int64_t* ids, double* latency_us, int32_t cap,
                                  int32_t* ndone) {
  return guarded_call([&] {
    Server& S = srv(h);
    LSP_REQUIRE(ndone && (cap == 0 || ids), LESPMV_ERR_ARG, "bad poll arguments");
    S.progress(wait != 0);
    const int n = (int)std::min<size_t>(S.finished.size(), (size_t)std::max(cap, 0));
    for (int i = 0; i < n; i++) {
      ids[i] = S.finished[i].first;
      if (latency_us) latency_us[i] = S.finished[i].second;
    }
    S.finished.erase(S.finished.begin(), S.finished.begin() + n);
    *ndone = n;
  });
}

extern "C" int lespmv_server_drain(lespmv_server_t h) {
  return guarded_call([&] {
    Server& S = srv(h);
    while (!S.queue.empty() ||
           std::any_of(S.slots.begin(), S.slots.end(), [](const Slot& s) { return s.busy; }))
      S.progress(true);
  });
}

extern "C" int lespmv_server_stats(lespmv_server_t h, lespmv_server_stats_t* st) {
  return guarded_call([&] {
    Server& S = srv(h);
    LSP_REQUIRE(st, LESPMV_ERR_ARG, "null stats");
    st->resident = (int32_t)std::count_if(S.mats.begin(), S.mats.end(), [](const ServerMat& m) { return m.live; });
    st->queued = (int32_t)S.queue.size();
    st->bytes_used = S.used;
    st->bytes_budget = S.budget;
    st->batches = S.batches;
    st->items = S.items;
    st->completed = S.done_count;
  });
}
