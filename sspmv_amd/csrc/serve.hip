// Batched SpMV / SpMM serving: a packed persistent kernel replayed from a HIP
// graph, and a continuous-batching request loop with admission of resident
// matrices by HBM footprint. No reference counterpart (the reference runs one
// matrix per process, test/benchmark_spmv_<fmt>.cpp).
//
// Resident matrices are the server's own CSR copies with a lanes-per-row value
// from the autotuner. A batch is a table of request descriptors (matrix arrays,
// X, Y, k, alpha, beta, rows per work item, first work item); a request's work
// items are consecutive row ranges, sized on the host from the matrix's mean
// row length and k so that an item is a few passes of ~kItemIters loads per
// lane. Each request is placed on one XCD (largest work first, least loaded
// XCD): its items form that XCD's queue, so its X is gathered into one L2.
// The persistent kernel's waves take a few items at a time (the batch's grab
// size, about eight grabs per wave) from their own XCD's counter and then help
// the other queues, find the request by a search over the first item numbers
// held in LDS, and compute the rows for all k right-hand sides: a lane covers
// W right-hand sides (1, 2 or 4, one vector access when aligned), a row gets G
// lanes (G = max(lanes, kk), kk the lanes that cover k), lane t handles
// right-hand sides (t mod kk)*W.. and every (G/kk)-th nonzero, and the partial
// sums meet in xor shuffles. W, G and the order of every sum depend only on
// the request, never on the batch or on which wave runs an item, so a
// request's result is bitwise the same alone and in any mix.
//
// Each in-flight slot owns a stream, pinned host table, device table (per-XCD
// counters, header, descriptors) and captured graphs: copy the table in
// (resized to the batch; the copy also zeroes the counters), run the
// persistent kernel.
#include <algorithm>
#include <chrono>
#include <deque>
#include <vector>

#include "common.hpp"
#include "rows_engine.hpp"

namespace lsp {

namespace {

constexpr int kMaxBatch = 256;
constexpr int kServeBlock = 256;

struct alignas(16) ReqDesc {
  const int* rp;
  const int* ci;
  const void* val;
  const void* X;
  void* Y;
  int64_t ldx, ldy;
  double alpha, beta;
  int32_t m, k, lanes, rows_per_item;
  int32_t item_begin, nitems;
  int32_t mode, pad[3];  // right-hand sides per lane and vector access (rhs_mode), + kBufferMode for buffer loads
};
constexpr int32_t kBufferMode = 8;

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
      const int r1 = r0 + q.rows_per_item < q.m ? r0 + q.rows_per_item : q.m;
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
        default: run_rows<V, 4, true, true>(q, r0, r1, lane); break;
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
  int64_t bytes = 0, last_use = 0, pending = 0;
  bool live = false;
};

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
    while (n < (int)std::min<size_t>(queue.size(), (size_t)max_batch)) {
      const Pending& p = queue[n];
      const int64_t m = mats[p.mat].A->m;
      const int64_t add = ceil_div(m, rows_per_item(mats[p.mat], p.k, kModeW[rhs_mode(p)]));
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
      work[i] = (double)A.m + (double)A.nnz * (double)ceil_div(p.k, kModeW[mode[i]]);
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
        const int rpi = rows_per_item(sm, p.k, kModeW[mode[i]]);
        const int32_t nitems = (int32_t)ceil_div(A.m, rpi);
        const int32_t dmode = mode[i] | (rows_buffer_ok(A.nnz, A.n, p.ldx, vsize(dtype)) ? kBufferMode : 0);
        R[d++] = ReqDesc{A.rp.as<int>(), A.ci.as<int>(), A.val.p, p.X, p.Y, p.ldx, p.ldy, p.alpha, p.beta,
                         (int32_t)A.m, p.k, lanes_of(sm, kModeW[mode[i]]), rpi, item, nitems, dmode, {0, 0, 0}};
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
    LSP_HIP(hipGraphExecMemcpyNodeSetParams1D(sl.exec[g], sl.copy[g], sl.d_table.p, sl.h_table,
                                              kCounterBytes + sizeof(BatchHeader) + sizeof(ReqDesc) * (size_t)n,
                                              hipMemcpyHostToDevice));
    for (const Pending& p : batch) mats[p.mat].last_use = ++tick;
    queue.erase(queue.begin(), queue.begin() + n);
    sl.reqs = std::move(batch);
    LSP_HIP(hipGraphLaunch(sl.exec[g], sl.s));
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
    const int64_t bytes = (m + 1) * 4 + nnz * (4 + vs);
    LSP_REQUIRE(bytes <= S.budget, LESPMV_ERR_OOM, "matrix larger than the server's HBM budget");
    LSP_REQUIRE(S.make_room(bytes), LESPMV_ERR_OOM, "no room: every resident matrix has requests in flight");
    lespmv_matrix_t mh = nullptr;
    int rc = lespmv_create_csr(&mh, S.dtype, m, n, nnz, rp, ci, val, 0);
    LSP_REQUIRE(rc == LESPMV_OK, rc, "creating the resident matrix failed");
    ServerMat sm;
    sm.A.reset(reinterpret_cast<Matrix*>(mh));
    sm.h = mh;
    sm.avg_row = (double)nnz / (double)m;
    const bool tune = lanes == 0;
    if (lanes == 0) {  // the autotuner's lanes per row (rows per wave = 64 / lanes) for this matrix
      int32_t count = 0;
      lespmv_tune_candidates(1 << LESPMV_CSR, S.dtype, nullptr, 0, &count);
      std::vector<lespmv_tune_result_t> cand(count);
      lespmv_tune_candidates(1 << LESPMV_CSR, S.dtype, cand.data(), count, &count);
      rc = lespmv_tune(mh, cand.data(), count, 5, 1);
      lanes = 8;
      double best = 1e300;
      for (const auto& c : cand)
        if (rc == LESPMV_OK && c.status == LESPMV_OK && c.variant >= 17 && c.variant <= 22 && c.ms < best) {
          best = c.ms;
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
