// serve_bench: batched serving against plain per-call CSR on one GPU.
//
// Builds the batched workload (matrices with rows log-uniform in [1k, 200k]
// drawn from 2-D / 3-D stencils, lognormal-row random matrices and 4x4-block
// random matrices with 8 blocks per block row; FP64), makes them resident in a
// lespmv server (lanes per row autotuned per matrix), and draws a request
// stream: matrix uniform, k uniform over --ks (default 1..32). It then measures
//   * the server at saturation (every request queued at once) and under
//     Poisson arrivals at the given rates: requests/s and p50 / p99 latency
//     (submit to completion, host clock);
//   * per-call CSR on the same requests: lespmv_spmm with the plain CSR SpMM
//     kernel and with the server's own row engine (same lanes per matrix),
//     (a) one synchronous call per request (the reference's LeSpMV_csr
//     usage), (b) all calls issued asynchronously on the matrices' own
//     streams and one synchronisation at the end (its throughput bound).
// Prints one JSON object on stdout, progress on stderr.
// Options: --matrices=N --requests=N --rates=r1,r2,... --seed=S --budget-gb=G --slots=N
//          --families=0123 (0 2-D Laplacian, 1 27-point stencil, 2 lognormal rows, 3 4x4 blocks)
//          --ks=k1,k2,... (right-hand-side counts, each in 1..32)
//          --lanes=L (lanes per row for every matrix instead of the server's tuning)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../../include/lespmv_hip.h"

static std::string arg(int argc, char** argv, const char* key, const char* def) {
  const size_t n = std::strlen(key);
  for (int i = 1; i < argc; i++)
    if (std::strncmp(argv[i], key, n) == 0 && argv[i][n] == '=') return argv[i] + n + 1;
  return def;
}

#define CHECK(call)                                                                   \
  do {                                                                                \
    int rc_ = (call);                                                                 \
    if (rc_ != LESPMV_OK) {                                                           \
      std::fprintf(stderr, "error %d at %s: %s\n", rc_, #call, lespmv_last_error()); \
      std::exit(3);                                                                   \
    }                                                                                 \
  } while (0)

static double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static uint64_t mix(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

static double pct(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  return v[std::min(v.size() - 1, (size_t)std::floor(q * (double)(v.size() - 1) + 0.5))];
}
static double mean(const std::vector<double>& v) {
  return v.empty() ? 0.0 : std::accumulate(v.begin(), v.end(), 0.0) / (double)v.size();
}

struct Mat {
  int32_t id = -1, lanes = 1;
  lespmv_matrix_t h = nullptr;
  int64_t m = 0, n = 0, nnz = 0;
  void *X = nullptr, *Y = nullptr, *Y2 = nullptr;
  int fam = 0;
};

struct Req {
  int mat, k;
  double arrival_us;
};

int main(int argc, char** argv) {
  const int nmat = std::stoi(arg(argc, argv, "--matrices", "256"));
  const int nreq = std::stoi(arg(argc, argv, "--requests", "4000"));
  const uint64_t seed = std::stoull(arg(argc, argv, "--seed", "1000"));
  const double budget_gb = std::stod(arg(argc, argv, "--budget-gb", "200"));
  const int nslots = std::stoi(arg(argc, argv, "--slots", "4"));
  const int fixed_lanes = std::stoi(arg(argc, argv, "--lanes", "0"));  // 0: the server tunes lanes per matrix
  // auto: the server's tuned layouts (sliced for k = 1, BSR 4x4 from a k up) where they timed faster;
  // csr: every request on the CSR row engine (lanes still tuned); bsr: BSR 4x4 for k >= 9 wherever its
  // fill stays within 2 stored values per nonzero, else as auto
  const std::string formats = arg(argc, argv, "--formats", "auto");
  if (formats != "auto" && formats != "csr" && formats != "bsr") {
    std::fprintf(stderr, "--formats is auto, csr or bsr\n");
    return 2;
  }
  std::vector<int> ks;  // right-hand-side counts to draw from (default 1..32)
  {
    const std::string r = arg(argc, argv, "--ks", "");
    size_t p = 0;
    while (p < r.size()) {
      size_t q = r.find(',', p);
      if (q == std::string::npos) q = r.size();
      ks.push_back(std::stoi(r.substr(p, q - p)));
      p = q + 1;
    }
    if (ks.empty())
      for (int k = 1; k <= 32; k++) ks.push_back(k);
    for (int k : ks)
      if (k < 1 || k > 32) {
        std::fprintf(stderr, "--ks values must lie in 1..32\n");
        return 2;
      }
  }
  std::vector<int> families;  // generator families to draw from (default all four)
  {
    const std::string f = arg(argc, argv, "--families", "0,1,2,3");
    for (char c : f)
      if (c >= '0' && c <= '3') families.push_back(c - '0');
    if (families.empty()) families = {0, 1, 2, 3};
  }
  std::vector<double> rates;
  {
    const std::string r = arg(argc, argv, "--rates", "5000,20000,50000");
    size_t p = 0;
    while (p < r.size()) {
      size_t q = r.find(',', p);
      if (q == std::string::npos) q = r.size();
      rates.push_back(std::stod(r.substr(p, q - p)));
      p = q + 1;
    }
  }
  CHECK(lespmv_init(0));
  lespmv_server_t S;
  CHECK(lespmv_server_create(&S, LESPMV_F64, (int64_t)(budget_gb * 1e9), 256, nslots));

  // ---- resident matrices and per-call handles ----
  std::vector<Mat> mats(nmat);
  int64_t total_nnz = 0, total_rows = 0;
  int sliced_mats = 0, bsr_mats = 0;  // matrices served from a sliced layout at k = 1, from BSR at k = 32
  const double t_build = now_us();
  for (int i = 0; i < nmat; i++) {
    const uint64_t h = mix(seed * 7919 + (uint64_t)i);
    const double rows = std::exp(std::log(1e3) + unit(h) * (std::log(2e5) - std::log(1e3)));
    const int fam = families[mix(h) % families.size()];
    std::vector<double> p;
    int gen;
    if (fam == 0) {
      const double nx = std::max(8.0, std::floor(std::sqrt(rows)));
      gen = LESPMV_GEN_LAPLACE2D;
      p = {nx, nx, 4.0, -1.0};
    } else if (fam == 1) {
      const double k = std::max(4.0, std::floor(std::cbrt(rows)));
      gen = LESPMV_GEN_STENCIL27;
      p = {k, 26.0, -1.0};
    } else if (fam == 2) {
      gen = LESPMV_GEN_LOGNORMAL;
      p = {std::floor(rows), std::floor(rows), 16.0, 1.0, 4096, 0.3, 64};
    } else {
      const double mb = std::max(1.0, std::floor(rows / 4));
      gen = LESPMV_GEN_BLOCKRAND;
      p = {mb, mb, 4, 4, 8};
    }
    lespmv_host_csr_t c;
    CHECK(lespmv_generate(gen, p.data(), (int32_t)p.size(), 1000 + (uint64_t)i, LESPMV_F64, &c));
    Mat& M = mats[i];
    M.m = c.m;
    M.n = c.n;
    M.nnz = c.nnz;
    M.fam = fam;
    CHECK(lespmv_server_add_csr(S, c.m, c.n, c.nnz, c.row_offset, c.col_index, c.values, fixed_lanes, &M.id));
    if (formats == "csr") CHECK(lespmv_server_use(S, M.id, LESPMV_CSR, nullptr, 0));
    if (formats == "bsr") {
      lespmv_params_t bp{};
      bp.max_fill = 2;
      (void)lespmv_server_use(S, M.id, LESPMV_BSR, &bp, 9);  // refused beyond the fill limit
    }
    CHECK(lespmv_server_matrix(S, M.id, nullptr, &M.lanes));
    int32_t f1 = 0, f32 = 0;
    CHECK(lespmv_server_format(S, M.id, 1, &f1));
    CHECK(lespmv_server_format(S, M.id, 32, &f32));
    sliced_mats += f1 != LESPMV_CSR;
    bsr_mats += f32 == LESPMV_BSR;
    CHECK(lespmv_create_csr(&M.h, LESPMV_F64, c.m, c.n, c.nnz, c.row_offset, c.col_index, c.values, 0));
    lespmv_host_csr_free(&c);
    CHECK(lespmv_malloc(&M.X, 8 * M.n * 32));
    CHECK(lespmv_malloc(&M.Y, 8 * M.m * 32));
    CHECK(lespmv_malloc(&M.Y2, 8 * M.m * 32));
    std::vector<double> x((size_t)M.n * 32);
    CHECK(lespmv_fill_uniform(x.data(), LESPMV_F64, (int64_t)x.size(), 77 + (uint64_t)i, -1.0, 1.0));
    CHECK(lespmv_memcpy_h2d(M.X, x.data(), 8 * (int64_t)x.size()));
    total_nnz += M.nnz;
    total_rows += M.m;
    if ((i + 1) % 32 == 0) std::fprintf(stderr, "built %d/%d matrices\n", i + 1, nmat);
  }
  CHECK(lespmv_synchronize());
  lespmv_server_stats_t st;
  CHECK(lespmv_server_stats(S, &st));
  const double build_s = (now_us() - t_build) * 1e-6;

  // ---- request stream ----
  std::vector<Req> reqs(nreq);
  for (int r = 0; r < nreq; r++) {
    const uint64_t h = mix(seed * 104729 + (uint64_t)r);
    reqs[r].mat = (int)(h % (uint64_t)nmat);
    reqs[r].k = ks[mix(h) % ks.size()];
  }
  // flops, and the bytes a request must move at least: its CSR matrix
  // (row pointers, indices, values), X once and Y once
  int64_t flops = 0;
  double bytes = 0;
  for (const auto& q : reqs) {
    const Mat& M = mats[q.mat];
    flops += 2 * M.nnz * q.k;
    bytes += 4.0 * (double)(M.m + 1) + 12.0 * (double)M.nnz + 8.0 * (double)q.k * (double)(M.n + M.m);
  }

  auto run_server = [&](double rate) {
    // Poisson arrivals at `rate` req/s, or everything at t = 0 (saturation)
    std::mt19937_64 rng(seed + (uint64_t)rate);
    std::exponential_distribution<double> gap(rate > 0 ? rate * 1e-6 : 1.0);
    double t = 0;
    for (auto& q : reqs) {
      q.arrival_us = rate > 0 ? t : 0.0;
      if (rate > 0) t += gap(rng);
    }
    std::vector<double> lat;
    lat.reserve(nreq);
    std::vector<int64_t> ids(4096);
    std::vector<double> l(4096);
    const double t0 = now_us();
    int next = 0, done = 0;
    while (done < nreq) {
      const double now = now_us() - t0;
      while (next < nreq && reqs[next].arrival_us <= now) {
        const Req& q = reqs[next];
        Mat& M = mats[q.mat];
        void* Y = (next & 1) ? M.Y2 : M.Y;  // results are checked by the tests, not here
        CHECK(lespmv_server_submit(S, M.id, q.k, 1.0, M.X, q.k, 0.0, Y, q.k, nullptr));
        next++;
      }
      int32_t n = 0;
      CHECK(lespmv_server_poll(S, next == nreq ? 1 : 0, ids.data(), l.data(), (int32_t)ids.size(), &n));
      for (int i = 0; i < n; i++) lat.push_back(l[i]);
      done += n;
    }
    const double secs = (now_us() - t0) * 1e-6;
    char buf[512];
    std::snprintf(buf, sizeof buf,
                  "{\"offered_req_per_s\": %s, \"req_per_s\": %.1f, \"p50_us\": %.1f, \"p99_us\": %.1f, "
                  "\"mean_us\": %.1f, \"gflops\": %.1f, \"GBs\": %.1f}",
                  rate > 0 ? std::to_string((long long)rate).c_str() : "\"saturation\"", nreq / secs,
                  pct(lat, 0.5), pct(lat, 0.99), mean(lat), flops / secs * 1e-9, bytes / secs * 1e-9);
    std::fprintf(stderr, "server %s\n", buf);
    return std::string(buf);
  };

  // per-call kernels: the plain CSR SpMM (variant AUTO), or the server's own
  // row engine with the matrix's tuned lanes (variant 16 + log2 lanes)
  auto variant_of = [&](const Mat& M, bool engine) {
    if (!engine) return (int)LESPMV_VARIANT_AUTO;
    int v = 16;
    while ((1 << (v - 16)) < M.lanes) v++;
    return v;
  };
  auto run_percall_sync = [&](bool engine) {
    std::vector<double> lat;
    const double t0 = now_us();
    for (const auto& q : reqs) {
      Mat& M = mats[q.mat];
      const double a = now_us();
      CHECK(lespmv_spmm(M.h, LESPMV_CSR, variant_of(M, engine), q.k, 1.0, M.X, q.k, 0.0, M.Y, q.k, nullptr));
      CHECK(lespmv_synchronize());
      lat.push_back(now_us() - a);
    }
    const double secs = (now_us() - t0) * 1e-6;
    char buf[384];
    std::snprintf(buf, sizeof buf,
                  "{\"req_per_s\": %.1f, \"p50_us\": %.1f, \"p99_us\": %.1f, \"gflops\": %.1f, \"GBs\": %.1f}",
                  nreq / secs, pct(lat, 0.5), pct(lat, 0.99), flops / secs * 1e-9, bytes / secs * 1e-9);
    std::fprintf(stderr, "per-call %s sync %s\n", engine ? "engine" : "plain", buf);
    return std::string(buf);
  };

  auto run_percall_async = [&](bool engine) {
    const double t0 = now_us();
    for (size_t r = 0; r < reqs.size(); r++) {
      const Req& q = reqs[r];
      Mat& M = mats[q.mat];
      CHECK(lespmv_spmm(M.h, LESPMV_CSR, variant_of(M, engine), q.k, 1.0, M.X, q.k, 0.0, (r & 1) ? M.Y2 : M.Y, q.k,
                        nullptr));
    }
    CHECK(lespmv_synchronize());
    const double secs = (now_us() - t0) * 1e-6;
    char buf[256];
    std::snprintf(buf, sizeof buf, "{\"req_per_s\": %.1f, \"gflops\": %.1f, \"GBs\": %.1f}", nreq / secs,
                  flops / secs * 1e-9, bytes / secs * 1e-9);
    std::fprintf(stderr, "per-call %s async %s\n", engine ? "engine" : "plain", buf);
    return std::string(buf);
  };

  // warm-up: every matrix once through both paths
  for (int i = 0; i < nmat; i++) {
    CHECK(lespmv_server_submit(S, mats[i].id, 4, 1.0, mats[i].X, 4, 0.0, mats[i].Y, 4, nullptr));
    CHECK(lespmv_spmm(mats[i].h, LESPMV_CSR, LESPMV_VARIANT_AUTO, 4, 1.0, mats[i].X, 4, 0.0, mats[i].Y2, 4, nullptr));
    CHECK(lespmv_spmm(mats[i].h, LESPMV_CSR, variant_of(mats[i], true), 4, 1.0, mats[i].X, 4, 0.0, mats[i].Y2, 4,
                      nullptr));
  }
  CHECK(lespmv_server_drain(S));
  CHECK(lespmv_synchronize());
  {
    int32_t n = 0;
    std::vector<int64_t> ids(nmat + 16);
    CHECK(lespmv_server_poll(S, 0, ids.data(), nullptr, (int32_t)ids.size(), &n));
  }

  // unloaded latency: one request at a time through each path
  auto run_unloaded = [&](int n) {
    std::vector<double> ls, lp;
    std::vector<int64_t> ids(16);
    std::vector<double> l(16);
    for (int r = 0; r < n; r++) {
      const Req& q = reqs[r];
      Mat& M = mats[q.mat];
      const double a = now_us();
      CHECK(lespmv_server_submit(S, M.id, q.k, 1.0, M.X, q.k, 0.0, M.Y, q.k, nullptr));
      int32_t got = 0;
      while (got == 0) CHECK(lespmv_server_poll(S, 1, ids.data(), l.data(), 16, &got));
      ls.push_back(now_us() - a);
      const double b = now_us();
      CHECK(lespmv_spmm(M.h, LESPMV_CSR, LESPMV_VARIANT_AUTO, q.k, 1.0, M.X, q.k, 0.0, M.Y2, q.k, nullptr));
      CHECK(lespmv_synchronize());
      lp.push_back(now_us() - b);
    }
    char buf[384];
    std::snprintf(buf, sizeof buf,
                  "{\"server_p50_us\": %.1f, \"server_p99_us\": %.1f, \"percall_p50_us\": %.1f, "
                  "\"percall_p99_us\": %.1f}",
                  pct(ls, 0.5), pct(ls, 0.99), pct(lp, 0.5), pct(lp, 0.99));
    std::fprintf(stderr, "unloaded %s\n", buf);
    return std::string(buf);
  };
  const std::string unloaded = run_unloaded(std::min(nreq, 400));
  const std::string sat = run_server(0.0);
  std::vector<std::string> loads;
  for (double r : rates) loads.push_back(run_server(r));
  const std::string sync = run_percall_sync(false);
  const std::string async = run_percall_async(false);
  const std::string esync = run_percall_sync(true);
  const std::string easync = run_percall_async(true);
  CHECK(lespmv_server_stats(S, &st));

  std::printf("{\"matrices\": %d, \"requests\": %d, \"bytes_per_request\": %.0f, \"total_rows\": %lld, \"total_nnz\": %lld, "
              "\"resident_bytes\": %lld, \"build_s\": %.1f, \"formats\": \"%s\", \"sliced_k1_matrices\": %d, "
              "\"bsr_k32_matrices\": %d, \"batches\": %lld, \"server_saturation\": %s, \"server_poisson\": [",
              nmat, nreq, bytes / nreq, (long long)total_rows, (long long)total_nnz, (long long)st.bytes_used, build_s,
              formats.c_str(), sliced_mats, bsr_mats, (long long)st.batches, sat.c_str());
  for (size_t i = 0; i < loads.size(); i++) std::printf("%s%s", i ? ", " : "", loads[i].c_str());
  std::printf("], \"unloaded\": %s, \"percall_sync\": %s, \"percall_async\": %s, \"percall_engine_sync\": %s, "
              "\"percall_engine_async\": %s}\n",
              unloaded.c_str(), sync.c_str(), async.c_str(), esync.c_str(), easync.c_str());
  for (auto& M : mats) {
    lespmv_free(M.X);
    lespmv_free(M.Y);
    lespmv_free(M.Y2);
    lespmv_destroy(M.h);
  }
  lespmv_server_destroy(S);
  return 0;
}
