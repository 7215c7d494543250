// Synthetic CSR inputs (host side). Every random draw comes from a
// counter-based splitmix64 stream, so a matrix depends only on its parameters
// and seed, whatever the thread count:
//   u(h) = (h >> 11) * 2^-53,  U[lo,hi) = lo + (hi-lo)*u.
// Columns inside a row are sorted ascending and distinct.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <thread>
#include <vector>

#include "common.hpp"

namespace lsp {
namespace {

inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
inline double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

// Sequential stream for one row: state advanced by the golden gamma.
struct RowRng {
  uint64_t s;
  RowRng(uint64_t seed, uint64_t row) : s(splitmix64(seed * 0x632be59bd9b4e019ULL ^ row)) {}
  uint64_t next() {
    s += 0x9e3779b97f4a7c15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double u() { return unit(next()); }
  int64_t below(int64_t n) { return (int64_t)(u() * (double)n); }
  double normal() {
    const double u1 = u(), u2 = u();
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
  }
};

template <class F>
void parallel_for(int64_t n, F&& f) {
  int threads = (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
  if (n < 4096) threads = 1;
  std::vector<std::thread> pool;
  const int64_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; t++) {
    const int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& th : pool) th.join();
}

struct Builder {
  int64_t m = 0, n = 0;
  std::vector<int64_t> len;  // row lengths, then prefix offsets
  std::vector<int32_t> rp, ci;
  std::vector<double> val;

  void offsets() {
    rp.assign(m + 1, 0);
    int64_t acc = 0;
    for (int64_t i = 0; i < m; i++) {
      rp[i] = (int32_t)acc;
      acc += len[i];
      LSP_REQUIRE(acc < INT32_MAX, LESPMV_ERR_UNSUPPORTED, "generated matrix exceeds 32-bit nonzero count");
    }
    rp[m] = (int32_t)acc;
    ci.resize(acc);
    val.resize(acc);
  }
  // values U[-1,1) indexed by global nonzero position
  void random_values(uint64_t seed) {
    parallel_for((int64_t)val.size(), [&](int64_t b, int64_t e) {
      for (int64_t k = b; k < e; k++) val[k] = 2.0 * unit(splitmix64(seed ^ (uint64_t)k)) - 1.0;
    });
  }
};

// Draw `r` distinct sorted columns for a row; draw() gives one candidate.
template <class Draw>
void distinct_sorted(int32_t* out, int64_t r, Draw&& draw) {
  for (int64_t j = 0; j < r; j++) out[j] = draw();
  for (int attempt = 0; attempt < 1000; attempt++) {
    std::sort(out, out + r);
    bool dup = false;
    for (int64_t j = 1; j < r; j++)
      if (out[j] == out[j - 1]) { out[j] = draw(); dup = true; }
    if (!dup) return;
  }
  std::sort(out, out + r);
}

void gen_laplace2d(Builder& B, int64_t nx, int64_t ny, double diag, double off) {
  B.m = B.n = nx * ny;
  B.len.assign(B.m, 0);
  parallel_for(B.m, [&](int64_t b, int64_t e) {
    for (int64_t r = b; r < e; r++) {
      const int64_t i = r % nx, j = r / nx;
      B.len[r] = 1 + (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1);
    }
  });
  B.offsets();
  parallel_for(B.m, [&](int64_t b, int64_t e) {
    for (int64_t r = b; r < e; r++) {
      const int64_t i = r % nx, j = r / nx;
      int64_t k = B.rp[r];
      auto put = [&](int64_t c, double v) { B.ci[k] = (int32_t)c; B.val[k] = v; k++; };
      if (j > 0) put(r - nx, off);
      if (i > 0) put(r - 1, off);
      put(r, diag);
      if (i < nx - 1) put(r + 1, off);
      if (j < ny - 1) put(r + nx, off);
    }
  });
}

void gen_stencil27(Builder& B, int64_t n, double diag, double off) {
  B.m = B.n = n * n * n;
  B.len.assign(B.m, 0);
  auto span = [n](int64_t c) { return (int64_t)1 + (c > 0) + (c < n - 1); };
  parallel_for(B.m, [&](int64_t b, int64_t e) {
    for (int64_t r = b; r < e; r++) B.len[r] = span(r % n) * span((r / n) % n) * span(r / (n * n));
  });
  B.offsets();
  parallel_for(B.m, [&](int64_t b, int64_t e) {
    for (int64_t r = b; r < e; r++) {
      const int64_t x = r % n, y = (r / n) % n, z = r / (n * n);
      int64_t k = B.rp[r];
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            const int64_t xx = x + dx, yy = y + dy, zz = z + dz;
            if (xx < 0 || yy < 0 || zz < 0 || xx >= n || yy >= n || zz >= n) continue;
            B.ci[k] = (int32_t)((zz * n + yy) * n + xx);
            B.val[k] = (dx == 0 && dy == 0 && dz == 0) ? diag : off;
            k++;
          }
    }
  });
}

// Row lengths r = clamp(round(exp(mu + sigma*Z)), 1, cap), mu = ln(mean) - sigma^2/2;
// a fraction `frac_local` of the columns lies within +-band of the diagonal,
// the rest is uniform over [0, n).
void gen_lognormal(Builder& B, int64_t m, int64_t n, double mean, double sigma, int64_t cap, double frac_local,
                   int64_t band, uint64_t seed) {
  B.m = m;
  B.n = n;
  B.len.assign(m, 0);
  const double mu = std::log(mean) - 0.5 * sigma * sigma;
  const int64_t maxlen = std::min<int64_t>(cap, n);
  parallel_for(m, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      const double u1 = unit(splitmix64(seed ^ (2 * (uint64_t)i))), u2 = unit(splitmix64(seed ^ (2 * (uint64_t)i + 1)));
      const double z = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
      int64_t r = (int64_t)std::llround(std::exp(mu + sigma * z));
      B.len[i] = std::max<int64_t>(1, std::min(r, maxlen));
    }
  });
  B.offsets();
  parallel_for(m, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      RowRng rng(seed, (uint64_t)i);
      const int64_t r = B.len[i];
      // a row cannot have more local candidates than its window holds
      const int64_t lo = std::max<int64_t>(0, i - band), hi = std::min<int64_t>(n - 1, i + band);
      distinct_sorted(&B.ci[B.rp[i]], r, [&]() -> int32_t {
        if (rng.u() < frac_local && hi >= lo) return (int32_t)(lo + rng.below(hi - lo + 1));
        return (int32_t)rng.below(n);
      });
    }
  });
  B.random_values(seed + 1);
}

void gen_blockrand(Builder& B, int64_t mb, int64_t nb, int64_t br, int64_t bc, int64_t bpr, uint64_t seed) {
  B.m = mb * br;
  B.n = nb * bc;
  bpr = std::min(bpr, nb);
  std::vector<int32_t> bcols(mb * bpr);
  parallel_for(mb, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      RowRng rng(seed, (uint64_t)i);
      distinct_sorted(&bcols[i * bpr], bpr, [&]() -> int32_t { return (int32_t)rng.below(nb); });
    }
  });
  B.len.assign(B.m, bpr * bc);
  B.offsets();
  parallel_for(B.m, [&](int64_t b, int64_t e) {
    for (int64_t r = b; r < e; r++) {
      int64_t k = B.rp[r];
      const int32_t* bc_row = &bcols[(r / br) * bpr];
      for (int64_t t = 0; t < bpr; t++)
        for (int64_t c = 0; c < bc; c++) B.ci[k++] = (int32_t)(bc_row[t] * bc + c);
    }
  });
  B.random_values(seed + 1);
}

void gen_banded(Builder& B, int64_t m, int64_t hb, double fill, uint64_t seed) {
  B.m = B.n = m;
  B.len.assign(m, 0);
  // decide membership with a hash of (row, offset) so both passes agree
  auto keep = [&](int64_t i, int64_t d) {
    return d == 0 || unit(splitmix64(seed ^ ((uint64_t)i * 0x100000001b3ULL + (uint64_t)(d + hb)))) < fill;
  };
  parallel_for(m, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      int64_t c = 0;
      for (int64_t d = -hb; d <= hb; d++)
        if (i + d >= 0 && i + d < m && keep(i, d)) c++;
      B.len[i] = c;
    }
  });
  B.offsets();
  parallel_for(m, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      int64_t k = B.rp[i];
      for (int64_t d = -hb; d <= hb; d++)
        if (i + d >= 0 && i + d < m && keep(i, d)) B.ci[k++] = (int32_t)(i + d);
    }
  });
  B.random_values(seed + 1);
}

// Power-law row lengths: r = clamp(floor(xmin * (1-u)^(-1/(a-1))), 1, max).
void gen_powerlaw(Builder& B, int64_t m, double a, double mean, int64_t maxlen, uint64_t seed) {
  B.m = B.n = m;
  maxlen = std::min<int64_t>(maxlen, m);
  const double xmin = a > 2.0 ? std::max(1.0, mean * (a - 2.0) / (a - 1.0)) : 1.0;
  B.len.assign(m, 0);
  parallel_for(m, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      const double u = unit(splitmix64(seed ^ (uint64_t)i));
      const double r = std::floor(xmin * std::pow(1.0 - u, -1.0 / (a - 1.0)));
      B.len[i] = std::max<int64_t>(1, std::min<int64_t>(maxlen, (int64_t)std::min(r, 1e12)));
    }
  });
  B.offsets();
  parallel_for(m, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      RowRng rng(seed, (uint64_t)i);
      distinct_sorted(&B.ci[B.rp[i]], B.len[i], [&]() -> int32_t { return (int32_t)rng.below(m); });
    }
  });
  B.random_values(seed + 1);
}

void gen_blockdiag(Builder& B, int64_t m, int64_t blk, uint64_t seed) {
  B.m = B.n = m;
  B.len.assign(m, 0);
  parallel_for(m, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      const int64_t s = (i / blk) * blk;
      B.len[i] = std::min(m, s + blk) - s;
    }
  });
  B.offsets();
  parallel_for(m, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      const int64_t s = (i / blk) * blk;
      int64_t k = B.rp[i];
      for (int64_t c = s; c < std::min(m, s + blk); c++) B.ci[k++] = (int32_t)c;
    }
  });
  B.random_values(seed + 1);
}

void gen_uniform(Builder& B, int64_t m, int64_t n, double density, uint64_t seed) {
  B.m = m;
  B.n = n;
  const double mean = density * (double)n, sd = std::sqrt(std::max(mean * (1.0 - density), 0.0));
  B.len.assign(m, 0);
  parallel_for(m, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      RowRng rng(seed ^ 0x5bd1e995ULL, (uint64_t)i);
      const int64_t r = (int64_t)std::llround(mean + sd * rng.normal());
      B.len[i] = std::max<int64_t>(0, std::min<int64_t>(r, n));
    }
  });
  B.offsets();
  parallel_for(m, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      RowRng rng(seed, (uint64_t)i);
      distinct_sorted(&B.ci[B.rp[i]], B.len[i], [&]() -> int32_t { return (int32_t)rng.below(n); });
    }
  });
  B.random_values(seed + 1);
}

}  // namespace

extern thread_local std::string g_last_error;

}  // namespace lsp

extern "C" int lespmv_generate(int family, const double* prm, int32_t np, uint64_t seed, int dtype,
                               lespmv_host_csr_t* out) {
  try {
    LSP_REQUIRE(out && (np == 0 || prm), LESPMV_ERR_ARG, "null argument");
    LSP_REQUIRE(dtype == LESPMV_F32 || dtype == LESPMV_F64, LESPMV_ERR_ARG, "bad dtype");
    auto P = [&](int i, double def) { return i < np ? prm[i] : def; };
    lsp::Builder B;
    switch (family) {
      case LESPMV_GEN_LAPLACE2D:
        lsp::gen_laplace2d(B, (int64_t)P(0, 1024), (int64_t)P(1, P(0, 1024)), P(2, 4.0), P(3, -1.0));
        break;
      case LESPMV_GEN_STENCIL27:
        lsp::gen_stencil27(B, (int64_t)P(0, 160), P(1, 26.0), P(2, -1.0));
        break;
      case LESPMV_GEN_LOGNORMAL:
        lsp::gen_lognormal(B, (int64_t)P(0, 4194304), (int64_t)P(1, P(0, 4194304)), P(2, 16.0), P(3, 1.0),
                           (int64_t)P(4, 4096), P(5, 0.3), (int64_t)P(6, 64), seed);
        break;
      case LESPMV_GEN_BLOCKRAND:
        lsp::gen_blockrand(B, (int64_t)P(0, 1024), (int64_t)P(1, P(0, 1024)), (int64_t)P(2, 4), (int64_t)P(3, 4),
                           (int64_t)P(4, 8), seed);
        break;
      case LESPMV_GEN_BANDED:
        lsp::gen_banded(B, (int64_t)P(0, 100000), (int64_t)P(1, 8), P(2, 1.0), seed);
        break;
      case LESPMV_GEN_POWERLAW:
        lsp::gen_powerlaw(B, (int64_t)P(0, 100000), P(1, 2.0), P(2, 8.0), (int64_t)P(3, 4096), seed);
        break;
      case LESPMV_GEN_BLOCKDIAG:
        lsp::gen_blockdiag(B, (int64_t)P(0, 100000), (int64_t)P(1, 16), seed);
        break;
      case LESPMV_GEN_UNIFORM:
        lsp::gen_uniform(B, (int64_t)P(0, 100000), (int64_t)P(1, P(0, 100000)), P(2, 1e-4), seed);
        break;
      default:
        throw lsp::Error(LESPMV_ERR_ARG, "unknown generator family");
    }
    out->m = B.m;
    out->n = B.n;
    out->nnz = (int64_t)B.ci.size();
    out->dtype = dtype;
    out->mm_flags = 0;
    out->row_offset = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * (B.m + 1)));
    out->col_index = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * std::max<size_t>(B.ci.size(), 1)));
    out->values = std::malloc(lsp::vsize(dtype) * std::max<size_t>(B.val.size(), 1));
    LSP_REQUIRE(out->row_offset && out->col_index && out->values, LESPMV_ERR_OOM, "host allocation failed");
    std::memcpy(out->row_offset, B.rp.data(), sizeof(int32_t) * (B.m + 1));
    if (!B.ci.empty()) std::memcpy(out->col_index, B.ci.data(), sizeof(int32_t) * B.ci.size());
    if (dtype == LESPMV_F64) {
      if (!B.val.empty()) std::memcpy(out->values, B.val.data(), sizeof(double) * B.val.size());
    } else {
      float* f = static_cast<float*>(out->values);
      for (size_t k = 0; k < B.val.size(); k++) f[k] = (float)B.val[k];
    }
    return LESPMV_OK;
  } catch (const lsp::Error& e) {
    lsp::g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    lsp::g_last_error = e.what();
    return LESPMV_ERR_OOM;
  }
}

extern "C" int lespmv_host_csr_free(lespmv_host_csr_t* c) {
  if (!c) return LESPMV_ERR_ARG;
  std::free(c->row_offset);
  std::free(c->col_index);
  std::free(c->values);
  c->row_offset = nullptr;
  c->col_index = nullptr;
  c->values = nullptr;
  return LESPMV_OK;
}

extern "C" int lespmv_fill_uniform(void* dst, int dtype, int64_t n, uint64_t seed, double lo, double hi) {
  if (!dst && n) return LESPMV_ERR_ARG;
  lsp::parallel_for(n, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      const double v = lo + (hi - lo) * lsp::unit(lsp::splitmix64(seed ^ (uint64_t)i));
      if (dtype == LESPMV_F64) static_cast<double*>(dst)[i] = v;
      else static_cast<float*>(dst)[i] = (float)v;
    }
  });
  return LESPMV_OK;
}
