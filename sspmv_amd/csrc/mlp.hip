// A small multi-layer perceptron trained and evaluated on the GPU: the
// learned format selector's network. No reference counterpart (the
// reference's selector is a Keras model, SmartAdpter/models/SpMV_Adapter.py).
//
// Network: in -> H1 -> H2 -> sum(head sizes), ReLU after the two hidden
// layers, one softmax cross-entropy per head (label -1 = no label), the loss of
// a head averaged over the labelled samples of the batch. Inputs are
// standardised with a stored mean / std inside the batch gather. FP32 math.
//
// Kernels: one LDS-tiled GEMM (64 x 64 output tile, K step 16, 4 x 4 outputs
// per thread, operands in either layout) serves the forward pass
// (X W^T + b, ReLU), the input gradient (dY W) and the weight gradient
// (dY^T X); bias gradients are column sums; ReLU backward masks by the stored
// activation; Adam updates in place. Every reduction runs in a fixed order, so
// training is deterministic for a given seed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace lsp {

namespace {

constexpr int kTM = 64, kTN = 64, kTK = 16, kGemmThreads = 256;

// C[M x N] (row-major, ldc) = op(A) op(B) (+ bias[n]) (ReLU) (+= C)
// op(A) is M x K: A stored M x K (TA = false) or K x M (TA = true).
// op(B) is K x N: B stored K x N (TB = false) or N x K (TB = true).
template <bool TA, bool TB>
__global__ __launch_bounds__(kGemmThreads) void gemm_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                            const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                            int ldc, const float* __restrict__ bias, int relu,
                                                            int accumulate) {
  __shared__ float As[kTK][kTM + 4];
  __shared__ float Bs[kTK][kTN + 4];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += kTK) {
    for (int i = threadIdx.x; i < kTM * kTK; i += kGemmThreads) {
      // consecutive threads walk the contiguous dimension of the stored operand
      const int mm = TA ? i % kTM : i / kTK, kk = TA ? i / kTM : i % kTK;
      const int m = m0 + mm, k = k0 + kk;
      As[kk][mm] = (m < M && k < K) ? (TA ? A[(int64_t)k * lda + m] : A[(int64_t)m * lda + k]) : 0.f;
    }
    for (int i = threadIdx.x; i < kTN * kTK; i += kGemmThreads) {
      const int nn = TB ? i / kTK : i % kTN, kk = TB ? i % kTK : i / kTN;
      const int n = n0 + nn, k = k0 + kk;
      Bs[kk][nn] = (n < N && k < K) ? (TB ? B[(int64_t)n * ldb + k] : B[(int64_t)k * ldb + n]) : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; kk++) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int m = m0 + ty * 4 + i;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int n = n0 + tx * 4 + j;
      if (n >= N) continue;
      float v = acc[i][j];
      if (bias) v += bias[n];
      if (relu) v = fmaxf(v, 0.f);
      float* c = C + (int64_t)m * ldc + n;
      *c = accumulate ? *c + v : v;
    }
  }
}

template <bool TA, bool TB>
void gemm(int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc, const float* bias,
          bool relu, bool accumulate, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  const dim3 grid((unsigned)ceil_div(N, kTN), (unsigned)ceil_div(M, kTM));
  hipLaunchKernelGGL((gemm_kernel<TA, TB>), grid, dim3(kGemmThreads), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias,
                     relu ? 1 : 0, accumulate ? 1 : 0);
  LSP_LAUNCHED();
}

// rows idx[b] of X (n x d), standardised: out[b][j] = (X - mean) / std
__global__ void gather_standardise(int B, int d, const int32_t* __restrict__ idx, const float* __restrict__ X,
                                   const float* __restrict__ mean, const float* __restrict__ inv_std,
                                   float* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (int64_t)B * d) {
    const int b = (int)(t / d), j = (int)(t % d);
    const int64_t row = idx ? idx[b] : b;
    out[t] = (X[row * d + j] - mean[j]) * inv_std[j];
  }
}

__global__ void gather_labels(int B, int h, const int32_t* __restrict__ idx, const int32_t* __restrict__ L,
                              int32_t* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (int64_t)B * h) {
    const int b = (int)(t / h), j = (int)(t % h);
    out[t] = L[(int64_t)(idx ? idx[b] : b) * h + j];
  }
}

struct Heads {
  int n;
  int off[8], size[8];
};

// One thread per (sample, head): softmax, loss, and dLogits = (p - onehot) / count.
__global__ void softmax_ce(int B, int O, Heads hd, const float* __restrict__ logits, const int32_t* __restrict__ labels,
                           const int32_t* __restrict__ counts, float* __restrict__ dlogits, float* __restrict__ loss) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B * hd.n) return;
  const int b = (int)(t / hd.n), h = (int)(t % hd.n);
  const float* z = logits + (int64_t)b * O + hd.off[h];
  float* g = dlogits + (int64_t)b * O + hd.off[h];
  const int lab = labels[(int64_t)b * hd.n + h];
  const int sz = hd.size[h];
  if (lab < 0 || counts[h] == 0) {
    for (int j = 0; j < sz; j++) g[j] = 0.f;
    loss[t] = 0.f;
    return;
  }
  float mx = z[0];
  for (int j = 1; j < sz; j++) mx = fmaxf(mx, z[j]);
  float se = 0.f;
  for (int j = 0; j < sz; j++) se += expf(z[j] - mx);
  const float lse = mx + logf(se);
  const float inv = 1.f / (float)counts[h];
  for (int j = 0; j < sz; j++) g[j] = (expf(z[j] - lse) - (j == lab ? 1.f : 0.f)) * inv;
  loss[t] = (lse - z[lab]) * inv;
}

__global__ void softmax_heads(int B, int O, Heads hd, const float* __restrict__ logits, float* __restrict__ probs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B * hd.n) return;
  const int b = (int)(t / hd.n), h = (int)(t % hd.n);
  const float* z = logits + (int64_t)b * O + hd.off[h];
  float* p = probs + (int64_t)b * O + hd.off[h];
  float mx = z[0];
  for (int j = 1; j < hd.size[h]; j++) mx = fmaxf(mx, z[j]);
  float se = 0.f;
  for (int j = 0; j < hd.size[h]; j++) se += expf(z[j] - mx);
  for (int j = 0; j < hd.size[h]; j++) p[j] = expf(z[j] - mx) / se;
}

// db[o] = sum_b dY[b][o], in order
__global__ void column_sums(int B, int O, const float* __restrict__ dY, float* __restrict__ db) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o < O) {
    float s = 0.f;
    for (int b = 0; b < B; b++) s += dY[(int64_t)b * O + o];
    db[o] = s;
  }
}

__global__ void relu_backward(int64_t n, const float* __restrict__ act, float* __restrict__ g) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && act[i] <= 0.f) g[i] = 0.f;
}

// sum of a batch's per-(sample, head) losses, in order, into slot
__global__ void sum_loss(int n, const float* __restrict__ loss, float* __restrict__ slot) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < n; i++) s += loss[i];
    *slot = s;
  }
}

__global__ void adam_step(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, float lr, float b1, float b2, float eps, float wd, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + wd * p[i]);
  }
}

uint64_t mix64(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
double unit01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct Mlp {
  int in = 0, h1 = 0, h2 = 0, out = 0;
  Heads heads{};
  int64_t nparams = 0;
  int64_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
  DBuf params, grads, adam_m, adam_v, mean, inv_std;
  std::vector<float> h_mean, h_std;
  hipStream_t stream = nullptr;
  int64_t adam_t = 0;
  // activations for a batch of up to `cap` samples
  int cap = 0;
  DBuf x, a1, a2, logits, dlogits, da2, da1, labels, counts, loss;

  ~Mlp() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  float* P() const { return params.as<float>(); }
  float* G() const { return grads.as<float>(); }

  void layout() {
    oW1 = 0;
    ob1 = oW1 + (int64_t)h1 * in;
    oW2 = ob1 + h1;
    ob2 = oW2 + (int64_t)h2 * h1;
    oW3 = ob2 + h2;
    ob3 = oW3 + (int64_t)out * h2;
    nparams = ob3 + out;
  }

  void ensure(int B) {
    if (B <= cap) return;
    cap = B;
    x.alloc(sizeof(float) * (size_t)B * in);
    a1.alloc(sizeof(float) * (size_t)B * h1);
    a2.alloc(sizeof(float) * (size_t)B * h2);
    logits.alloc(sizeof(float) * (size_t)B * out);
    dlogits.alloc(sizeof(float) * (size_t)B * out);
    da2.alloc(sizeof(float) * (size_t)B * h2);
    da1.alloc(sizeof(float) * (size_t)B * h1);
    labels.alloc(sizeof(int32_t) * (size_t)B * heads.n);
    loss.alloc(sizeof(float) * (size_t)B * heads.n);
  }

  // x (B x in, already standardised) -> logits
  void forward(int B) {
    gemm<false, true>(B, h1, in, x.as<float>(), in, P() + oW1, in, a1.as<float>(), h1, P() + ob1, true, false, stream);
    gemm<false, true>(B, h2, h1, a1.as<float>(), h1, P() + oW2, h1, a2.as<float>(), h2, P() + ob2, true, false, stream);
    gemm<false, true>(B, out, h2, a2.as<float>(), h2, P() + oW3, h2, logits.as<float>(), out, P() + ob3, false, false,
                      stream);
  }

  // loss and gradients of the batch in x / labels; counts = labelled samples per head
  void backward(int B) {
    hipLaunchKernelGGL(softmax_ce, grid1d((int64_t)B * heads.n, 256), dim3(256), 0, stream, B, out, heads,
                       logits.as<float>(), labels.as<int32_t>(), counts.as<int32_t>(), dlogits.as<float>(),
                       loss.as<float>());
    LSP_LAUNCHED();
    // layer 3
    gemm<true, false>(out, h2, B, dlogits.as<float>(), out, a2.as<float>(), h2, G() + oW3, h2, nullptr, false, false,
                      stream);
    hipLaunchKernelGGL(column_sums, grid1d(out, 256), dim3(256), 0, stream, B, out, dlogits.as<float>(), G() + ob3);
    LSP_LAUNCHED();
    gemm<false, false>(B, h2, out, dlogits.as<float>(), out, P() + oW3, h2, da2.as<float>(), h2, nullptr, false, false,
                       stream);
    hipLaunchKernelGGL(relu_backward, grid1d((int64_t)B * h2, 256), dim3(256), 0, stream, (int64_t)B * h2,
                       a2.as<float>(), da2.as<float>());
    LSP_LAUNCHED();
    // layer 2
    gemm<true, false>(h2, h1, B, da2.as<float>(), h2, a1.as<float>(), h1, G() + oW2, h1, nullptr, false, false, stream);
    hipLaunchKernelGGL(column_sums, grid1d(h2, 256), dim3(256), 0, stream, B, h2, da2.as<float>(), G() + ob2);
    LSP_LAUNCHED();
    gemm<false, false>(B, h1, h2, da2.as<float>(), h2, P() + oW2, h1, da1.as<float>(), h1, nullptr, false, false,
                       stream);
    hipLaunchKernelGGL(relu_backward, grid1d((int64_t)B * h1, 256), dim3(256), 0, stream, (int64_t)B * h1,
                       a1.as<float>(), da1.as<float>());
    LSP_LAUNCHED();
    // layer 1
    gemm<true, false>(h1, in, B, da1.as<float>(), h1, x.as<float>(), in, G() + oW1, in, nullptr, false, false, stream);
    hipLaunchKernelGGL(column_sums, grid1d(h1, 256), dim3(256), 0, stream, B, h1, da1.as<float>(), G() + ob1);
    LSP_LAUNCHED();
  }

  void load_batch(int B, const int32_t* d_idx, const float* dX, const int32_t* dL) {
    ensure(B);
    hipLaunchKernelGGL(gather_standardise, grid1d((int64_t)B * in, 256), dim3(256), 0, stream, B, in, d_idx, dX,
                       mean.as<float>(), inv_std.as<float>(), x.as<float>());
    LSP_LAUNCHED();
    if (dL) {
      hipLaunchKernelGGL(gather_labels, grid1d((int64_t)B * heads.n, 256), dim3(256), 0, stream, B, heads.n, d_idx,
                         dL, labels.as<int32_t>());
      LSP_LAUNCHED();
    }
  }

  // the per-head label counts of the current batch (host copy kept alive
  // until the stream has consumed it)
  std::vector<int32_t> h_counts;
  void set_counts(const std::vector<int32_t>& c) {
    if (counts.bytes < sizeof(int32_t) * 8) counts.alloc(sizeof(int32_t) * 8);
    LSP_HIP(hipStreamSynchronize(stream));
    h_counts = c;
    LSP_HIP(hipMemcpyAsync(counts.p, h_counts.data(), sizeof(int32_t) * h_counts.size(), hipMemcpyHostToDevice,
                           stream));
  }

  void upload_normaliser() {
    std::vector<float> inv(in);
    for (int j = 0; j < in; j++) inv[j] = h_std[j] > 0.f ? 1.f / h_std[j] : 1.f;
    LSP_HIP(hipMemcpyAsync(mean.p, h_mean.data(), sizeof(float) * in, hipMemcpyHostToDevice, stream));
    LSP_HIP(hipMemcpyAsync(inv_std.p, inv.data(), sizeof(float) * in, hipMemcpyHostToDevice, stream));
    LSP_HIP(hipStreamSynchronize(stream));
  }
};

Mlp* make_mlp(int in, int h1, int h2, int nheads, const int32_t* sizes) {
  LSP_REQUIRE(in > 0 && h1 > 0 && h2 > 0 && nheads > 0 && nheads <= 8 && sizes, LESPMV_ERR_ARG, "bad MLP shape");
  auto* M = new Mlp();
  std::unique_ptr<Mlp> guard(M);
  M->in = in;
  M->h1 = h1;
  M->h2 = h2;
  M->heads.n = nheads;
  int off = 0;
  for (int h = 0; h < nheads; h++) {
    LSP_REQUIRE(sizes[h] > 0, LESPMV_ERR_ARG, "head sizes must be positive");
    M->heads.off[h] = off;
    M->heads.size[h] = sizes[h];
    off += sizes[h];
  }
  M->out = off;
  M->layout();
  LSP_HIP(hipStreamCreate(&M->stream));
  M->params.alloc(sizeof(float) * M->nparams);
  M->grads.alloc(sizeof(float) * M->nparams);
  M->adam_m.alloc(sizeof(float) * M->nparams);
  M->adam_v.alloc(sizeof(float) * M->nparams);
  M->mean.alloc(sizeof(float) * in);
  M->inv_std.alloc(sizeof(float) * in);
  M->h_mean.assign(in, 0.f);
  M->h_std.assign(in, 1.f);
  M->upload_normaliser();
  return guard.release();
}

void init_params(Mlp& M, uint64_t seed) {
  std::vector<float> p(M.nparams, 0.f);
  auto fill = [&](int64_t off, int rows, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int64_t i = 0; i < (int64_t)rows * fan_in; i++)
      p[off + i] = (float)((2.0 * unit01(mix64(seed ^ (uint64_t)(off + i))) - 1.0) * a);
  };
  fill(M.oW1, M.h1, M.in);
  fill(M.oW2, M.h2, M.h1);
  fill(M.oW3, M.out, M.h2);
  LSP_HIP(hipMemcpy(M.params.p, p.data(), sizeof(float) * M.nparams, hipMemcpyHostToDevice));
  LSP_HIP(hipMemset(M.adam_m.p, 0, sizeof(float) * M.nparams));
  LSP_HIP(hipMemset(M.adam_v.p, 0, sizeof(float) * M.nparams));
  M.adam_t = 0;
}

std::vector<int32_t> head_counts(const Mlp& M, const int32_t* labels, const int32_t* idx, int B) {
  std::vector<int32_t> c(8, 0);
  for (int b = 0; b < B; b++) {
    const int64_t row = idx ? idx[b] : b;
    for (int h = 0; h < M.heads.n; h++) c[h] += labels[row * M.heads.n + h] >= 0;
  }
  return c;
}

}  // namespace lsp

using namespace lsp;

namespace {
Mlp& mlp_of(lespmv_mlp_t h) {
  LSP_REQUIRE(h, LESPMV_ERR_ARG, "null MLP handle");
  return *reinterpret_cast<Mlp*>(h);
}
const char kMagic[8] = {'L', 'S', 'P', 'M', 'L', 'P', '1', 0};
}  // namespace

extern "C" int lespmv_mlp_create(lespmv_mlp_t* out, int32_t in_dim, int32_t hidden1, int32_t hidden2, int32_t nheads,
                                 const int32_t* head_sizes, uint64_t seed) {
  return guarded_call([&] {
    LSP_REQUIRE(out, LESPMV_ERR_ARG, "null output");
    Mlp* M = make_mlp(in_dim, hidden1, hidden2, nheads, head_sizes);
    std::unique_ptr<Mlp> guard(M);
    init_params(*M, seed);
    *out = reinterpret_cast<lespmv_mlp_t>(guard.release());
  });
}

extern "C" int lespmv_mlp_destroy(lespmv_mlp_t h) {
  return guarded_call([&] { delete reinterpret_cast<Mlp*>(h); });
}

extern "C" int lespmv_mlp_shape(lespmv_mlp_t h, int32_t* in_dim, int32_t* out_dim, int64_t* nparams) {
  return guarded_call([&] {
    Mlp& M = mlp_of(h);
    if (in_dim) *in_dim = M.in;
    if (out_dim) *out_dim = M.out;
    if (nparams) *nparams = M.nparams;
  });
}

extern "C" int lespmv_mlp_get_params(lespmv_mlp_t h, float* host) {
  return guarded_call([&] {
    Mlp& M = mlp_of(h);
    LSP_HIP(hipMemcpy(host, M.params.p, sizeof(float) * M.nparams, hipMemcpyDeviceToHost));
  });
}

extern "C" int lespmv_mlp_set_params(lespmv_mlp_t h, const float* host) {
  return guarded_call([&] {
    Mlp& M = mlp_of(h);
    LSP_HIP(hipMemcpy(M.params.p, host, sizeof(float) * M.nparams, hipMemcpyHostToDevice));
  });
}

extern "C" int lespmv_mlp_set_normalizer(lespmv_mlp_t h, const float* mean, const float* stddev) {
  return guarded_call([&] {
    Mlp& M = mlp_of(h);
    LSP_REQUIRE(mean && stddev, LESPMV_ERR_ARG, "null normaliser");
    M.h_mean.assign(mean, mean + M.in);
    M.h_std.assign(stddev, stddev + M.in);
    M.upload_normaliser();
  });
}

extern "C" int lespmv_mlp_loss_grad(lespmv_mlp_t h, const float* X, const int32_t* labels, int32_t n, double* loss,
                                    float* grad) {
  return guarded_call([&] {
    Mlp& M = mlp_of(h);
    LSP_REQUIRE(X && labels && n > 0, LESPMV_ERR_ARG, "bad batch");
    DBuf dX, dL;
    dX.alloc(sizeof(float) * (size_t)n * M.in);
    dL.alloc(sizeof(int32_t) * (size_t)n * M.heads.n);
    LSP_HIP(hipMemcpy(dX.p, X, dX.bytes, hipMemcpyHostToDevice));
    LSP_HIP(hipMemcpy(dL.p, labels, dL.bytes, hipMemcpyHostToDevice));
    M.load_batch(n, nullptr, dX.as<float>(), dL.as<int32_t>());
    M.set_counts(head_counts(M, labels, nullptr, n));
    M.forward(n);
    M.backward(n);
    std::vector<float> l((size_t)n * M.heads.n);
    LSP_HIP(hipMemcpyAsync(l.data(), M.loss.p, sizeof(float) * l.size(), hipMemcpyDeviceToHost, M.stream));
    if (grad) LSP_HIP(hipMemcpyAsync(grad, M.grads.p, sizeof(float) * M.nparams, hipMemcpyDeviceToHost, M.stream));
    LSP_HIP(hipStreamSynchronize(M.stream));
    double s = 0.0;
    for (float v : l) s += v;
    if (loss) *loss = s;
  });
}

extern "C" int lespmv_mlp_train(lespmv_mlp_t h, const float* X, const int32_t* labels, int32_t n, int32_t epochs,
                                int32_t batch, float lr, float weight_decay, uint64_t seed, float* epoch_loss) {
  return guarded_call([&] {
    Mlp& M = mlp_of(h);
    LSP_REQUIRE(X && labels && n > 0 && epochs > 0 && batch > 0, LESPMV_ERR_ARG, "bad training arguments");
    DBuf dX, dL, dIdx, dSlots;
    dX.alloc(sizeof(float) * (size_t)n * M.in);
    dL.alloc(sizeof(int32_t) * (size_t)n * M.heads.n);
    dIdx.alloc(sizeof(int32_t) * (size_t)n);
    const int nb = (int)ceil_div(n, batch);
    dSlots.alloc(sizeof(float) * (size_t)nb);
    LSP_HIP(hipMemcpy(dX.p, X, dX.bytes, hipMemcpyHostToDevice));
    LSP_HIP(hipMemcpy(dL.p, labels, dL.bytes, hipMemcpyHostToDevice));
    std::vector<int32_t> perm(n);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; e++) {
      for (int i = 0; i < n; i++) perm[i] = i;
      for (int i = n - 1; i > 0; i--) {  // Fisher-Yates with a counter-based stream
        const int j = (int)(mix64(seed ^ ((uint64_t)e << 32) ^ (uint64_t)i) % (uint64_t)(i + 1));
        std::swap(perm[i], perm[j]);
      }
      LSP_HIP(hipMemcpyAsync(dIdx.p, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, M.stream));
      for (int b = 0; b < nb; b++) {
        const int start = b * batch, B = std::min(batch, n - start);
        const int32_t* idx = dIdx.as<int32_t>() + start;
        M.load_batch(B, idx, dX.as<float>(), dL.as<int32_t>());
        M.set_counts(head_counts(M, labels, perm.data() + start, B));
        M.forward(B);
        M.backward(B);
        hipLaunchKernelGGL(sum_loss, dim3(1), dim3(64), 0, M.stream, B * M.heads.n, M.loss.as<float>(),
                           dSlots.as<float>() + b);
        LSP_LAUNCHED();
        M.adam_t++;
        const float c1 = 1.f - std::pow(b1, (float)M.adam_t), c2 = 1.f - std::pow(b2, (float)M.adam_t);
        hipLaunchKernelGGL(adam_step, grid1d(M.nparams, 256), dim3(256), 0, M.stream, M.nparams, M.P(), M.G(),
                           M.adam_m.as<float>(), M.adam_v.as<float>(), lr, b1, b2, eps, weight_decay, c1, c2);
        LSP_LAUNCHED();
      }
      LSP_HIP(hipStreamSynchronize(M.stream));
      std::vector<float> slots(nb);
      LSP_HIP(hipMemcpy(slots.data(), dSlots.p, sizeof(float) * nb, hipMemcpyDeviceToHost));
      double s = 0.0;
      for (float v : slots) s += v;
      if (epoch_loss) epoch_loss[e] = (float)(s / nb);
    }
  });
}

extern "C" int lespmv_mlp_forward(lespmv_mlp_t h, const float* X, int32_t n, float* probs) {
  return guarded_call([&] {
    Mlp& M = mlp_of(h);
    LSP_REQUIRE(X && probs && n > 0, LESPMV_ERR_ARG, "bad forward arguments");
    DBuf dX, dP;
    dX.alloc(sizeof(float) * (size_t)n * M.in);
    dP.alloc(sizeof(float) * (size_t)n * M.out);
    LSP_HIP(hipMemcpy(dX.p, X, dX.bytes, hipMemcpyHostToDevice));
    M.load_batch(n, nullptr, dX.as<float>(), nullptr);
    M.forward(n);
    hipLaunchKernelGGL(softmax_heads, grid1d((int64_t)n * M.heads.n, 256), dim3(256), 0, M.stream, n, M.out, M.heads,
                       M.logits.as<float>(), dP.as<float>());
    LSP_LAUNCHED();
    LSP_HIP(hipMemcpyAsync(probs, dP.p, dP.bytes, hipMemcpyDeviceToHost, M.stream));
    LSP_HIP(hipStreamSynchronize(M.stream));
  });
}

extern "C" int lespmv_mlp_save(lespmv_mlp_t h, const char* path) {
  return guarded_call([&] {
    Mlp& M = mlp_of(h);
    std::vector<float> p(M.nparams);
    LSP_HIP(hipMemcpy(p.data(), M.params.p, sizeof(float) * M.nparams, hipMemcpyDeviceToHost));
    FILE* f = std::fopen(path, "wb");
    LSP_REQUIRE(f, LESPMV_ERR_ARG, std::string("cannot write ") + path);
    std::unique_ptr<FILE, int (*)(FILE*)> g(f, std::fclose);
    const int32_t dims[4] = {M.in, M.h1, M.h2, M.heads.n};
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(dims, 4, 4, f) == 4 &&
              std::fwrite(M.heads.size, 4, M.heads.n, f) == (size_t)M.heads.n &&
              std::fwrite(M.h_mean.data(), 4, M.in, f) == (size_t)M.in &&
              std::fwrite(M.h_std.data(), 4, M.in, f) == (size_t)M.in &&
              std::fwrite(p.data(), 4, p.size(), f) == p.size();
    LSP_REQUIRE(ok, LESPMV_ERR_ARG, std::string("short write to ") + path);
  });
}

extern "C" int lespmv_mlp_load(lespmv_mlp_t* out, const char* path) {
  return guarded_call([&] {
    LSP_REQUIRE(out && path, LESPMV_ERR_ARG, "null argument");
    FILE* f = std::fopen(path, "rb");
    LSP_REQUIRE(f, LESPMV_ERR_ARG, std::string("cannot open ") + path);
    std::unique_ptr<FILE, int (*)(FILE*)> g(f, std::fclose);
    char magic[8];
    int32_t dims[4];
    LSP_REQUIRE(std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
                    std::fread(dims, 4, 4, f) == 4 && dims[3] > 0 && dims[3] <= 8,
                LESPMV_ERR_ARG, std::string("not an MLP file: ") + path);
    int32_t sizes[8];
    LSP_REQUIRE(std::fread(sizes, 4, dims[3], f) == (size_t)dims[3], LESPMV_ERR_ARG, "truncated MLP file");
    Mlp* M = make_mlp(dims[0], dims[1], dims[2], dims[3], sizes);
    std::unique_ptr<Mlp> guard(M);
    std::vector<float> p(M->nparams);
    LSP_REQUIRE(std::fread(M->h_mean.data(), 4, M->in, f) == (size_t)M->in &&
                    std::fread(M->h_std.data(), 4, M->in, f) == (size_t)M->in &&
                    std::fread(p.data(), 4, p.size(), f) == p.size(),
                LESPMV_ERR_ARG, "truncated MLP file");
    M->upload_normaliser();
    LSP_HIP(hipMemcpy(M->params.p, p.data(), sizeof(float) * M->nparams, hipMemcpyHostToDevice));
    *out = reinterpret_cast<lespmv_mlp_t>(guard.release());
  });
}
