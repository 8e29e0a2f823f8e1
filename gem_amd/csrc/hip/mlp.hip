// Order-cost model: a 3-layer MLP (in -> 128 -> 128 -> 1, ReLU) with bf16
// weights and fp32 accumulation, trained and evaluated on the GPU by the
// kernels below (no library GEMM).  It predicts the cost of a join order from
// int32 order features; the batched path uses it to pick among candidate
// orders (the reference has no learned component; GEM's order stays a
// candidate and the fallback).
//
// Parameters (fp32 master copy, flat): W1[in x h], b1[h], W2[h x h], b2[h],
// W3[h], b3[1]; forward and backward use bf16 copies of W1..W3 (rounded to
// nearest even).  Loss: mean squared error / 2.  Optimiser: Adam.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "../host/host_graph.h"
#include "ctx.h"

namespace gemd {

// fp32 -> bf16 bits, round to nearest even (NaN kept quiet)
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t x = __builtin_bit_cast(uint32_t, f);
  if ((x & 0x7F800000u) == 0x7F800000u && (x & 0x007FFFFFu)) return (uint16_t)((x >> 16) | 0x40);
  x += 0x7FFFu + ((x >> 16) & 1u);
  return (uint16_t)(x >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

constexpr int TM = 64, TN = 64, TK = 16;

// C[M x N] = op(A) op(B) (+ bias[N]) with an epilogue:
//   EPI 0: store; 1: bias + ReLU; 2: bias; 3: multiply by (mask[m][n] > 0)
// A is fp32: element (m, k) at A[m*lda + k] (TA = false) or A[k*lda + m].
// B is fp32 or bf16 (BBF): element (k, n) at B[k*ldb + n] (TB = false) or
// B[n*ldb + k].  256 threads, each a 4x4 block of C.
template <bool TA, bool TB, bool BBF, int EPI>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const float *__restrict__ A, int lda,
                                             const void *__restrict__ Bv, int ldb, const float *__restrict__ bias,
                                             const float *__restrict__ mask, float *__restrict__ C, int ldc) {
  __shared__ float As[TK][TM + 1];
  __shared__ float Bs[TK][TN + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  float acc[4][4] = {};
  const float *Bf = reinterpret_cast<const float *>(Bv);
  const uint16_t *Bh = reinterpret_cast<const uint16_t *>(Bv);
  for (int k0 = 0; k0 < K; k0 += TK) {
    for (int i = threadIdx.x; i < TK * TM; i += 256) {
      const int kk = i / TM, mm = i % TM;
      const int m = m0 + mm, k = k0 + kk;
      float v = 0.f;
      if (m < M && k < K) v = TA ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k];
      As[kk][mm] = v;
    }
    for (int i = threadIdx.x; i < TK * TN; i += 256) {
      const int kk = i / TN, nn = i % TN;
      const int n = n0 + nn, k = k0 + kk;
      float v = 0.f;
      if (n < N && k < K) {
        const size_t idx = TB ? (size_t)n * ldb + k : (size_t)k * ldb + n;
        v = BBF ? bf16_to_f32(Bh[idx]) : Bf[idx];
      }
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + ty * 4 + i;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + tx * 4 + j;
      if (n >= N) continue;
      float v = acc[i][j];
      if (EPI == 1) v = fmaxf(v + bias[n], 0.f);
      if (EPI == 2) v = v + bias[n];
      if (EPI == 3) v = mask[(size_t)m * ldc + n] > 0.f ? v : 0.f;
      C[(size_t)m * ldc + n] = v;
    }
  }
}

// column sums of X[rows x cols] into out[cols]
__global__ void k_colsum(const float *__restrict__ X, int rows, int cols, float *__restrict__ out) {
  const int c = blockIdx.x;
  float s = 0.f;
  for (int r = threadIdx.x; r < rows; r += blockDim.x) s += X[(size_t)r * cols + c];
  __shared__ float sm[256];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sm[0];
}

__global__ void k_feat_to_f32(const int32_t *__restrict__ f, size_t n, float scale, float *__restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (float)f[i] * scale;
}

// dY = (Y - T) / n ; partial loss sums (one per block)
__global__ void k_mse_grad(const float *__restrict__ Y, const float *__restrict__ T, int n, float *__restrict__ dY,
                           float *__restrict__ part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < n) {
    const float d = Y[i] - T[i];
    dY[i] = d / (float)n;
    l = 0.5f * d * d / (float)n;
  }
  __shared__ float sm[256];
  sm[threadIdx.x] = l;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       size_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_to_bf16(const float *__restrict__ src, size_t n, uint16_t *__restrict__ dst) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f32_to_bf16(src[i]);
}

}  // namespace gemd

using namespace gemd;

struct gem_mlp {
  gem_ctx *ctx;
  uint32_t in, h;
  size_t P;
  // offsets into the flat parameter vector
  size_t oW1, ob1, oW2, ob2, oW3, ob3;
  DevBuf params, grads, adam_m, adam_v, wbf;  // wbf: bf16 W1 | W2 | W3
  uint64_t adam_t = 0;
  float margin = 0.f;  // a candidate replaces GEM's order only if predicted cheaper by more
};

namespace {

void refresh_bf16(gem_mlp *m) {
  hipStream_t s = m->ctx->stream;
  const float *p = m->params.as<float>();
  uint16_t *b = m->wbf.as<uint16_t>();
  const size_t n1 = (size_t)m->in * m->h, n2 = (size_t)m->h * m->h, n3 = m->h;
  k_to_bf16<<<(uint32_t)((n1 + 255) / 256), 256, 0, s>>>(p + m->oW1, n1, b);
  k_to_bf16<<<(uint32_t)((n2 + 255) / 256), 256, 0, s>>>(p + m->oW2, n2, b + n1);
  k_to_bf16<<<(uint32_t)((n3 + 255) / 256), 256, 0, s>>>(p + m->oW3, n3, b + n1 + n2);
  GEM_LAUNCH_CHECK("k_to_bf16");
}

template <bool TA, bool TB, bool BBF, int EPI>
void gemm(hipStream_t s, int M, int N, int K, const float *A, int lda, const void *B, int ldb, const float *bias,
          const float *mask, float *C, int ldc) {
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
  k_gemm<TA, TB, BBF, EPI><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, bias, mask, C, ldc);
  GEM_LAUNCH_CHECK("k_gemm");
}

constexpr float FEAT_SCALE = 1.0f / 64.0f;

struct Acts {
  DevBuf X, H1, H2, Y;
};

void forward(gem_mlp *m, const int32_t *h_feat, uint32_t n, Acts &a) {
  hipStream_t s = m->ctx->stream;
  const size_t nf = (size_t)n * m->in;
  DevBuf fi;
  fi.alloc(nf * 4, s);
  GEM_HIP_CHECK(hipMemcpyAsync(fi.p, h_feat, nf * 4, hipMemcpyHostToDevice, s));
  a.X.alloc(nf * 4, s);
  a.H1.alloc((size_t)n * m->h * 4, s);
  a.H2.alloc((size_t)n * m->h * 4, s);
  a.Y.alloc((size_t)n * 4, s);
  k_feat_to_f32<<<(uint32_t)((nf + 255) / 256), 256, 0, s>>>(fi.as<int32_t>(), nf, FEAT_SCALE, a.X.as<float>());
  GEM_LAUNCH_CHECK("k_feat_to_f32");
  const float *p = m->params.as<float>();
  const uint16_t *wb = m->wbf.as<uint16_t>();
  const int H = (int)m->h, F = (int)m->in;
  gemm<false, false, true, 1>(s, n, H, F, a.X.as<float>(), F, wb, H, p + m->ob1, nullptr, a.H1.as<float>(), H);
  gemm<false, false, true, 1>(s, n, H, H, a.H1.as<float>(), H, wb + (size_t)F * H, H, p + m->ob2, nullptr,
                              a.H2.as<float>(), H);
  gemm<false, false, true, 2>(s, n, 1, H, a.H2.as<float>(), H, wb + (size_t)F * H + (size_t)H * H, 1, p + m->ob3,
                              nullptr, a.Y.as<float>(), 1);
}

float backward(gem_mlp *m, const float *h_target, uint32_t n, Acts &a) {
  hipStream_t s = m->ctx->stream;
  const int H = (int)m->h, F = (int)m->in;
  DevBuf T, dY, dH2, dH1, part;
  T.alloc((size_t)n * 4, s);
  dY.alloc((size_t)n * 4, s);
  dH2.alloc((size_t)n * H * 4, s);
  dH1.alloc((size_t)n * H * 4, s);
  const uint32_t nb = (n + 255) / 256;
  part.alloc(nb * 4, s);
  GEM_HIP_CHECK(hipMemcpyAsync(T.p, h_target, (size_t)n * 4, hipMemcpyHostToDevice, s));
  k_mse_grad<<<nb, 256, 0, s>>>(a.Y.as<float>(), T.as<float>(), (int)n, dY.as<float>(), part.as<float>());
  GEM_LAUNCH_CHECK("k_mse_grad");
  float *g = m->grads.as<float>();
  const uint16_t *wb = m->wbf.as<uint16_t>();
  const uint16_t *W2b = wb + (size_t)F * H, *W3b = wb + (size_t)F * H + (size_t)H * H;
  // layer 3
  gemm<true, false, false, 0>(s, H, 1, n, a.H2.as<float>(), H, dY.as<float>(), 1, nullptr, nullptr, g + m->oW3, 1);
  k_colsum<<<1, 256, 0, s>>>(dY.as<float>(), (int)n, 1, g + m->ob3);
  gemm<false, false, true, 3>(s, n, H, 1, dY.as<float>(), 1, W3b, H, nullptr, a.H2.as<float>(), dH2.as<float>(), H);
  // layer 2
  gemm<true, false, false, 0>(s, H, H, n, a.H1.as<float>(), H, dH2.as<float>(), H, nullptr, nullptr, g + m->oW2, H);
  k_colsum<<<H, 256, 0, s>>>(dH2.as<float>(), (int)n, H, g + m->ob2);
  gemm<false, true, true, 3>(s, n, H, H, dH2.as<float>(), H, W2b, H, nullptr, a.H1.as<float>(), dH1.as<float>(), H);
  // layer 1
  gemm<true, false, false, 0>(s, F, H, n, a.X.as<float>(), F, dH1.as<float>(), H, nullptr, nullptr, g + m->oW1, H);
  k_colsum<<<H, 256, 0, s>>>(dH1.as<float>(), (int)n, H, g + m->ob1);
  GEM_LAUNCH_CHECK("k_colsum");
  std::vector<float> hp(nb);
  GEM_HIP_CHECK(hipMemcpyAsync(hp.data(), part.p, nb * 4, hipMemcpyDeviceToHost, s));
  GEM_HIP_CHECK(hipStreamSynchronize(s));
  double loss = 0;
  for (float x : hp) loss += x;
  return (float)loss;
}

template <class F>
int mguard(gem_mlp *m, F &&f) {
  if (!m) return GEM_E_INVALID;
  if (!m->ctx) return GEM_E_STATE;  // its context was destroyed first
  try {
    GEM_HIP_CHECK(hipSetDevice(m->ctx->device));
    f();
    return GEM_OK;
  } catch (const Error &e) {
    m->ctx->err = e.msg;
    return e.code;
  } catch (const std::exception &e) {
    m->ctx->err = e.what();
    return GEM_E_INVALID;
  }
}

}  // namespace

namespace gemd {
float mlp_margin(const gem_mlp *m) { return m->margin; }

void mlp_detach(gem_mlp *m) {
  m->params.reset();
  m->grads.reset();
  m->adam_m.reset();
  m->adam_v.reset();
  m->wbf.reset();
  m->ctx = nullptr;
}
}  // namespace gemd

extern "C" {

int gem_mlp_create(gem_ctx *ctx, uint32_t in_dim, uint32_t hidden, uint64_t seed, gem_mlp **out) {
  if (!ctx || !out || in_dim == 0 || hidden == 0) return GEM_E_INVALID;
  auto *m = new gem_mlp();
  m->ctx = ctx;
  m->in = in_dim;
  m->h = hidden;
  m->oW1 = 0;
  m->ob1 = m->oW1 + (size_t)in_dim * hidden;
  m->oW2 = m->ob1 + hidden;
  m->ob2 = m->oW2 + (size_t)hidden * hidden;
  m->oW3 = m->ob2 + hidden;
  m->ob3 = m->oW3 + hidden;
  m->P = m->ob3 + 1;
  int st = mguard(m, [&]() {
    hipStream_t s = ctx->stream;
    m->params.alloc(m->P * 4, s);
    m->grads.alloc(m->P * 4, s);
    m->adam_m.alloc(m->P * 4, s);
    m->adam_v.alloc(m->P * 4, s);
    m->wbf.alloc(((size_t)in_dim * hidden + (size_t)hidden * hidden + hidden) * 2, s);
    // He-uniform weights, zero biases (seeded, host side)
    std::vector<float> p(m->P, 0.f);
    gemh::Rng rng(seed);
    auto fill = [&](size_t o, size_t cnt, double fan_in) {
      const double lim = std::sqrt(6.0 / fan_in);
      for (size_t i = 0; i < cnt; ++i) p[o + i] = (float)((2.0 * rng.uniform() - 1.0) * lim);
    };
    fill(m->oW1, (size_t)in_dim * hidden, in_dim);
    fill(m->oW2, (size_t)hidden * hidden, hidden);
    fill(m->oW3, hidden, hidden);
    GEM_HIP_CHECK(hipMemcpyAsync(m->params.p, p.data(), m->P * 4, hipMemcpyHostToDevice, s));
    GEM_HIP_CHECK(hipMemsetAsync(m->adam_m.p, 0, m->P * 4, s));
    GEM_HIP_CHECK(hipMemsetAsync(m->adam_v.p, 0, m->P * 4, s));
    refresh_bf16(m);
    GEM_HIP_CHECK(hipStreamSynchronize(s));
  });
  if (st) {
    delete m;
    return st;
  }
  ctx->mlps.push_back(m);
  *out = m;
  return GEM_OK;
}

void gem_mlp_destroy(gem_mlp *m) {
  if (!m) return;
  if (gem_ctx *c = m->ctx) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->mlps.erase(std::remove(c->mlps.begin(), c->mlps.end(), m), c->mlps.end());
  }
  delete m;
}

uint64_t gem_mlp_num_params(const gem_mlp *m) { return m ? m->P : 0; }

int gem_mlp_set_margin(gem_mlp *m, float margin) {
  if (!m || !(margin >= 0.f)) return GEM_E_INVALID;
  m->margin = margin;
  return GEM_OK;
}

int gem_mlp_get_params(gem_mlp *m, float *out) {
  return mguard(m, [&]() {
    if (!out) throw Error{GEM_E_INVALID, "null output"};
    GEM_HIP_CHECK(hipMemcpyAsync(out, m->params.p, m->P * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    GEM_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
  });
}

int gem_mlp_set_params(gem_mlp *m, const float *in) {
  return mguard(m, [&]() {
    if (!in) throw Error{GEM_E_INVALID, "null input"};
    GEM_HIP_CHECK(hipMemcpyAsync(m->params.p, in, m->P * 4, hipMemcpyHostToDevice, m->ctx->stream));
    refresh_bf16(m);
    GEM_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
  });
}

int gem_mlp_predict(gem_mlp *m, const int32_t *feats, uint32_t n, float *out) {
  return mguard(m, [&]() {
    if (!n) return;
    if (!feats || !out) throw Error{GEM_E_INVALID, "null argument"};
    Acts a;
    forward(m, feats, n, a);
    GEM_HIP_CHECK(hipMemcpyAsync(out, a.Y.p, (size_t)n * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    GEM_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
  });
}

int gem_mlp_loss_grad(gem_mlp *m, const int32_t *feats, const float *targets, uint32_t n, float *loss, float *grad) {
  return mguard(m, [&]() {
    if (!n || !feats || !targets) throw Error{GEM_E_INVALID, "empty or null batch"};
    Acts a;
    forward(m, feats, n, a);
    const float l = backward(m, targets, n, a);
    if (loss) *loss = l;
    if (grad) {
      GEM_HIP_CHECK(hipMemcpyAsync(grad, m->grads.p, m->P * 4, hipMemcpyDeviceToHost, m->ctx->stream));
      GEM_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
    }
  });
}

int gem_mlp_train(gem_mlp *m, const int32_t *feats, const float *targets, uint32_t n, uint32_t iters, float lr,
                  float *losses) {
  return mguard(m, [&]() {
    if (!n || !feats || !targets) throw Error{GEM_E_INVALID, "empty or null batch"};
    hipStream_t s = m->ctx->stream;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t it = 0; it < iters; ++it) {
      Acts a;
      forward(m, feats, n, a);
      const float l = backward(m, targets, n, a);
      if (losses) losses[it] = l;
      ++m->adam_t;
      const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
      k_adam<<<(uint32_t)((m->P + 255) / 256), 256, 0, s>>>(m->params.as<float>(), m->grads.as<float>(),
                                                            m->adam_m.as<float>(), m->adam_v.as<float>(), m->P, lr, b1,
                                                            b2, eps, c1, c2);
      GEM_LAUNCH_CHECK("k_adam");
      refresh_bf16(m);
    }
    GEM_HIP_CHECK(hipStreamSynchronize(s));
  });
}

}  // extern "C"
