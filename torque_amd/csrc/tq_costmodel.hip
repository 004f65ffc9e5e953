// tq_costmodel.hip -- learned cost model for the grouped-GEMM tile
// configuration: a multilayer perceptron trained and run on the GPU.
//
// No reference counterpart (the reference has no tile choice: it calls
// cublas{S,D}gemmStridedBatched per A block, src/gpu/block_sparse.cuh:870-930).
//
// Input: TQ_COST_FEATURES numbers describing one MFMA tile class of a plan
// (flop-weighted histograms of the pairs' M, N and K, totals, and the tile
// shape, tile count and padding efficiency of one configuration). Output:
// the predicted log of that class's kernel time under that configuration.
// Network: x -> relu(W1^T x + b1) -> relu(W2^T h1 + b2) -> w3 . h2 + b3 with
// 64 hidden units per layer, one per lane of a wavefront, so one wavefront
// evaluates one sample with its weights spread across lanes.
//
// Training (tq_costmodel_fit): full-batch Adam on the mean squared error of
// the normalised log-times, with three kernels per step: per-wavefront
// gradients (forward and backward of every sample of the wave, accumulated
// in registers), a fixed-order reduction of the per-wave partial gradients
// (deterministic), and the Adam update.
#include <algorithm>
#include <cmath>
#include <memory>
#include <random>
#include <vector>

#include "tq_common.h"
#include "tq_plan.h"

namespace {

constexpr int kF = TQ_COST_FEATURES;
constexpr int kH = TQ_COST_HIDDEN;
static_assert(kH == 64, "one hidden unit per lane");

// Parameter layout (float): W1 [kF][kH] (input i, unit u at i*kH + u), b1,
// W2 [kH][kH] (unit j of layer 1, unit u of layer 2 at j*kH + u), b2, w3, b3.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kF * kH;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH;
constexpr int kParams = kOffB3 + 1;
constexpr int kMaxWaves = 256;  // gradient partials: one row per wavefront

__device__ inline float lane_value(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Weights shared by the 4 wavefronts of a workgroup. W2 rows are padded to
// kH + 1 so that lane u reading row u (backward) hits distinct banks.
struct MlpSmem {
  float w1[kF * kH];
  float w2[kH * (kH + 1)];
};

__device__ inline void load_weights(MlpSmem &s, const float *__restrict__ prm) {
  for (int e = threadIdx.x; e < kF * kH; e += blockDim.x) s.w1[e] = prm[kOffW1 + e];
  for (int e = threadIdx.x; e < kH * kH; e += blockDim.x)
    s.w2[(e / kH) * (kH + 1) + e % kH] = prm[kOffW2 + e];
  __syncthreads();
}

// Forward pass of sample b on one wavefront (lane = hidden unit).
struct Fwd {
  float z1, h1, z2, h2, out;
};

__device__ inline Fwd forward(const MlpSmem &s, const float *__restrict__ xb, float b1,
                              float b2, float w3, float b3, int lane) {
  Fwd f;
  f.z1 = b1;
#pragma unroll 2
  for (int i = 0; i < kF; ++i) f.z1 += xb[i] * s.w1[i * kH + lane];
  f.h1 = f.z1 > 0.f ? f.z1 : 0.f;
  f.z2 = b2;
#pragma unroll 8
  for (int j = 0; j < kH; ++j) f.z2 += lane_value(f.h1, j) * s.w2[j * (kH + 1) + lane];
  f.h2 = f.z2 > 0.f ? f.z2 : 0.f;
  // lane 0's sum for every lane: the butterfly order differs per lane
  f.out = lane_value(wave_sum(f.h2 * w3), 0) + b3;
  return f;
}

__global__ __launch_bounds__(256) void mlp_grad_kernel(const float *__restrict__ prm,
                                                       const float *__restrict__ x,
                                                       const float *__restrict__ y, int64_t n,
                                                       float *__restrict__ partial,
                                                       float *__restrict__ loss_partial) {
  __shared__ MlpSmem s;
  load_weights(s, prm);
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t n_waves = (int64_t)gridDim.x * 4;
  const float b1 = prm[kOffB1 + lane], b2 = prm[kOffB2 + lane];
  const float w3 = prm[kOffW3 + lane], b3 = prm[kOffB3];
  const float inv_n = 1.0f / (float)n;
  float g_w1[kF], g_w2[kH];
#pragma unroll
  for (int i = 0; i < kF; ++i) g_w1[i] = 0.f;
#pragma unroll
  for (int j = 0; j < kH; ++j) g_w2[j] = 0.f;
  float g_b1 = 0.f, g_b2 = 0.f, g_w3 = 0.f, g_b3 = 0.f, loss = 0.f;
  for (int64_t b = w; b < n; b += n_waves) {
    const float *xb = x + b * kF;
    const Fwd f = forward(s, xb, b1, b2, w3, b3, lane);
    const float r = f.out - y[b];
    loss += r * r * inv_n;
    const float d = 2.f * r * inv_n;  // dL/dout
    g_w3 += d * f.h2;
    g_b3 += d;
    const float dz2 = f.z2 > 0.f ? d * w3 : 0.f;
    g_b2 += dz2;
#pragma unroll
    for (int j = 0; j < kH; ++j) g_w2[j] += lane_value(f.h1, j) * dz2;
    // dL/dh1 for unit `lane`: row `lane` of W2 against every dz2
    float dh1 = 0.f;
#pragma unroll 8
    for (int u = 0; u < kH; ++u) dh1 += s.w2[lane * (kH + 1) + u] * lane_value(dz2, u);
    const float dz1 = f.z1 > 0.f ? dh1 : 0.f;
    g_b1 += dz1;
#pragma unroll
    for (int i = 0; i < kF; ++i) g_w1[i] += xb[i] * dz1;
  }
  float *g = partial + w * kParams;
#pragma unroll
  for (int i = 0; i < kF; ++i) g[kOffW1 + i * kH + lane] = g_w1[i];
  g[kOffB1 + lane] = g_b1;
#pragma unroll
  for (int j = 0; j < kH; ++j) g[kOffW2 + j * kH + lane] = g_w2[j];
  g[kOffB2 + lane] = g_b2;
  g[kOffW3 + lane] = g_w3;
  if (lane == 0) {
    g[kOffB3] = g_b3;
    loss_partial[w] = loss;
  }
}

// grad[p] = sum over wavefronts of partial[w][p], in wavefront order; the
// extra thread p = kParams sums the loss.
__global__ void mlp_reduce_kernel(const float *__restrict__ partial,
                                  const float *__restrict__ loss_partial, int n_waves,
                                  float *__restrict__ grad, float *__restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kParams) {
    float acc = 0.f;
    for (int w = 0; w < n_waves; ++w) acc += partial[(int64_t)w * kParams + p];
    grad[p] = acc;
  } else if (p == kParams) {
    float acc = 0.f;
    for (int w = 0; w < n_waves; ++w) acc += loss_partial[w];
    *loss = acc;
  }
}

__global__ void mlp_adam_kernel(float *__restrict__ prm, const float *__restrict__ grad,
                                float *__restrict__ m, float *__restrict__ v, float lr,
                                float beta1, float beta2, float eps, float corr1,
                                float corr2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  const float g = grad[p];
  const float mp = beta1 * m[p] + (1.f - beta1) * g;
  const float vp = beta2 * v[p] + (1.f - beta2) * g * g;
  m[p] = mp;
  v[p] = vp;
  prm[p] -= lr * (mp / corr1) / (sqrtf(vp / corr2) + eps);
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(const float *__restrict__ prm,
                                                          const float *__restrict__ x,
                                                          int64_t n, float *__restrict__ out) {
  __shared__ MlpSmem s;
  load_weights(s, prm);
  const int lane = threadIdx.x & 63;
  const int64_t n_waves = (int64_t)gridDim.x * 4;
  const float b1 = prm[kOffB1 + lane], b2 = prm[kOffB2 + lane];
  const float w3 = prm[kOffW3 + lane], b3 = prm[kOffB3];
  for (int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); b < n; b += n_waves) {
    const Fwd f = forward(s, x + b * kF, b1, b2, w3, b3, lane);
    if (lane == 0) out[b] = f.out;
  }
}

}  // namespace

struct tq_costmodel_s {
  tq_handle_t h = nullptr;
  DevBuf params, m, v, grad, partial, loss_partial, loss, x, y, out;
  // input and target normalisation (set by fit; identity until then)
  std::vector<float> x_mean = std::vector<float>(kF, 0.f);
  std::vector<float> x_std = std::vector<float>(kF, 1.f);
  float y_mean = 0.f, y_std = 1.f;
  int64_t steps_done = 0;
};

namespace {

tq_status upload_f(tq_handle_t h, DevBuf &dst, const float *src, int64_t n) {
  TQ_HIP(h, dst.alloc(sizeof(float) * std::max<int64_t>(n, 1)));
  if (n > 0)
    TQ_HIP(h, hipMemcpyAsync(dst.p, src, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
  return TQ_SUCCESS;
}

int grad_grid(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, (int64_t)4), kMaxWaves / 4));
}

// Gradient of the mean squared error over (x, y) already on the device
// (model->x, model->y) into model->grad and model->loss.
tq_status run_gradient(tq_costmodel_s *md, int64_t n) {
  tq_handle_t h = md->h;
  const int grid = grad_grid(n);
  TQ_HIP(h, md->partial.alloc(sizeof(float) * (size_t)grid * 4 * kParams));
  TQ_HIP(h, md->loss_partial.alloc(sizeof(float) * grid * 4));
  TQ_HIP(h, md->loss.alloc(sizeof(float)));
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(grid), dim3(256), 0, h->stream,
                     md->params.as<float>(), md->x.as<float>(), md->y.as<float>(), n,
                     md->partial.as<float>(), md->loss_partial.as<float>());
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3(ceil_div(kParams + 1, 256)), dim3(256), 0,
                     h->stream, md->partial.as<float>(), md->loss_partial.as<float>(),
                     grid * 4, md->grad.as<float>(), md->loss.as<float>());
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

}  // namespace

// ---- features ------------------------------------------------------------

// Histogram bucket of a dimension: <16, 16-31, ..., 512-1023, >=1024.
static int dim_bucket(int64_t d) {
  int b = 0;
  for (int64_t t = 16; b < 7 && d >= t; t <<= 1) ++b;
  return b;
}

int tq_cost_features_host(const int64_t *m, const int64_t *n, const int64_t *k,
                          int64_t n_pairs, const GemmConfig &g, int n_cu, float *feat) {
  for (int i = 0; i < kF; ++i) feat[i] = 0.f;
  double flops = 0, padded = 0, tiles = 0, out_bytes = 0;
  double hist[24] = {0};
  for (int64_t p = 0; p < n_pairs; ++p) {
    const double f = 2.0 * m[p] * (double)n[p] * k[p];
    flops += f;
    hist[dim_bucket(m[p])] += f;
    hist[8 + dim_bucket(n[p])] += f;
    hist[16 + dim_bucket(k[p])] += f;
    const double tm = std::ceil((double)m[p] / g.bm), tn = std::ceil((double)n[p] / g.bn);
    tiles += tm * tn;
    padded += 2.0 * tm * g.bm * tn * g.bn * std::ceil((double)k[p] / g.bk) * g.bk;
    out_bytes += 8.0 * m[p] * (double)n[p];
  }
  if (n_pairs == 0 || flops <= 0) return 0;
  for (int i = 0; i < 24; ++i) feat[i] = (float)(hist[i] / flops);
  const double concurrency = (double)n_cu * (g.wave_tile ? 16 : 4);
  feat[24] = (float)std::log10(flops);
  feat[25] = (float)std::log10((double)n_pairs);
  feat[26] = (float)std::log10(tiles);
  feat[27] = (float)(flops / padded);
  feat[28] = (float)std::log2((double)g.bm);
  feat[29] = (float)std::log2((double)g.bn);
  feat[30] = (float)std::log2((double)g.bk);
  feat[31] = (float)(2 * g.wave_tile + g.min_waves);
  feat[32] = (float)std::log10(tiles / concurrency);
  feat[33] = (float)std::log10(out_bytes);
  return 1;
}

extern "C" {

int64_t tq_costmodel_n_params(void) { return kParams; }

tq_status tq_cost_features(const int64_t *m, const int64_t *n, const int64_t *k,
                           int64_t n_pairs, tq_precision precision, int32_t config,
                           int32_t n_cu, float *features) {
  if (!features || n_pairs < 0 || (n_pairs > 0 && (!m || !n || !k)))
    return TQ_ERROR_INVALID_VALUE;
  if (config < 0 || config >= tq_gemm_n_configs(precision)) return TQ_ERROR_INVALID_VALUE;
  tq_cost_features_host(m, n, k, n_pairs, tq_gemm_config(precision, config),
                        n_cu > 0 ? n_cu : 256, features);
  return TQ_SUCCESS;
}

tq_status tq_plan_features(tq_plan_t pl, int32_t cls, int32_t config, float *features) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (cls < 0 || cls > 1 || !features || config < 0 ||
      config >= tq_gemm_n_configs(pl->prec))
    return TQ_ERROR_INVALID_VALUE;
  std::vector<int64_t> m, n, k;
  for (int64_t p = 0; p < pl->n_pairs; ++p)
    if (pair_class(*pl, pl->h_m[p], pl->h_n[p], pl->h_k[p]) == cls) {
      m.push_back(pl->h_m[p]);
      n.push_back(pl->h_n[p]);
      k.push_back(pl->h_k[p]);
    }
  tq_cost_features_host(m.data(), n.data(), k.data(), (int64_t)m.size(),
                        tq_gemm_config(pl->prec, config), pl->h ? pl->h->n_cu : 256, features);
  return TQ_SUCCESS;
}

tq_status tq_costmodel_create(tq_handle_t h, uint64_t seed, tq_costmodel_t *model) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (!model) return TQ_ERROR_INVALID_VALUE;
  *model = nullptr;
  std::unique_ptr<tq_costmodel_s> md(new tq_costmodel_s());
  md->h = h;
  // He-normal weights, zero biases
  std::vector<float> prm(kParams, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> n1(0.f, std::sqrt(2.f / kF)), n2(0.f, std::sqrt(2.f / kH));
  for (int e = 0; e < kF * kH; ++e) prm[kOffW1 + e] = n1(rng);
  for (int e = 0; e < kH * kH; ++e) prm[kOffW2 + e] = n2(rng);
  for (int e = 0; e < kH; ++e) prm[kOffW3 + e] = n2(rng);
  std::vector<float> zeros(kParams, 0.f);
  TQ_CHECK(upload_f(h, md->params, prm.data(), kParams));
  TQ_CHECK(upload_f(h, md->m, zeros.data(), kParams));
  TQ_CHECK(upload_f(h, md->v, zeros.data(), kParams));
  TQ_HIP(h, md->grad.alloc(sizeof(float) * kParams));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  *model = md.release();
  return TQ_SUCCESS;
}

tq_status tq_costmodel_destroy(tq_costmodel_t md) {
  delete md;
  return TQ_SUCCESS;
}

tq_status tq_costmodel_get_params(tq_costmodel_t md, float *params) {
  if (!md) return TQ_ERROR_NOT_INITIALIZED;
  if (!params) return TQ_ERROR_INVALID_VALUE;
  TQ_HIP(md->h, hipMemcpyAsync(params, md->params.p, sizeof(float) * kParams,
                               hipMemcpyDeviceToHost, md->h->stream));
  TQ_HIP(md->h, hipStreamSynchronize(md->h->stream));
  return TQ_SUCCESS;
}

tq_status tq_costmodel_set_params(tq_costmodel_t md, const float *params) {
  if (!md) return TQ_ERROR_NOT_INITIALIZED;
  if (!params) return TQ_ERROR_INVALID_VALUE;
  TQ_CHECK(upload_f(md->h, md->params, params, kParams));
  TQ_HIP(md->h, hipStreamSynchronize(md->h->stream));
  return TQ_SUCCESS;
}

tq_status tq_costmodel_get_scaling(tq_costmodel_t md, float *x_mean, float *x_std,
                                   float *y) {
  if (!md) return TQ_ERROR_NOT_INITIALIZED;
  if (!x_mean || !x_std || !y) return TQ_ERROR_INVALID_VALUE;
  std::copy(md->x_mean.begin(), md->x_mean.end(), x_mean);
  std::copy(md->x_std.begin(), md->x_std.end(), x_std);
  y[0] = md->y_mean;
  y[1] = md->y_std;
  return TQ_SUCCESS;
}

tq_status tq_costmodel_set_scaling(tq_costmodel_t md, const float *x_mean, const float *x_std,
                                   const float *y) {
  if (!md) return TQ_ERROR_NOT_INITIALIZED;
  if (!x_mean || !x_std || !y) return TQ_ERROR_INVALID_VALUE;
  for (int i = 0; i < kF; ++i)
    if (!(x_std[i] > 0.f)) return TQ_ERROR_INVALID_VALUE;
  if (!(y[1] > 0.f)) return TQ_ERROR_INVALID_VALUE;
  md->x_mean.assign(x_mean, x_mean + kF);
  md->x_std.assign(x_std, x_std + kF);
  md->y_mean = y[0];
  md->y_std = y[1];
  return TQ_SUCCESS;
}

tq_status tq_costmodel_gradient(tq_costmodel_t md, const float *x, const float *y, int64_t n,
                                float *grad, float *loss) {
  if (!md) return TQ_ERROR_NOT_INITIALIZED;
  if (!x || !y || n < 1) return TQ_ERROR_INVALID_VALUE;
  tq_handle_t h = md->h;
  TQ_CHECK(upload_f(h, md->x, x, n * kF));
  TQ_CHECK(upload_f(h, md->y, y, n));
  TQ_CHECK(run_gradient(md, n));
  if (grad)
    TQ_HIP(h, hipMemcpyAsync(grad, md->grad.p, sizeof(float) * kParams, hipMemcpyDeviceToHost,
                             h->stream));
  float l = 0.f;
  TQ_HIP(h, hipMemcpyAsync(&l, md->loss.p, sizeof(float), hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  if (loss) *loss = l;
  return TQ_SUCCESS;
}

tq_status tq_costmodel_fit(tq_costmodel_t md, const float *x, const float *y, int64_t n,
                           int32_t steps, float lr, float *loss) {
  if (!md) return TQ_ERROR_NOT_INITIALIZED;
  if (!x || !y || n < 2 || steps < 0 || !(lr > 0.f)) return TQ_ERROR_INVALID_VALUE;
  tq_handle_t h = md->h;
  // normalisation from this data set
  std::vector<double> mean(kF, 0.0), var(kF, 0.0);
  double ym = 0.0, yv = 0.0;
  for (int64_t b = 0; b < n; ++b) {
    for (int i = 0; i < kF; ++i) mean[i] += x[b * kF + i];
    ym += y[b];
  }
  for (int i = 0; i < kF; ++i) mean[i] /= n;
  ym /= n;
  for (int64_t b = 0; b < n; ++b) {
    for (int i = 0; i < kF; ++i) {
      const double d = x[b * kF + i] - mean[i];
      var[i] += d * d;
    }
    yv += (y[b] - ym) * (y[b] - ym);
  }
  for (int i = 0; i < kF; ++i) {
    md->x_mean[i] = (float)mean[i];
    const double sd = std::sqrt(var[i] / n);
    md->x_std[i] = sd > 1e-6 ? (float)sd : 1.f;
  }
  md->y_mean = (float)ym;
  const double ysd = std::sqrt(yv / n);
  md->y_std = ysd > 1e-6 ? (float)ysd : 1.f;
  std::vector<float> xn((size_t)n * kF), yn(n);
  for (int64_t b = 0; b < n; ++b) {
    for (int i = 0; i < kF; ++i)
      xn[b * kF + i] = (x[b * kF + i] - md->x_mean[i]) / md->x_std[i];
    yn[b] = (y[b] - md->y_mean) / md->y_std;
  }
  TQ_CHECK(upload_f(h, md->x, xn.data(), n * kF));
  TQ_CHECK(upload_f(h, md->y, yn.data(), n));
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  for (int32_t s = 0; s < steps; ++s) {
    TQ_CHECK(run_gradient(md, n));
    const int64_t t = ++md->steps_done;
    const float c1 = 1.f - std::pow(beta1, (float)t), c2 = 1.f - std::pow(beta2, (float)t);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(ceil_div(kParams, 256)), dim3(256), 0,
                       h->stream, md->params.as<float>(), md->grad.as<float>(),
                       md->m.as<float>(), md->v.as<float>(), lr, beta1, beta2, eps, c1, c2);
    TQ_HIP(h, hipGetLastError());
  }
  // loss at the final parameters
  TQ_CHECK(run_gradient(md, n));
  float l = 0.f;
  TQ_HIP(h, hipMemcpyAsync(&l, md->loss.p, sizeof(float), hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  if (loss) *loss = l;
  return TQ_SUCCESS;
}

// raw = 1: x is fed to the network as given and the raw output returned
// (for gradient checks); raw = 0: the fit's normalisation is applied.
tq_status tq_costmodel_predict(tq_costmodel_t md, const float *x, int64_t n, int32_t raw,
                               float *y) {
  if (!md) return TQ_ERROR_NOT_INITIALIZED;
  if (!x || !y || n < 0) return TQ_ERROR_INVALID_VALUE;
  if (n == 0) return TQ_SUCCESS;
  tq_handle_t h = md->h;
  std::vector<float> xn((size_t)n * kF);
  for (int64_t b = 0; b < n; ++b)
    for (int i = 0; i < kF; ++i)
      xn[b * kF + i] = raw ? x[b * kF + i] : (x[b * kF + i] - md->x_mean[i]) / md->x_std[i];
  DevBuf dx, dy;
  TQ_CHECK(upload_f(h, dx, xn.data(), n * kF));
  TQ_HIP(h, dy.alloc(sizeof(float) * n));
  const int grid = (int)std::min<int64_t>(ceil_div(n, (int64_t)4), 1024);
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(grid), dim3(256), 0, h->stream,
                     md->params.as<float>(), dx.as<float>(), n, dy.as<float>());
  TQ_HIP(h, hipGetLastError());
  TQ_HIP(h, hipMemcpyAsync(y, dy.p, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  if (!raw)
    for (int64_t b = 0; b < n; ++b) y[b] = y[b] * md->y_std + md->y_mean;
  return TQ_SUCCESS;
}

tq_status tq_plan_choose_config(tq_plan_t pl, tq_costmodel_t md, int32_t *best) {
  if (!pl || !md) return TQ_ERROR_NOT_INITIALIZED;
  const int nc = tq_gemm_n_configs(pl->prec);
  const int n_cu = pl->h ? pl->h->n_cu : 256;
  for (int c = 0; c < 2; ++c) {
    std::vector<int64_t> m, n, k;
    for (int64_t p = 0; p < pl->n_pairs; ++p)
      if (pair_class(*pl, pl->h_m[p], pl->h_n[p], pl->h_k[p]) == c) {
        m.push_back(pl->h_m[p]);
        n.push_back(pl->h_n[p]);
        k.push_back(pl->h_k[p]);
      }
    int arg = -1, arg_split = 1;
    double arg_t = 0;
    // split-K enters through the shapes it produces: a split class is the
    // grouped GEMM of its (pair, slice) GEMMs, predicted by the same model,
    // plus the reduction's traffic (its slices read, the output written).
    // Splitting adds parallelism, so it is weighed only for a class whose
    // whole-pair tiles under the best unsplit configuration leave the GPU
    // short of two rounds of workgroups (the model, trained on whole
    // pairs, misjudges slices of classes that already fill it), and a split
    // is taken only when it is predicted at least 20% faster than the best
    // whole-pair configuration (a margin for that extrapolation).
    double whole_t = 0;
    for (int sp : {1, 2, 4, 8, 16}) {
      if (sp > 1 && arg >= 0) {
        const GemmConfig g = tq_gemm_config(pl->prec, arg);
        double tiles = 0;
        for (size_t p = 0; p < m.size(); ++p)
          tiles += (double)ceil_div(m[p], (int64_t)g.bm) * (double)ceil_div(n[p], (int64_t)g.bn);
        if (tiles >= 2.0 * n_cu) break;
      }
      std::vector<int64_t> vm, vn, vk;
      double out_elems = 0, slices = 0;
      for (size_t p = 0; p < m.size(); ++p) {
        const int64_t ks = split_slice_k(k[p], sp);
        const int64_t ns = ceil_div(k[p], ks);
        for (int64_t q = 0; q < ns; ++q) {
          vm.push_back(m[p]);
          vn.push_back(n[p]);
          vk.push_back(std::min<int64_t>(ks, k[p] - q * ks));
        }
        out_elems += (double)m[p] * n[p];
        slices += (double)ns * m[p] * n[p];
      }
      if (sp > 1 && vm.size() == m.size()) continue;  // no pair would split
      std::vector<float> feat((size_t)nc * kF), pred(nc);
      for (int q = 0; q < nc; ++q)
        tq_cost_features_host(vm.data(), vn.data(), vk.data(), (int64_t)vm.size(),
                              tq_gemm_config(pl->prec, q), n_cu, &feat[(size_t)q * kF]);
      TQ_CHECK(tq_costmodel_predict(md, feat.data(), nc, 0, pred.data()));
      const double reduce_ms = sp > 1 ? 8.0 * (slices + out_elems) / 5e9 : 0.0;
      for (int q = 0; q < nc; ++q) {
        if (!plan_config_ok(*pl, q)) continue;
        const double t = std::exp((double)pred[q]) + reduce_ms;
        if (sp > 1 && t > 0.8 * whole_t) continue;
        if (arg < 0 || t < arg_t) {
          arg = q;
          arg_t = t;
          arg_split = sp;
        }
      }
      if (sp == 1) whole_t = arg_t;
    }
    pl->drop_graph();
    pl->cls[c].cfg = arg < 0 ? -1 : arg;
    pl->cls[c].split = arg_split;
    pl->cls[c].split_user = true;
    if (best) best[c] = pl->cls[c].current();
  }
  return TQ_SUCCESS;
}

}  // extern "C"
