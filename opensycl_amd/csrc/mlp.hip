// Cost model network: a 16 -> 64 -> 64 -> 1 multilayer perceptron with tanh
// hidden units, trained and evaluated by the kernels below.  It predicts the
// (standardised) log run time of a launch configuration on a problem shape
// and is used to pick launch configurations for shapes the exhaustive tuner
// has not seen.  No counterpart in the reference, whose launch geometry is
// fixed (numeric.hpp:144-156, decoupled_lookback_scan.hpp:481-488).
//
// Parameter layout (one float array of ACPP_MLP_PARAMS entries):
//   W1[64][16], b1[64], W2[64][64], b2[64], w3[64], b3
// Loss: mean squared error over the batch, L = (1/B) sum_b (y_b - t_b)^2.
//
// Kernels:
//   fwd_bwd  -- one work-item per sample: forward pass with all weights in
//               LDS (broadcast reads), then the backward pass to the hidden
//               deltas; activations and deltas are stored feature-major
//               ([unit][batch]) so that the stores coalesce.
//   grad     -- one work-item per (parameter, batch chunk): the weight
//               gradient as a fixed-order sum over the chunk's samples.
//   combine  -- sums the chunk partials in chunk order (deterministic: same
//               inputs give the same bits), optionally applies Adam.
#include <math.h>

#include <algorithm>
#include <cmath>

#include "internal.hpp"
#include "launch_util.hpp"

namespace acpp_impl {
namespace mlp {

constexpr int IN = ACPP_MLP_INPUTS;
constexpr int H = ACPP_MLP_HIDDEN;
constexpr int P = ACPP_MLP_PARAMS;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + H * IN;
constexpr int OFF_W2 = OFF_B1 + H;
constexpr int OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H;
constexpr int OFF_B3 = OFF_W3 + H;
static_assert(OFF_B3 + 1 == P, "parameter layout");

constexpr int CHUNK = 256;  // samples per gradient partial
constexpr uint32_t max_batch = 1u << 20;

__device__ __forceinline__ void load_params(const float *params, float *s) {
  for (int i = threadIdx.x; i < P; i += blockDim.x) s[i] = params[i];
  __syncthreads();
}

// Forward only: y[b].
__global__ __launch_bounds__(256) void forward_kernel(const float *params, const float *x, uint32_t batch,
                                                      float *y) {
  __shared__ float s[P];
  load_params(params, s);
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  float in[IN];
#pragma unroll
  for (int i = 0; i < IN; ++i) in[i] = x[static_cast<uint64_t>(b) * IN + i];
  float h1[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float a = s[OFF_B1 + j];
#pragma unroll
    for (int i = 0; i < IN; ++i) a += s[OFF_W1 + j * IN + i] * in[i];
    h1[j] = tanhf(a);
  }
  float out = s[OFF_B3];
  for (int k = 0; k < H; ++k) {
    float a = s[OFF_B2 + k];
#pragma unroll
    for (int j = 0; j < H; ++j) a += s[OFF_W2 + k * H + j] * h1[j];
    out += s[OFF_W3 + k] * tanhf(a);
  }
  y[b] = out;
}

// Forward + backward to the hidden deltas.
//   h1, h2, d1, d2: [H][batch]; dy, err: [batch]
__global__ __launch_bounds__(256) void fwd_bwd_kernel(const float *params, const float *x, const float *target,
                                                      uint32_t batch, float inv_batch, float *h1o, float *h2o,
                                                      float *d1o, float *d2o, float *dyo, float *erro) {
  __shared__ float s[P];
  load_params(params, s);
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  float in[IN];
#pragma unroll
  for (int i = 0; i < IN; ++i) in[i] = x[static_cast<uint64_t>(b) * IN + i];
  float h1[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float a = s[OFF_B1 + j];
#pragma unroll
    for (int i = 0; i < IN; ++i) a += s[OFF_W1 + j * IN + i] * in[i];
    h1[j] = tanhf(a);
    h1o[static_cast<uint64_t>(j) * batch + b] = h1[j];
  }
  float h2[H];
  float out = s[OFF_B3];
#pragma unroll
  for (int k = 0; k < H; ++k) {
    float a = s[OFF_B2 + k];
#pragma unroll
    for (int j = 0; j < H; ++j) a += s[OFF_W2 + k * H + j] * h1[j];
    h2[k] = tanhf(a);
    h2o[static_cast<uint64_t>(k) * batch + b] = h2[k];
    out += s[OFF_W3 + k] * h2[k];
  }
  const float e = out - target[b];
  erro[b] = e * e;
  const float dy = 2.0f * e * inv_batch;
  dyo[b] = dy;
  // d2[k] = dy * w3[k] * (1 - h2[k]^2), kept in h2[] in place.
#pragma unroll
  for (int k = 0; k < H; ++k) {
    h2[k] = dy * s[OFF_W3 + k] * (1.0f - h2[k] * h2[k]);
    d2o[static_cast<uint64_t>(k) * batch + b] = h2[k];
  }
  // d1[j] = (sum_k W2[k][j] d2[k]) * (1 - h1[j]^2)
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float a = 0.0f;
#pragma unroll
    for (int k = 0; k < H; ++k) a += s[OFF_W2 + k * H + j] * h2[k];
    d1o[static_cast<uint64_t>(j) * batch + b] = a * (1.0f - h1[j] * h1[j]);
  }
}

// partial[c][p] = gradient of parameter p over samples of chunk c.
__global__ __launch_bounds__(256) void grad_kernel(const float *x, const float *h1, const float *h2,
                                                   const float *d1, const float *d2, const float *dy,
                                                   uint32_t batch, float *partial) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const uint32_t c = blockIdx.y;
  const uint32_t lo = c * CHUNK;
  const uint32_t hi = min(lo + CHUNK, batch);
  float acc = 0.0f;
  if (p < OFF_B1) {
    const int j = p / IN, i = p % IN;
    const float *dj = d1 + static_cast<uint64_t>(j) * batch;
    for (uint32_t b = lo; b < hi; ++b) acc += dj[b] * x[static_cast<uint64_t>(b) * IN + i];
  } else if (p < OFF_W2) {
    const float *dj = d1 + static_cast<uint64_t>(p - OFF_B1) * batch;
    for (uint32_t b = lo; b < hi; ++b) acc += dj[b];
  } else if (p < OFF_B2) {
    const int k = (p - OFF_W2) / H, j = (p - OFF_W2) % H;
    const float *dk = d2 + static_cast<uint64_t>(k) * batch;
    const float *hj = h1 + static_cast<uint64_t>(j) * batch;
    for (uint32_t b = lo; b < hi; ++b) acc += dk[b] * hj[b];
  } else if (p < OFF_W3) {
    const float *dk = d2 + static_cast<uint64_t>(p - OFF_B2) * batch;
    for (uint32_t b = lo; b < hi; ++b) acc += dk[b];
  } else if (p < OFF_B3) {
    const float *hk = h2 + static_cast<uint64_t>(p - OFF_W3) * batch;
    for (uint32_t b = lo; b < hi; ++b) acc += dy[b] * hk[b];
  } else {
    for (uint32_t b = lo; b < hi; ++b) acc += dy[b];
  }
  partial[static_cast<uint64_t>(c) * P + p] = acc;
}

struct adam_args {
  float lr, beta1, beta2, eps, weight_decay;
  float corr1, corr2;  // 1 - beta^step
  int update;
};

__global__ __launch_bounds__(256) void combine_kernel(const float *partial, uint32_t chunks, float *grad_out,
                                                      float *params, float *m, float *v, adam_args a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  float g = 0.0f;
  for (uint32_t c = 0; c < chunks; ++c) g += partial[static_cast<uint64_t>(c) * P + p];
  if (grad_out) grad_out[p] = g;
  if (!a.update) return;
  const float w = params[p];
  g += a.weight_decay * w;
  const float mm = a.beta1 * m[p] + (1.0f - a.beta1) * g;
  const float vv = a.beta2 * v[p] + (1.0f - a.beta2) * g * g;
  m[p] = mm;
  v[p] = vv;
  params[p] = w - a.lr * (mm / a.corr1) / (sqrtf(vv / a.corr2) + a.eps);
}

// Fixed-order sum of `n` floats by one work-group (256 work-items).
__global__ __launch_bounds__(256) void sum_kernel(const float *in, uint32_t n, float scale, float *out) {
  __shared__ float s[256];
  float acc = 0.0f;
  for (uint32_t i = threadIdx.x; i < n; i += 256) acc += in[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0] * scale;
}

acpp_status check_batch(const char *name, uint32_t batch) {
  if (batch == 0 || batch > max_batch)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT,
                     std::string(name) + ": batch must be in [1, " + std::to_string(max_batch) + "]");
  return ACPP_SUCCESS;
}

// Gradient (and optional Adam update) through the kernels above.
acpp_status train(acpp_allocation_group *g, hipStream_t stream, float *params, float *m, float *v,
                  uint32_t step, const float *x, const float *target, uint32_t batch, float lr,
                  float weight_decay, float *grad_out, float *loss, bool update, const char *name) {
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": scratch group is NULL");
  if (!params || !x || !target || (update && (!m || !v)))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": NULL pointer argument");
  acpp_status st = check_batch(name, batch);
  if (st != ACPP_SUCCESS) return st;
  if (update && step == 0)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": step counts from 1");
  int cus = 0;
  st = enter_device(g, &cus);
  if (st != ACPP_SUCCESS) return st;
  const uint32_t chunks = (batch + CHUNK - 1) / CHUNK;
  const size_t act = static_cast<size_t>(H) * batch;
  const size_t floats = 4 * act + 2 * static_cast<size_t>(batch) + static_cast<size_t>(chunks) * P + 1;
  void *mem = nullptr;
  st = group_obtain(g, floats * sizeof(float), &mem);
  if (st != ACPP_SUCCESS) return st;
  float *h1 = static_cast<float *>(mem);
  float *h2 = h1 + act;
  float *d1 = h2 + act;
  float *d2 = d1 + act;
  float *dy = d2 + act;
  float *err = dy + batch;
  float *partial = err + batch;
  const uint32_t grid = (batch + 255) / 256;
  hipLaunchKernelGGL(fwd_bwd_kernel, dim3(grid), dim3(256), 0, stream, params, x, target, batch,
                     1.0f / static_cast<float>(batch), h1, h2, d1, d2, dy, err);
  hipLaunchKernelGGL(grad_kernel, dim3((P + 255) / 256, chunks), dim3(256), 0, stream, x, h1, h2, d1, d2,
                     dy, batch, partial);
  if (loss)
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, stream, err, batch,
                       1.0f / static_cast<float>(batch), loss);
  adam_args a{};
  a.update = update ? 1 : 0;
  if (update) {
    a.lr = lr;
    a.beta1 = 0.9f;
    a.beta2 = 0.999f;
    a.eps = 1e-8f;
    a.weight_decay = weight_decay;
    a.corr1 = 1.0f - powf(a.beta1, static_cast<float>(step));
    a.corr2 = 1.0f - powf(a.beta2, static_cast<float>(step));
  }
  hipLaunchKernelGGL(combine_kernel, dim3((P + 255) / 256), dim3(256), 0, stream, partial, chunks, grad_out,
                     params, m, v, a);
  st = check_launch(name);
  return st == ACPP_SUCCESS ? clear_error() : st;
}

}  // namespace mlp
}  // namespace acpp_impl

// A trained cost model held on one device, with its input / target
// standardisation, for configuration queries.
struct acpp_cost_model {
  int device = 0;
  float *params = nullptr;  // device, ACPP_MLP_PARAMS
  float *x = nullptr;       // device, candidates x 16
  float *y = nullptr;       // device, candidates
  double x_mean[ACPP_MLP_INPUTS];
  double x_std[ACPP_MLP_INPUTS];
  double y_mean = 0, y_std = 1;
};

namespace acpp_impl {
namespace mlp {

// Candidate configurations of the segmented forms, in the order the model
// was trained on: work-group size x rows.
constexpr uint32_t seg_wgs[] = {64, 128, 256, 512, 1024};
constexpr uint32_t seg_rows[] = {2, 4, 8};
constexpr int seg_candidates = 15;

// The 16 features of (shape, configuration): log2(n+1), log2(S+1), mean,
// deviation, maximum and skew (clipped to +-5) of log2(length+1), fraction
// of empty segments, log2(element bytes), is-float, is-scan, is-plus;
// log2(work-group size), log2(rows), log2(tile elements),
// log2(tiles + 1), log2(S / (tiles + 1) + 1).
void segmented_features(const uint64_t *offsets, uint64_t S, int elem_bytes, bool is_float, bool is_scan,
                        bool is_plus, double shape[11]) {
  const uint64_t n = offsets[S] - offsets[0];
  double sum = 0, mx = 0, empty = 0;
  for (uint64_t s = 0; s < S; ++s) {
    const double len = static_cast<double>(offsets[s + 1] - offsets[s]);
    const double l = std::log2(len + 1.0);
    sum += l;
    mx = std::max(mx, l);
    empty += (len == 0) ? 1.0 : 0.0;
  }
  const double mu = S ? sum / S : 0.0;
  double m2 = 0, m3 = 0;
  for (uint64_t s = 0; s < S; ++s) {
    const double d = std::log2(static_cast<double>(offsets[s + 1] - offsets[s]) + 1.0) - mu;
    m2 += d * d;
    m3 += d * d * d;
  }
  const double sd = S ? std::sqrt(m2 / S) : 0.0;
  double skew = (sd > 1e-9) ? (m3 / S) / (sd * sd * sd) : 0.0;
  skew = std::min(5.0, std::max(-5.0, skew));
  shape[0] = std::log2(static_cast<double>(n) + 1.0);
  shape[1] = std::log2(static_cast<double>(S) + 1.0);
  shape[2] = mu;
  shape[3] = sd;
  shape[4] = mx;
  shape[5] = skew;
  shape[6] = S ? empty / S : 0.0;
  shape[7] = std::log2(static_cast<double>(elem_bytes));
  shape[8] = is_float ? 1.0 : 0.0;
  shape[9] = is_scan ? 1.0 : 0.0;
  shape[10] = is_plus ? 1.0 : 0.0;
}

}  // namespace mlp
}  // namespace acpp_impl

using namespace acpp_impl;

extern "C" {

acpp_status acpp_mlp_forward(acpp_stream_t stream, const float *params, const float *x, uint32_t batch,
                             float *y) {
  const char *name = "acpp_mlp_forward";
  if (!params || !x || !y) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": NULL pointer argument");
  acpp_status st = mlp::check_batch(name, batch);
  if (st != ACPP_SUCCESS) return st;
  hipLaunchKernelGGL(mlp::forward_kernel, dim3((batch + 255) / 256), dim3(256), 0,
                     static_cast<hipStream_t>(stream), params, x, batch, y);
  st = check_launch(name);
  return st == ACPP_SUCCESS ? clear_error() : st;
}

acpp_status acpp_mlp_gradient(acpp_allocation_group *scratch, acpp_stream_t stream, const float *params,
                              const float *x, const float *target, uint32_t batch, float *grad, float *loss) {
  if (!grad) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "acpp_mlp_gradient: grad is NULL");
  return mlp::train(scratch, static_cast<hipStream_t>(stream), const_cast<float *>(params), nullptr, nullptr, 0,
                    x, target, batch, 0.0f, 0.0f, grad, loss, false, "acpp_mlp_gradient");
}

acpp_status acpp_cost_model_create(int device, const float *params, const double *x_mean, const double *x_std,
                                   double y_mean, double y_std, acpp_cost_model **out) {
  if (!params || !x_mean || !x_std || !out)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, "acpp_cost_model_create: NULL pointer argument");
  int prev = 0;
  ACPP_CHECK_HIP(hipGetDevice(&prev));
  if (prev != device) ACPP_CHECK_HIP(hipSetDevice(device));
  auto *m = new acpp_cost_model;
  m->device = device;
  for (int i = 0; i < ACPP_MLP_INPUTS; ++i) {
    m->x_mean[i] = x_mean[i];
    m->x_std[i] = x_std[i] != 0.0 ? x_std[i] : 1.0;
  }
  m->y_mean = y_mean;
  m->y_std = y_std;
  hipError_t e = hipMalloc(&m->params, ACPP_MLP_PARAMS * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->x, mlp::seg_candidates * ACPP_MLP_INPUTS * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->y, mlp::seg_candidates * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(m->params, params, ACPP_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice);
  if (prev != device) (void)hipSetDevice(prev);
  if (e != hipSuccess) {
    (void)hipFree(m->params);
    (void)hipFree(m->x);
    (void)hipFree(m->y);
    delete m;
    return hip_error(e, "acpp_cost_model_create");
  }
  *out = m;
  return clear_error();
}

acpp_status acpp_cost_model_destroy(acpp_cost_model *m) {
  if (!m) return clear_error();
  (void)hipFree(m->params);
  (void)hipFree(m->x);
  (void)hipFree(m->y);
  delete m;
  return clear_error();
}

acpp_status acpp_cost_model_choose_segmented(acpp_cost_model *m, acpp_stream_t stream, const uint64_t *offsets,
                                             uint64_t num_segments, int elem, int kind, acpp_algorithm_op op,
                                             acpp_launch_config *config, float *predicted_us) {
  const char *name = "acpp_cost_model_choose_segmented";
  if (!m || !offsets || !config) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": NULL pointer argument");
  if (num_segments == 0) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": no segments");
  int bytes = 0;
  bool is_float = false;
  switch (elem) {
    case ACPP_ELEM_I32: bytes = 4; break;
    case ACPP_ELEM_I64: bytes = 8; break;
    case ACPP_ELEM_F32: bytes = 4; is_float = true; break;
    case ACPP_ELEM_F64: bytes = 8; is_float = true; break;
    default: return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": element type must be i32, i64, f32 or f64");
  }
  if (kind != ACPP_SEGMENTED_SCAN && kind != ACPP_SEGMENTED_REDUCE)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": kind must be scan or reduce");
  double shape[11];
  mlp::segmented_features(offsets, num_segments, bytes, is_float, kind == ACPP_SEGMENTED_SCAN, op == ACPP_OP_PLUS,
                          shape);
  const double n = static_cast<double>(offsets[num_segments] - offsets[0]);
  float x[mlp::seg_candidates * ACPP_MLP_INPUTS];
  int c = 0;
  for (uint32_t wg : mlp::seg_wgs)
    for (uint32_t rows : mlp::seg_rows) {
      double f[ACPP_MLP_INPUTS];
      for (int i = 0; i < 11; ++i) f[i] = shape[i];
      const double tile = static_cast<double>(wg) * rows * (16 / bytes);
      const double tiles = n / tile;
      f[11] = std::log2(static_cast<double>(wg));
      f[12] = std::log2(static_cast<double>(rows));
      f[13] = std::log2(tile);
      f[14] = std::log2(tiles + 1.0);
      f[15] = std::log2(static_cast<double>(num_segments) / (tiles + 1.0) + 1.0);
      for (int i = 0; i < ACPP_MLP_INPUTS; ++i)
        x[c * ACPP_MLP_INPUTS + i] = static_cast<float>((f[i] - m->x_mean[i]) / m->x_std[i]);
      ++c;
    }
  int prev = 0;
  ACPP_CHECK_HIP(hipGetDevice(&prev));
  if (prev != m->device) ACPP_CHECK_HIP(hipSetDevice(m->device));
  hipStream_t s = static_cast<hipStream_t>(stream);
  float y[mlp::seg_candidates];
  hipError_t e = hipMemcpyAsync(m->x, x, sizeof(x), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(mlp::forward_kernel, dim3(1), dim3(256), 0, s, m->params, m->x,
                       static_cast<uint32_t>(mlp::seg_candidates), m->y);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(y, m->y, sizeof(y), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // x and y live on this stack frame
  if (prev != m->device) (void)hipSetDevice(prev);
  if (e != hipSuccess) return hip_error(e, name);
  int best = 0;
  for (int i = 1; i < mlp::seg_candidates; ++i)
    if (y[i] < y[best]) best = i;
  config->work_group_size = mlp::seg_wgs[best / 3];
  config->items_per_thread = mlp::seg_rows[best % 3];
  config->groups_per_cu = 0;
  config->flags = 0;
  if (predicted_us) *predicted_us = static_cast<float>(std::exp(y[best] * m->y_std + m->y_mean));
  return clear_error();
}

acpp_status acpp_mlp_adam_step(acpp_allocation_group *scratch, acpp_stream_t stream, float *params,
                               float *adam_m, float *adam_v, uint32_t step, const float *x,
                               const float *target, uint32_t batch, float lr, float weight_decay,
                               float *loss) {
  return mlp::train(scratch, static_cast<hipStream_t>(stream), params, adam_m, adam_v, step, x, target, batch,
                    lr, weight_decay, nullptr, loss, true, "acpp_mlp_adam_step");
}

}  // extern "C"
