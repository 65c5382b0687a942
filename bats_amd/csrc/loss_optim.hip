// Losses, optimizer steps and event-list conversion for gfx950.
//
// Each kernel computes in float64 and stores float32 at the points where the reference
// keeps float32 arrays, so that results match a float64 evaluation of the reference's
// expressions in the reference's operation order.
#include "common.hpp"

#include <algorithm>

namespace bats {

constexpr int kThreads = 256;

static unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, 1 << 20));
}

// ---------------------------------------------------------------- spike-count loss
// bats/Losses/SpikeCountLoss.py:17-38, targets from SpikeCountClassLoss.py:15-18.
__global__ __launch_bounds__(kThreads) void count_loss_kernel(
    int B, int O, int K, const int *__restrict__ n, const int64_t *__restrict__ labels,
    const float *__restrict__ targets, float target_true, float target_false,
    float *__restrict__ loss, float *__restrict__ errors, int64_t *__restrict__ pred) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  // per-sample loss and prediction
  if (gid < B && (loss || pred)) {
    const int b = gid;
    const int64_t lab = labels ? labels[b] : -1;
    double sq = 0.0;
    int best = 0, bestn = 0;
    for (int j = 0; j < O; ++j) {
      const int nj = n[(size_t)b * O + j];
      const float tg = targets ? targets[(size_t)b * O + j] : (j == lab ? target_true : target_false);
      const double d = (double)tg - (double)(float)nj;
      sq += d * d;
      if (j == 0 || nj > bestn) {
        bestn = nj;
        best = j;
      }
    }
    if (loss) loss[b] = (float)(sq / 2.0);
    if (pred) pred[b] = best;
  }
  // errors, one coalesced element per thread
  if (errors) {
    const int64_t total = (int64_t)B * O * K;
    for (int64_t e = gid; e < total; e += (int64_t)gridDim.x * blockDim.x) {
      const int64_t bj = e / K;
      const int b = (int)(bj / O), j = (int)(bj - (int64_t)b * O);
      const int nj = n[bj];
      float tg;
      if (targets) {
        tg = targets[bj];
      } else {
        const int64_t lab = labels[b];
        tg = (j == lab) ? target_true : target_false;
      }
      errors[e] = (float)((double)tg - (double)(float)nj);
    }
  }
}

// ---------------------------------------------------------------- TTFS softmax CE
// bats/Losses/TTFSSoftmaxCrossEntropy.py:20-56.
__device__ __forceinline__ float nan_to_num(float v) {
  if (isnan(v)) return 0.f;
  if (isinf(v)) return v > 0.f ? 3.4028234663852886e38f : -3.4028234663852886e38f;
  return v;
}

__global__ __launch_bounds__(kThreads) void ttfs_loss_kernel(
    int B, int O, int K, const float *__restrict__ t, const int64_t *__restrict__ labels,
    double tau, float *__restrict__ loss, float *__restrict__ errors, int64_t *__restrict__ pred) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int64_t lab = labels ? labels[b] : 0;
  double sum = 0.0;
  for (int j = 0; j < O; ++j) {
    const float e = (float)exp(-(double)t[((size_t)b * O + j) * K] / tau);
    sum += (double)e;
  }
  const float sums = (float)sum;
  if (loss) {
    if (lab < 0 || lab >= O) {
      loss[b] = __builtin_nanf("");
    } else {
      const float el = (float)exp(-(double)t[((size_t)b * O + lab) * K] / tau);
      loss[b] = (float)(-log((double)el / (double)sums));
    }
  }
  if (errors) {
    for (int j = 0; j < O; ++j) {
      const float e = (float)exp(-(double)t[((size_t)b * O + j) * K] / tau);
      float neg = (float)(-(double)e / (double)sums);
      if (j == lab) neg = (float)((double)neg + 1.0);
      neg = (float)((double)neg / tau);
      float *er = errors + ((size_t)b * O + j) * K;
      er[0] = nan_to_num(neg);
      for (int k = 1; k < K; ++k) er[k] = 0.f;
    }
  }
  if (pred) {
    // argmin over neurons for every slot k (the reference flattens [B, K])
    for (int k = 0; k < K; ++k) {
      int best = 0;
      float bv = t[((size_t)b * O) * K + k];
      for (int j = 1; j < O; ++j) {
        const float v = t[((size_t)b * O + j) * K + k];
        if (v < bv) {
          bv = v;
          best = j;
        }
      }
      pred[(size_t)b * K + k] = best;
    }
  }
}

// ---------------------------------------------------------------- optimizers
// bats/Optimizers/AdamOptimizer.py:7-19,35-55.
// When ``step`` is not NULL the bias corrections 1 - beta^t are computed on the device from
// the step counter it points to (graph-capturable form); otherwise bias1/bias2 are used.
__global__ __launch_bounds__(kThreads) void adam_kernel(
    int64_t n, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
    float *__restrict__ delta, float *__restrict__ w, double lr, double b1, double omb1,
    double b2, double omb2, double eps, double bias1, double bias2,
    const int64_t *__restrict__ step, int nadam) {
  if (step) {
    __shared__ double s_bias[2];
    if (threadIdx.x == 0) {
      const double t = (double)*step;
      s_bias[0] = 1.0 - pow(b1, t);
      s_bias[1] = 1.0 - pow(b2, t);
    }
    __syncthreads();
    bias1 = s_bias[0];
    bias2 = s_bias[1];
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double gi = (double)g[i];
    const float mi = (float)(b1 * (double)m[i] + omb1 * gi);
    const float vi = (float)(b2 * (double)v[i] + omb2 * gi * gi);
    m[i] = mi;
    v[i] = vi;
    const double mh = (double)(float)((double)mi / bias1);
    const double vh = (double)(float)((double)vi / bias2);
    float d;
    if (nadam)
      d = (float)(-(lr / (sqrt(vh) + eps)) * (b1 * mh + omb1 * gi / bias1));
    else
      d = (float)(-(lr * mh / (sqrt(vh) + eps)));
    if (delta) delta[i] = d;
    if (w) w[i] = w[i] + d;
  }
}

// Several tensors in one launch (one optimizer step over every layer): the tensors' pointers
// and sizes travel by value in the kernel arguments; blockIdx.y selects the tensor and the
// blocks of a row stride over it. The element update is adam_kernel's.
constexpr int kMultiMax = 16;
struct MultiTensors {
  int64_t n[kMultiMax];
  const float *g[kMultiMax];
  float *m[kMultiMax];
  float *v[kMultiMax];
  float *delta[kMultiMax];
  float *w[kMultiMax];
};

__global__ __launch_bounds__(kThreads) void adam_multi_kernel(
    MultiTensors T, double lr, double b1, double omb1, double b2, double omb2, double eps,
    double bias1, double bias2, const int64_t *__restrict__ step, int nadam) {
  if (step) {
    __shared__ double s_bias[2];
    if (threadIdx.x == 0) {
      const double t = (double)*step;
      s_bias[0] = 1.0 - pow(b1, t);
      s_bias[1] = 1.0 - pow(b2, t);
    }
    __syncthreads();
    bias1 = s_bias[0];
    bias2 = s_bias[1];
  }
  const int k = blockIdx.y;
  const int64_t n = T.n[k];
  const float *__restrict__ g = T.g[k];
  float *__restrict__ m = T.m[k];
  float *__restrict__ v = T.v[k];
  float *__restrict__ delta = T.delta[k];
  float *__restrict__ w = T.w[k];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double gi = (double)g[i];
    const float mi = (float)(b1 * (double)m[i] + omb1 * gi);
    const float vi = (float)(b2 * (double)v[i] + omb2 * gi * gi);
    m[i] = mi;
    v[i] = vi;
    const double mh = (double)(float)((double)mi / bias1);
    const double vh = (double)(float)((double)vi / bias2);
    float d;
    if (nadam)
      d = (float)(-(lr / (sqrt(vh) + eps)) * (b1 * mh + omb1 * gi / bias1));
    else
      d = (float)(-(lr * mh / (sqrt(vh) + eps)));
    if (delta) delta[i] = d;
    if (w) w[i] = w[i] + d;
  }
}

__global__ __launch_bounds__(kThreads) void add_multi_kernel(MultiTensors T) {
  const int k = blockIdx.y;
  const int64_t n = T.n[k];
  float *__restrict__ w = T.w[k];
  const float *__restrict__ d = T.g[k];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = w[i] + d[i];
}

__global__ __launch_bounds__(kThreads) void sgd_kernel(int64_t n, const float *__restrict__ g,
                                                        float *__restrict__ delta,
                                                        float *__restrict__ w, double lr) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float d = (float)(-lr * (double)g[i]);
    if (delta) delta[i] = d;
    if (w) w[i] = w[i] + d;
  }
}

__global__ __launch_bounds__(kThreads) void add_kernel(int64_t n, float *__restrict__ w,
                                                       const float *__restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = w[i] + d[i];
}

// ---------------------------------------------------------------- event lists
// bats/CudaKernels/CuFiles/Inference/convert_to_spike_per_neuron.cu:9-24. One thread per
// sample keeps the reference's sequential per-sample order (the list is time-ordered, so
// each neuron's buffer comes out ascending); the +inf fill is a separate coalesced pass.
// Used only for layers wider than the LDS counters of the wave form below.
__global__ __launch_bounds__(kThreads) void events_to_neuron_kernel(
    int B, int P, int N, int K, const int *__restrict__ idx, const float *__restrict__ times,
    float *__restrict__ out_t, int *__restrict__ out_n) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  int *cnt = out_n + (size_t)b * N;
  float *ot = out_t + (size_t)b * N * K;
  for (int p = 0; p < P; ++p) {
    const int i = idx[(size_t)b * P + p];
    if (i < 0) break;
    if (i >= N) continue;
    const int c = cnt[i];
    if (c < K) ot[(size_t)i * K + c] = times[(size_t)b * P + p];
    cnt[i] = c + 1;
  }
}

// Largest neuron count whose per-neuron counters and lane masks fit the LDS of one wave's
// workgroup (12 bytes per neuron).
constexpr int kEventsLdsNeurons = 4096;

// The same conversion with one wave per sample and the sample's counters in LDS. The wave
// reads its list 64 events at a time. Each lane sets its bit in its neuron's 64-bit lane mask
// (an LDS atomic OR; a wave's LDS operations complete in order), so a lane's slot is the
// neuron's count before the chunk plus the number of lower lanes in the mask, and the highest
// lane in the mask stores the new count: the c-th slot of a neuron holds its c-th event in
// list order, as in the sequential loop. A negative index ends the sample's list; indices
// >= N are skipped. The wave first fills its sample's [N,K] block with +inf (the barrier's
// fence orders those stores before the scattered ones), and it writes every count of the
// sample plus, when row_max is given, the sample's largest count (events dropped beyond K
// included). Rows are [b*P, b*P + P) of a dense [B,P] list, or [row_off[b], row_off[b+1])
// of a packed list when row_off is given.
__global__ __launch_bounds__(64) void events_to_neuron_wave_kernel(
    int N, int K, int64_t P, const int *__restrict__ row_off, const int *__restrict__ idx,
    const float *__restrict__ times, float *__restrict__ out_t, int *__restrict__ out_n,
    int *__restrict__ row_max) {
  extern __shared__ unsigned long long lds_masks[];
  unsigned long long *mask = lds_masks;                     // [N]
  int *cnt = reinterpret_cast<int *>(lds_masks + N);        // [N]
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t start = row_off ? (int64_t)row_off[b] : (int64_t)b * P;
  const int64_t end = row_off ? (int64_t)row_off[b + 1] : start + P;
  float *ot = out_t + (size_t)b * N * K;
  const int64_t nk = (int64_t)N * K;
  for (int64_t q = lane; q < nk; q += 64) ot[q] = kInfF;
  for (int i = lane; i < N; i += 64) {
    cnt[i] = 0;
    mask[i] = 0ull;
  }
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int64_t p0 = start; p0 < end; p0 += 64) {
    const int64_t p = p0 + lane;
    const int i = p < end ? idx[p] : -1;
    const float t = p < end ? times[p] : 0.0f;
    const unsigned long long neg = __ballot(i < 0);
    const int n_valid = neg ? __builtin_ctzll(neg) : 64;
    const bool live = lane < n_valid && i < N;
    if (live) atomicOr(&mask[i], 1ull << lane);
    __syncthreads();
    unsigned long long m = 0ull;
    int c = 0;
    if (live) {
      m = mask[i];
      c = cnt[i] + __popcll(m & below);
      if (c < K) ot[(size_t)i * K + c] = t;
    }
    __syncthreads();
    if (live && (m >> lane) == 1ull) {  // highest lane of the neuron in this chunk
      cnt[i] = c + 1;
      mask[i] = 0ull;
    }
    __syncthreads();
    if (n_valid < 64) break;
  }
  __syncthreads();
  int mx = 0;
  int *on = out_n + (size_t)b * N;
  for (int i = lane; i < N; i += 64) {
    const int c = cnt[i];
    on[i] = c;
    mx = c > mx ? c : mx;
  }
  if (row_max) {
    for (int off = 32; off > 0; off >>= 1) {
      const int o = __shfl_xor(mx, off);
      mx = o > mx ? o : mx;
    }
    if (lane == 0) row_max[b] = mx;
  }
}

// ---------------------------------------------------------------- input encoding
// experiments/mnist/Dataset.py:66-71: t = window * (1 - px / max) in float64, +inf where that
// equals the window (px = 0), stored as float32; one spike slot per input neuron. Grid-stride,
// coalesced.
template <typename T>
__global__ __launch_bounds__(kThreads) void latency_encode_kernel(
    int64_t n, const T *__restrict__ px, double window, double max_value,
    float *__restrict__ times, int *__restrict__ counts) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double t = window * (1.0 - (double)px[i] / max_value);
    const bool fire = t != window;
    times[i] = fire ? (float)t : kInfF;
    counts[i] = fire ? 1 : 0;
  }
}

}  // namespace bats

using namespace bats;

extern "C" int bats_latency_encode(void *stream, int64_t n, const void *pixels, int pixel_type,
                                   double time_window, double max_value, float *times,
                                   int *counts) {
  BATS_REQUIRE(n >= 0, "negative size");
  BATS_REQUIRE(pixel_type == 0 || pixel_type == 1, "pixel_type must be 0 (uint8) or 1 (float32)");
  BATS_REQUIRE(max_value != 0.0, "max_value must be non-zero");
  if (n == 0) return 0;
  BATS_REQUIRE(pixels && times && counts, "null pointer");
  const unsigned grid = std::min<unsigned>(grid_for(n), 4096);
  if (pixel_type == 0)
    hipLaunchKernelGGL(latency_encode_kernel<uint8_t>, dim3(grid), dim3(kThreads), 0,
                       as_stream(stream), n, static_cast<const uint8_t *>(pixels),
                       time_window, max_value, times, counts);
  else
    hipLaunchKernelGGL(latency_encode_kernel<float>, dim3(grid), dim3(kThreads), 0,
                       as_stream(stream), n, static_cast<const float *>(pixels),
                       time_window, max_value, times, counts);
  return check_launch("latency_encode_kernel");
}

extern "C" int bats_count_loss(void *stream, int B, int O, int K, const int *n,
                               const int64_t *labels, float target_true, float target_false,
                               float *loss, float *errors, int64_t *pred) {
  BATS_REQUIRE(B >= 0 && O >= 1 && K >= 0, "bad sizes");
  if (B == 0) return 0;
  BATS_REQUIRE(n && labels, "null n or labels");
  const int64_t work = std::max<int64_t>(B, errors ? (int64_t)B * O * K : 0);
  hipLaunchKernelGGL(count_loss_kernel, dim3(std::min<unsigned>(grid_for(work), 2048)),
                     dim3(kThreads), 0, as_stream(stream), B, O, K, n, labels,
                     (const float *)nullptr, target_true, target_false, loss, errors, pred);
  return check_launch("count_loss_kernel");
}

extern "C" int bats_count_loss_targets(void *stream, int B, int O, int K, const int *n,
                                       const float *targets, float *loss, float *errors,
                                       int64_t *pred) {
  BATS_REQUIRE(B >= 0 && O >= 1 && K >= 0, "bad sizes");
  if (B == 0) return 0;
  BATS_REQUIRE(n && targets, "null n or targets");
  const int64_t work = std::max<int64_t>(B, errors ? (int64_t)B * O * K : 0);
  hipLaunchKernelGGL(count_loss_kernel, dim3(std::min<unsigned>(grid_for(work), 2048)),
                     dim3(kThreads), 0, as_stream(stream), B, O, K, n, (const int64_t *)nullptr,
                     targets, 0.f, 0.f, loss, errors, pred);
  return check_launch("count_loss_kernel");
}

extern "C" int bats_ttfs_loss(void *stream, int B, int O, int K, const float *t,
                              const int64_t *labels, float tau, float *loss, float *errors,
                              int64_t *pred) {
  BATS_REQUIRE(B >= 0 && O >= 1 && K >= 1, "bad sizes");
  if (B == 0) return 0;
  BATS_REQUIRE(t, "null t");
  BATS_REQUIRE(labels || (!loss && !errors), "labels required for loss/errors");
  hipLaunchKernelGGL(ttfs_loss_kernel, dim3(grid_for(B)), dim3(kThreads), 0, as_stream(stream), B,
                     O, K, t, labels, (double)tau, loss, errors, pred);
  return check_launch("ttfs_loss_kernel");
}

static int adam_like(void *stream, int64_t n, const float *grad, float *m, float *v, float *delta,
                     float *weights, float lr, float beta_1, float omb1, float beta_2, float omb2,
                     float eps, double bias1, double bias2, int nadam,
                     const int64_t *step = nullptr) {
  BATS_REQUIRE(n >= 0, "negative size");
  if (n == 0) return 0;
  BATS_REQUIRE(grad && m && v && (delta || weights), "null pointer");
  hipLaunchKernelGGL(adam_kernel, dim3(std::min<unsigned>(grid_for(n), 4096)), dim3(kThreads), 0,
                     as_stream(stream), n, grad, m, v, delta, weights, (double)lr,
                     (double)beta_1, (double)omb1, (double)beta_2, (double)omb2, (double)eps,
                     bias1, bias2, step, nadam);
  return check_launch("adam_kernel");
}

extern "C" int bats_adam_step_dev(void *stream, int64_t n, const float *grad, float *m, float *v,
                                  float *delta, float *weights, float lr, float beta_1,
                                  float one_minus_beta_1, float beta_2, float one_minus_beta_2,
                                  float epsilon, const int64_t *step, int nadam) {
  BATS_REQUIRE(step != nullptr, "null step counter");
  return adam_like(stream, n, grad, m, v, delta, weights, lr, beta_1, one_minus_beta_1, beta_2,
                   one_minus_beta_2, epsilon, 0.0, 0.0, nadam ? 1 : 0, step);
}

extern "C" int bats_adam_step(void *stream, int64_t n, const float *grad, float *m, float *v,
                              float *delta, float *weights, float lr, float beta_1,
                              float one_minus_beta_1, float beta_2, float one_minus_beta_2,
                              float epsilon, double bias1, double bias2) {
  return adam_like(stream, n, grad, m, v, delta, weights, lr, beta_1, one_minus_beta_1, beta_2,
                   one_minus_beta_2, epsilon, bias1, bias2, 0);
}

extern "C" int bats_nadam_step(void *stream, int64_t n, const float *grad, float *m, float *v,
                               float *delta, float *weights, float lr, float beta_1,
                               float one_minus_beta_1, float beta_2, float one_minus_beta_2,
                               float epsilon, double bias1, double bias2) {
  return adam_like(stream, n, grad, m, v, delta, weights, lr, beta_1, one_minus_beta_1, beta_2,
                   one_minus_beta_2, epsilon, bias1, bias2, 1);
}

extern "C" int bats_adam_step_multi(void *stream, int count, const int64_t *sizes,
                                    const float *const *grads, float *const *m,
                                    float *const *v, float *const *deltas,
                                    float *const *weights, float lr, float beta_1,
                                    float one_minus_beta_1, float beta_2, float one_minus_beta_2,
                                    float epsilon, double bias1, double bias2,
                                    const int64_t *step, int nadam) {
  BATS_REQUIRE(count >= 0 && count <= kMultiMax, "count must be in [0, 16]");
  if (count == 0) return 0;
  BATS_REQUIRE(sizes && grads && m && v, "null array");
  MultiTensors T{};
  int64_t nmax = 0;
  for (int k = 0; k < count; ++k) {
    BATS_REQUIRE(sizes[k] >= 0, "negative size");
    BATS_REQUIRE(sizes[k] == 0 || (grads[k] && m[k] && v[k] &&
                                   ((deltas && deltas[k]) || (weights && weights[k]))),
                 "null pointer");
    T.n[k] = sizes[k];
    T.g[k] = grads[k];
    T.m[k] = m[k];
    T.v[k] = v[k];
    T.delta[k] = deltas ? deltas[k] : nullptr;
    T.w[k] = weights ? weights[k] : nullptr;
    nmax = std::max(nmax, sizes[k]);
  }
  if (nmax == 0) return 0;
  hipLaunchKernelGGL(adam_multi_kernel, dim3(std::min<unsigned>(grid_for(nmax), 4096), count),
                     dim3(kThreads), 0, as_stream(stream), T, (double)lr, (double)beta_1,
                     (double)one_minus_beta_1, (double)beta_2, (double)one_minus_beta_2,
                     (double)epsilon, bias1, bias2, step, nadam ? 1 : 0);
  return check_launch("adam_multi_kernel");
}

extern "C" int bats_add_deltas_multi(void *stream, int count, const int64_t *sizes,
                                     float *const *weights, const float *const *deltas) {
  BATS_REQUIRE(count >= 0 && count <= kMultiMax, "count must be in [0, 16]");
  if (count == 0) return 0;
  BATS_REQUIRE(sizes && weights && deltas, "null array");
  MultiTensors T{};
  int64_t nmax = 0;
  for (int k = 0; k < count; ++k) {
    BATS_REQUIRE(sizes[k] >= 0, "negative size");
    BATS_REQUIRE(sizes[k] == 0 || (weights[k] && deltas[k]), "null pointer");
    T.n[k] = sizes[k];
    T.w[k] = weights[k];
    T.g[k] = deltas[k];
    nmax = std::max(nmax, sizes[k]);
  }
  if (nmax == 0) return 0;
  hipLaunchKernelGGL(add_multi_kernel, dim3(std::min<unsigned>(grid_for(nmax), 4096), count),
                     dim3(kThreads), 0, as_stream(stream), T);
  return check_launch("add_multi_kernel");
}

extern "C" int bats_sgd_step(void *stream, int64_t n, const float *grad, float *delta,
                             float *weights, float lr) {
  BATS_REQUIRE(n >= 0, "negative size");
  if (n == 0) return 0;
  BATS_REQUIRE(grad && (delta || weights), "null pointer");
  hipLaunchKernelGGL(sgd_kernel, dim3(std::min<unsigned>(grid_for(n), 4096)), dim3(kThreads), 0,
                     as_stream(stream), n, grad, delta, weights, (double)lr);
  return check_launch("sgd_kernel");
}

extern "C" int bats_add_deltas(void *stream, int64_t n, float *weights, const float *delta) {
  BATS_REQUIRE(n >= 0, "negative size");
  if (n == 0) return 0;
  BATS_REQUIRE(weights && delta, "null pointer");
  hipLaunchKernelGGL(add_kernel, dim3(std::min<unsigned>(grid_for(n), 4096)), dim3(kThreads), 0,
                     as_stream(stream), n, weights, delta);
  return check_launch("add_kernel");
}

extern "C" int bats_events_to_per_neuron(void *stream, int B, int P, int N, int K,
                                         const int *indices, const float *times,
                                         float *out_times, int *out_counts) {
  BATS_REQUIRE(B >= 0 && P >= 0 && N >= 0 && K >= 0, "bad sizes");
  if (B == 0 || N == 0) return 0;
  BATS_REQUIRE(out_times && out_counts, "null output");
  BATS_REQUIRE(P == 0 || (indices && times), "null input");
  hipStream_t st = as_stream(stream);
  if (N <= kEventsLdsNeurons) {
    hipLaunchKernelGGL(events_to_neuron_wave_kernel, dim3(B), dim3(64), (size_t)12 * N, st, N, K,
                       (int64_t)P, nullptr, indices, times, out_times, out_counts, nullptr);
    return check_launch("events_to_neuron_wave_kernel");
  }
  const int64_t nt = (int64_t)B * N * K, nc = (int64_t)B * N;
  if (nt > 0) {
    hipLaunchKernelGGL(fill_kernel<float>, dim3(std::min<unsigned>(grid_for(nt), 4096)),
                       dim3(kThreads), 0, st, out_times, nt, kInfF);
    if (int r = check_launch("fill times")) return r;
  }
  hipLaunchKernelGGL(fill_kernel<int>, dim3(std::min<unsigned>(grid_for(nc), 4096)), dim3(kThreads),
                     0, st, out_counts, nc, 0);
  if (int r = check_launch("fill counts")) return r;
  if (P > 0) {
    hipLaunchKernelGGL(events_to_neuron_kernel, dim3(grid_for(B)), dim3(kThreads), 0, st, B, P, N,
                       K, indices, times, out_times, out_counts);
    if (int r = check_launch("events_to_neuron_kernel")) return r;
  }
  return 0;
}

extern "C" int bats_events_csr_to_per_neuron(void *stream, int B, int N, int K,
                                             const int *row_offsets, const int *indices,
                                             const float *times, float *out_times,
                                             int *out_counts, int *out_row_max) {
  BATS_REQUIRE(B >= 0 && N >= 0 && K >= 0, "bad sizes");
  if (B == 0) return 0;
  BATS_REQUIRE(N > 0 && N <= kEventsLdsNeurons, "N must be in [1, 4096]");
  BATS_REQUIRE(row_offsets && indices && times && out_times && out_counts, "null pointer");
  hipLaunchKernelGGL(events_to_neuron_wave_kernel, dim3(B), dim3(64), (size_t)12 * N,
                     as_stream(stream), N, K, (int64_t)0, row_offsets, indices, times, out_times,
                     out_counts, out_row_max);
  return check_launch("events_to_neuron_wave_kernel");
}
