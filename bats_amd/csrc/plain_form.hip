// The plain FP32 straight-translation form: the measured baseline of the optimised path.
//
// Each kernel keeps the reference's launch geometry -- one thread per sample, block = batch
// (so the batch is limited to 1024), one block per neuron / spike slot -- and its
// materialisations (the gathered [B, J, P] spike weights, the per-sample [B, J, I] weight
// gradient, the [B, I, K_pre, J] pre-spike error terms), and computes in float32 like the
// reference (with the float64 "4.0 *" subexpression of the dense solver's discriminant).
// It is deliberately not optimised; it exists so that the optimised kernels can be timed
// against the reference's structure on the same GPU. Kernel-by-kernel correspondence:
//   plain_pre_exps       bats/CudaKernels/CuFiles/Inference/compute_pre_exps.cu
//   plain_spike_weights  bats/CudaKernels/CuFiles/Inference/get_spike_weights.cu
//   plain_solve          bats/CudaKernels/CuFiles/Inference/compute_spike_times.cu
//   plain_recurrent      bats/CudaKernels/CuFiles/Backpropagation/propagate_recurrent_errors.cu
//   plain_factors        bats/CudaKernels/CuFiles/Backpropagation/compute_factors.cu
//   plain_wgrad          bats/CudaKernels/CuFiles/Backpropagation/compute_weights_gradient.cu
//   plain_pre_errors     bats/CudaKernels/CuFiles/Backpropagation/propagate_errors_to_pre_spikes.cu
#include "common.hpp"

namespace bats {
namespace plain {

__global__ void pre_exps(const float *__restrict__ t, float *__restrict__ es,
                         float *__restrict__ e, float tau_s, float tau) {
  const int K = gridDim.x, N = gridDim.y;
  const int b = threadIdx.x, k = blockIdx.x, i = blockIdx.y;
  const size_t s = ((size_t)b * N + i) * K + k;
  if (blockIdx.z == 0)
    es[s] = __expf(t[s] / tau_s);
  else
    e[s] = __expf(t[s] / tau);
}

__global__ void spike_weights(const float *__restrict__ W, const int *__restrict__ idx,
                              float *__restrict__ out, int I) {
  const int P = gridDim.x, J = gridDim.y;
  const int b = threadIdx.x, q = blockIdx.x, j = blockIdx.y;
  const int i = idx[(size_t)b * P + q];
  out[((size_t)b * J + j) * P + q] = W[(size_t)j * I + i];
}

__global__ void solve(const float *__restrict__ st, const float *__restrict__ ses,
                      const float *__restrict__ se, const float *__restrict__ sw, float c,
                      float dtt, float tau, float tmax, int P, int K, int *__restrict__ n_out,
                      float *__restrict__ a_out, float *__restrict__ x_out,
                      float *__restrict__ t_out, float *__restrict__ z_out) {
  const int J = gridDim.x;
  const int b = threadIdx.x, j = blockIdx.x;
  const float *times = st + (size_t)b * P;
  const float *es = ses + (size_t)b * P;
  const float *e = se + (size_t)b * P;
  const float *w = sw + ((size_t)b * J + j) * P;
  const size_t nrow = (size_t)b * J + j;
  const size_t row = nrow * K;
  float ca = 0.f, cb = 0.f;
  int n = 0;
  for (int q = 0; q < P; ++q) {
    if (times[q] == kInfF) break;
    ca += w[q] * es[q];
    cb += w[q] * e[q];
    const float next = (q + 1 < P) ? times[q + 1] : kInfF;
    float last = times[q];
    bool full = false;
    while (true) {
      // float32 result of the mixed float32 / float64 expression, as the reference stores it
      const float disc = (float)((double)(cb * cb) - 4.0 * (double)ca * (double)c);
      if (disc < 0.f) break;
      const float x = sqrtf(disc);
      const float den = x + cb;
      if (den == 0.f) break;
      const float z = 2.f * ca / den;
      if (z < 0.f) break;
      const float tt = tau * __logf(z);
      if (tt <= last || tt > tmax || tt > next) break;
      a_out[row + n] = ca;
      x_out[row + n] = x;
      t_out[row + n] = tt;
      z_out[row + n] = z;
      last = tt;
      cb -= dtt * z;
      ++n;
      if (n >= K) {
        full = true;
        break;
      }
    }
    if (full) break;
  }
  n_out[nrow] = n;
}

__global__ void recurrent(const float *__restrict__ x, const float *__restrict__ z,
                          float *__restrict__ err, float dtt, int K) {
  const int J = gridDim.x;
  const int b = threadIdx.x, j = blockIdx.x;
  const size_t row = ((size_t)b * J + j) * K;
  float C = 0.f;
  for (int k = K - 1; k >= 0; --k) {
    if (isinf(z[row + k])) continue;
    err[row + k] += C * z[row + k];
    C += err[row + k] * dtt / x[row + k];
  }
}

__global__ void factors(const float *__restrict__ t, const float *__restrict__ a, float c,
                        const float *__restrict__ x, const float *__restrict__ z, float tau,
                        float *__restrict__ f1, float *__restrict__ f2) {
  const int K = gridDim.x, J = gridDim.y;
  const int b = threadIdx.x, k = blockIdx.x, j = blockIdx.y;
  const size_t s = ((size_t)b * J + j) * K + k;
  if (isinf(t[s])) {
    if (blockIdx.z == 0)
      f1[s] = 0.f;
    else
      f2[s] = 0.f;
    return;
  }
  if (blockIdx.z == 0)
    f1[s] = tau / a[s] * (1.f + z[s] * c / x[s]);
  else
    f2[s] = tau / x[s];
}

__global__ void wgrad(const float *__restrict__ f1, const float *__restrict__ f2,
                      const float *__restrict__ post_t, const float *__restrict__ pre_t,
                      const float *__restrict__ pre_es, const float *__restrict__ pre_e,
                      const float *__restrict__ err, float *__restrict__ grad, int K, int Kp) {
  const int J = gridDim.x, I = gridDim.y;
  const int b = threadIdx.x, j = blockIdx.x, i = blockIdx.y;
  const size_t prow = ((size_t)b * J + j) * K;
  const size_t qrow = ((size_t)b * I + i) * Kp;
  float sa = 0.f, sb = 0.f, g = 0.f;
  int m = 0;
  for (int k = 0; k < K; ++k) {
    const float tk = post_t[prow + k];
    if (isinf(tk)) break;
    while (m < Kp && pre_t[qrow + m] < tk) {
      sa += pre_es[qrow + m];
      sb += pre_e[qrow + m];
      ++m;
    }
    g += err[prow + k] * (f1[prow + k] * sa - f2[prow + k] * sb);
  }
  grad[((size_t)b * J + j) * I + i] += g;
}

__global__ void pre_errors(const float *__restrict__ f1, const float *__restrict__ f2,
                           const float *__restrict__ post_t, const float *__restrict__ pre_t,
                           const float *__restrict__ pre_es, const float *__restrict__ pre_e,
                           const float *__restrict__ W, const float *__restrict__ err,
                           float *__restrict__ out, int K, float tau_s, float tau) {
  const int I = gridDim.x, Kp = gridDim.y, J = gridDim.z;
  const int b = threadIdx.x, i = blockIdx.x, m = blockIdx.y, j = blockIdx.z;
  const size_t ps = ((size_t)b * I + i) * Kp + m;
  const float tp = pre_t[ps];
  if (isinf(tp)) return;
  const size_t prow = ((size_t)b * J + j) * K;
  float c1 = 0.f, c2 = 0.f;
  for (int k = 0; k < K; ++k) {
    const float tk = post_t[prow + k];
    if (isinf(tk)) break;
    if (tk <= tp) continue;
    c1 += err[prow + k] * f1[prow + k];
    c2 += err[prow + k] * f2[prow + k];
  }
  out[ps * J + j] += W[(size_t)j * I + i] * (c1 * pre_es[ps] / tau_s - c2 * pre_e[ps] / tau);
}

}  // namespace plain
}  // namespace bats

using namespace bats;

#define PLAIN_CHECK_B(B) \
  BATS_REQUIRE((B) >= 1 && (B) <= 1024, "the plain form needs 1 <= batch <= 1024 (one thread per sample)")

extern "C" int bats_plain_pre_exps(void *stream, int B, int N, int K, const float *t, float *es,
                                   float *e, float tau_s, float tau) {
  PLAIN_CHECK_B(B);
  if (N == 0 || K == 0) return 0;
  hipLaunchKernelGGL(plain::pre_exps, dim3(K, N, 2), dim3(B), 0, as_stream(stream), t, es, e,
                     tau_s, tau);
  return check_launch("plain::pre_exps");
}

extern "C" int bats_plain_spike_weights(void *stream, int B, int P, int J, int I, const float *W,
                                        const int *idx, float *out) {
  PLAIN_CHECK_B(B);
  if (P == 0 || J == 0) return 0;
  hipLaunchKernelGGL(plain::spike_weights, dim3(P, J), dim3(B), 0, as_stream(stream), W, idx, out,
                     I);
  return check_launch("plain::spike_weights");
}

extern "C" int bats_plain_solve(void *stream, int B, int P, int J, int K, const float *st,
                                const float *ses, const float *se, const float *sw, float c,
                                float dtt, float tau, float tmax, int *n, float *a, float *x,
                                float *t, float *z) {
  PLAIN_CHECK_B(B);
  if (J == 0) return 0;
  hipLaunchKernelGGL(plain::solve, dim3(J), dim3(B), 0, as_stream(stream), st, ses, se, sw, c,
                     dtt, tau, tmax, P, K, n, a, x, t, z);
  return check_launch("plain::solve");
}

extern "C" int bats_plain_recurrent(void *stream, int B, int J, int K, const float *x,
                                    const float *z, float *errors, float dtt) {
  PLAIN_CHECK_B(B);
  if (J == 0) return 0;
  hipLaunchKernelGGL(plain::recurrent, dim3(J), dim3(B), 0, as_stream(stream), x, z, errors, dtt,
                     K);
  return check_launch("plain::recurrent");
}

extern "C" int bats_plain_factors(void *stream, int B, int J, int K, const float *t,
                                  const float *a, float c, const float *x, const float *z,
                                  float tau, float *f1, float *f2) {
  PLAIN_CHECK_B(B);
  if (J == 0 || K == 0) return 0;
  hipLaunchKernelGGL(plain::factors, dim3(K, J, 2), dim3(B), 0, as_stream(stream), t, a, c, x, z,
                     tau, f1, f2);
  return check_launch("plain::factors");
}

extern "C" int bats_plain_wgrad(void *stream, int B, int J, int I, int K, int K_pre,
                                const float *f1, const float *f2, const float *post_t,
                                const float *pre_t, const float *pre_es, const float *pre_e,
                                const float *errors, float *grad) {
  PLAIN_CHECK_B(B);
  if (J == 0 || I == 0) return 0;
  hipLaunchKernelGGL(plain::wgrad, dim3(J, I), dim3(B), 0, as_stream(stream), f1, f2, post_t,
                     pre_t, pre_es, pre_e, errors, grad, K, K_pre);
  return check_launch("plain::wgrad");
}

extern "C" int bats_plain_pre_errors(void *stream, int B, int I, int K_pre, int J, int K,
                                     const float *f1, const float *f2, const float *post_t,
                                     const float *pre_t, const float *pre_es,
                                     const float *pre_e, const float *W, const float *errors,
                                     float *out, float tau_s, float tau) {
  PLAIN_CHECK_B(B);
  if (I == 0 || K_pre == 0 || J == 0) return 0;
  BATS_REQUIRE(J <= 65535 && K_pre <= 65535, "grid too large for the plain form");
  hipLaunchKernelGGL(plain::pre_errors, dim3(I, K_pre, J), dim3(B), 0, as_stream(stream), f1, f2,
                     post_t, pre_t, pre_es, pre_e, W, errors, out, K, tau_s, tau);
  return check_launch("plain::pre_errors");
}
