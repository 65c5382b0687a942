// Dense LIF layer BATS backward for gfx950.
//
// Both the weight gradient (reference compute_weights_gradient.cu:33-45) and the errors of
// the pre-synaptic spikes (propagate_errors_to_pre_spikes.cu:41-53) are functions of the
// same quantity: for a pre spike at time t_m and a post neuron j of the same sample,
//   V1_j(t_m) = sum_{k: t_jk > t_m} e_jk f1_jk,   V2_j(t_m) = sum_{k: t_jk > t_m} e_jk f2_jk.
// Regrouping the reference's per-post-spike sums gives
//   dW[j,i]      = sum_b sum_m  es_m V1_j(t_m) - e_m V2_j(t_m)       (m: spikes of pre i)
//   pre_err[b,i,m] = sum_j W[j,i] (V1_j(t_m) es_m / tau_s - V2_j(t_m) e_m / tau)
// so the backward is three kernels:
//   1. post_prep_kernel  one lane per (sample, post neuron): the reset-chain error scan
//      (propagate_recurrent_errors.cu:17-22, written back in place like the reference),
//      the factors f1, f2 (compute_factors.cu:20-30, rounded to float32 as stored there),
//      and the float64 suffix sums S1[k] = sum_{k'>=k} e f1, S2[k] = sum_{k'>=k} e f2.
//      V_j(t) is then S[first k with t_jk > t].
//   2. wgrad_kernel      lanes over pre neurons i, a tile of post neurons j, a slice of the
//      batch; the post rows of a few samples at a time are staged in LDS (wave-uniform
//      reads), the lookup counts the post spikes at or before t_m; float64 accumulation;
//      a deterministic fixed-order reduction over batch slices (wgrad_reduce_kernel)
//      finishes the mean.
//   3. pre_err_kernel    lanes over pre neurons i of one sample, post rows staged in LDS
//      a tile at a time.
#include "common.hpp"

#include <algorithm>

namespace bats {

constexpr int kBwdThreads = 256;

__global__ __launch_bounds__(kBwdThreads) void post_prep_kernel(
    int BJ, int K, const int *__restrict__ n, const float *__restrict__ t,
    const float *__restrict__ a, const float *__restrict__ x, const float *__restrict__ z,
    float *__restrict__ errors, double c, double tau, double dtt, int *__restrict__ n_eff,
    double *__restrict__ S1, double *__restrict__ S2) {
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= BJ) return;
  const size_t row = (size_t)lane * K;
  const int nk = min(n[lane], K);
  // reset chain, reverse order (propagate_recurrent_errors.cu:17-22)
  double C = 0.0;
  for (int k = nk - 1; k >= 0; --k) {
    const float zk = z[row + k];
    if (isinf(zk)) continue;
    const double ek = (double)errors[row + k] + C * (double)zk;
    const float ek32 = (float)ek;
    errors[row + k] = ek32;
    C = C + (double)ek32 * dtt / (double)x[row + k];
  }
  // spikes used by the gradient: those before the first infinite time
  int ne = 0;
  while (ne < nk && !isinf(t[row + ne])) ++ne;
  double s1 = 0.0, s2 = 0.0;
  for (int k = ne - 1; k >= 0; --k) {
    const double ak = (double)a[row + k], xk = (double)x[row + k], zk = (double)z[row + k];
    const float f1 = (float)(tau / ak * (1.0 + zk * c / xk));
    const float f2 = (float)(tau / xk);
    const double ek = (double)errors[row + k];
    s1 += ek * (double)f1;
    s2 += ek * (double)f2;
    S1[row + k] = s1;
    S2[row + k] = s2;
  }
  n_eff[lane] = ne;
}

// Post-neuron rows staged in LDS for SB samples at a time: spike counts, times and suffix
// sums of rows (b, j0 .. j0+jn-1), b = b0 .. b0+sb-1. Rows are copied whole (slots past a
// row's spike count are never read), so one copy and two barriers serve SB samples.
struct RowStage {
  int *ne;     // [SB][JT]
  float *t;    // [SB][JT][K]
  double *s1;  // [SB][JT][K]
  double *s2;  // [SB][JT][K]
};

__device__ __forceinline__ RowStage row_stage(char *smem, int SB, int JT, int K) {
  RowStage r;
  const size_t n = (size_t)SB * JT * K;
  r.s1 = reinterpret_cast<double *>(smem);
  r.s2 = r.s1 + n;
  r.t = reinterpret_cast<float *>(r.s2 + n);
  r.ne = reinterpret_cast<int *>(r.t + n);
  return r;
}

__host__ __device__ constexpr size_t row_stage_bytes(int SB, int JT, int K) {
  return (size_t)SB * JT * K * (2 * sizeof(double) + sizeof(float)) + (size_t)SB * JT * sizeof(int);
}

// Copies rows of samples b0 .. b0+sb-1 into the stage. Contains barriers: call from every
// thread of the block.
__device__ __forceinline__ void load_rows(const RowStage &st, int b0, int sb, int J, int j0,
                                          int jn, int JT, int K,
                                          const float *__restrict__ post_t,
                                          const int *__restrict__ n_eff,
                                          const double *__restrict__ S1,
                                          const double *__restrict__ S2) {
  const int tid = threadIdx.x, nth = blockDim.x;
  __syncthreads();  // previous contents consumed
  const int per = jn * K;
  for (int e = tid; e < sb * per; e += nth) {
    const int s = e / per;
    const int r = e - s * per;
    const size_t src = ((size_t)(b0 + s) * J + j0) * K + r;
    const int dst = s * JT * K + r;
    st.t[dst] = post_t[src];
    st.s1[dst] = S1[src];
    st.s2[dst] = S2[src];
  }
  for (int e = tid; e < sb * jn; e += nth) {
    const int s = e / jn;
    const int r = e - s * jn;
    st.ne[s * JT + r] = n_eff[(size_t)(b0 + s) * J + j0 + r];
  }
  __syncthreads();
}

// V(t): S[first k with tpost[k] > tp], 0 when there is none; the post times of a row are
// ascending, so the index is the count of times <= tp.
__device__ __forceinline__ void lookup_count(const float *__restrict__ tr,
                                             const double *__restrict__ s1,
                                             const double *__restrict__ s2, int nj, float tp,
                                             double &v1, double &v2) {
  int kk = 0;
  for (int k = 0; k < nj; ++k) kk += (tr[k] <= tp) ? 1 : 0;
  v1 = (kk < nj) ? s1[kk] : 0.0;
  v2 = (kk < nj) ? s2[kk] : 0.0;
}

// Number of leading finite slots of a spike row (independent loads, no dependent chain).
__device__ __forceinline__ int leading_finite(const float *__restrict__ row, int K) {
  int n = 0;
  bool run = true;
#pragma unroll 8
  for (int m = 0; m < K; ++m) {
    run = run && isfinite(row[m]);
    n += run ? 1 : 0;
  }
  return n;
}

// Lanes are pre neurons (a block covers up to 1024 of them), a block owns kWJT post neurons
// and a slice of the batch. Per group of samples the block stages its post rows in LDS;
// every lane accumulates es*V1 - e*V2 over its pre spikes in float64.
constexpr int kWJT = 16;

__global__ __launch_bounds__(1024) void wgrad_kernel(
    int B, int I, int K_pre, int J, int K, int b_per_slice, int SB,
    const float *__restrict__ pre_t, const float *__restrict__ pre_es,
    const float *__restrict__ pre_e, const float *__restrict__ post_t,
    const int *__restrict__ n_eff, const double *__restrict__ S1,
    const double *__restrict__ S2, int per_sample, float *__restrict__ grad,
    float *__restrict__ partial) {
  extern __shared__ __align__(16) char smem[];
  const RowStage st = row_stage(smem, SB, kWJT, K);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < I;
  const int ic = active ? i : 0;
  const int j0 = blockIdx.y * kWJT;
  const int jn = min(kWJT, J - j0);
  const int slice = blockIdx.z;
  const int b0 = slice * b_per_slice;
  const int b1 = min(B, b0 + b_per_slice);
  double acc[kWJT];
#pragma unroll
  for (int u = 0; u < kWJT; ++u) acc[u] = 0.0;
  for (int bg = b0; bg < b1; bg += SB) {
    const int sb = min(SB, b1 - bg);
    load_rows(st, bg, sb, J, j0, jn, kWJT, K, post_t, n_eff, S1, S2);
    for (int s = 0; s < sb; ++s) {
      const int b = bg + s;
      const size_t pbase = ((size_t)b * I + ic) * K_pre;
      const int np = active ? leading_finite(pre_t + pbase, K_pre) : 0;
      for (int m = 0; m < np; ++m) {
        const float tp = pre_t[pbase + m];
        const double es = (double)pre_es[pbase + m];
        const double e = (double)pre_e[pbase + m];
#pragma unroll
        for (int u = 0; u < kWJT; ++u) {
          const int nj = (u < jn) ? st.ne[s * kWJT + u] : 0;
          if (nj > 0) {
            const int off = (s * kWJT + u) * K;
            double v1, v2;
            lookup_count(st.t + off, st.s1 + off, st.s2 + off, nj, tp, v1, v2);
            acc[u] = fma(es, v1, acc[u]);
            acc[u] = fma(-e, v2, acc[u]);
          }
        }
      }
      if (per_sample) {
#pragma unroll
        for (int u = 0; u < kWJT; ++u) {
          if (u < jn && active) grad[((size_t)b * J + j0 + u) * I + i] = (float)acc[u];
          acc[u] = 0.0;
        }
      }
    }
  }
  if (!per_sample && active) {
#pragma unroll
    for (int u = 0; u < kWJT; ++u)
      if (u < jn) partial[((size_t)slice * J + j0 + u) * I + i] = (float)acc[u];
  }
}

__global__ __launch_bounds__(kBwdThreads) void wgrad_reduce_kernel(
    int slices, int64_t JI, double inv_b, const float *__restrict__ partial,
    float *__restrict__ grad) {
  const int64_t e = (int64_t)blockIdx.x * kBwdThreads + threadIdx.x;
  if (e >= JI) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int z = 0;
  for (; z + 4 <= slices; z += 4) {
    s0 += (double)partial[(size_t)z * JI + e];
    s1 += (double)partial[(size_t)(z + 1) * JI + e];
    s2 += (double)partial[(size_t)(z + 2) * JI + e];
    s3 += (double)partial[(size_t)(z + 3) * JI + e];
  }
  for (; z < slices; ++z) s0 += (double)partial[(size_t)z * JI + e];
  grad[e] = (float)(((s0 + s1) + (s2 + s3)) * inv_b);
}

// Lanes are pre neurons of one sample; the post rows are staged kWJT at a time and every
// pre spike of the lane accumulates W[j,i] (V1 es/tau_s - V2 e/tau) over all post neurons.
// Pre-spike sums of one lane are kept in float64 scratch rows across the row chunks.
__global__ __launch_bounds__(1024) void pre_err_kernel(
    int I, int K_pre, int J, int K, const float *__restrict__ pre_t,
    const float *__restrict__ pre_es, const float *__restrict__ pre_e,
    const float *__restrict__ W, const float *__restrict__ post_t,
    const int *__restrict__ n_eff, const double *__restrict__ S1,
    const double *__restrict__ S2, double tau_s, double tau, double *__restrict__ acc_scratch,
    float *__restrict__ pre_errors) {
  extern __shared__ __align__(16) char smem[];
  const RowStage st = row_stage(smem, 1, kWJT, K);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const bool active = i < I;
  const int ic = active ? i : 0;
  const size_t pbase = ((size_t)b * I + ic) * K_pre;
  const int np = active ? leading_finite(pre_t + pbase, K_pre) : 0;
  const bool single_chunk = J <= kWJT;
  for (int j0 = 0; j0 < J; j0 += kWJT) {
    const int jn = min(kWJT, J - j0);
    load_rows(st, b, 1, J, j0, jn, kWJT, K, post_t, n_eff, S1, S2);
    for (int m = 0; m < np; ++m) {
      const float tp = pre_t[pbase + m];
      const double es = (double)pre_es[pbase + m];
      const double e = (double)pre_e[pbase + m];
      double acc = (j0 == 0) ? 0.0 : acc_scratch[pbase + m];
      for (int u = 0; u < jn; ++u) {
        const int nj = st.ne[u];
        if (nj == 0) continue;
        double v1, v2;
        lookup_count(st.t + u * K, st.s1 + u * K, st.s2 + u * K, nj, tp, v1, v2);
        const double w = (double)W[(size_t)(j0 + u) * I + i];
        acc += w * ((v1 * es) / tau_s - (v2 * e) / tau);
      }
      if (single_chunk || j0 + kWJT >= J)
        pre_errors[pbase + m] = (float)acc;
      else
        acc_scratch[pbase + m] = acc;
    }
  }
  if (active)
    for (int m = np; m < K_pre; ++m) pre_errors[pbase + m] = 0.f;
}

struct BwdLayout {
  size_t neff, s1, s2, partial, scratch, total;
  int slices, b_per_slice, lanes, gi;
};

static BwdLayout bwd_layout(int B, int I, int K_pre, int J, int K) {
  BwdLayout L{};
  L.gi = std::max(1, (I + 1023) / 1024);
  const int per_block = (I + L.gi - 1) / L.gi;
  L.lanes = std::max(64, (per_block + 63) / 64 * 64);
  const int gj = (J + kWJT - 1) / kWJT;
  const int base = std::max(1, L.gi * gj);
  int slices = std::max(1, std::min(B, (512 + base - 1) / base));
  L.b_per_slice = B > 0 ? (B + slices - 1) / slices : 1;
  L.slices = B > 0 ? (B + L.b_per_slice - 1) / L.b_per_slice : 1;
  size_t off = 0;
  L.neff = off;
  off = align_up(off + sizeof(int) * (size_t)B * J, 256);
  L.s1 = off;
  off = align_up(off + sizeof(double) * (size_t)B * J * K, 256);
  L.s2 = off;
  off = align_up(off + sizeof(double) * (size_t)B * J * K, 256);
  L.partial = off;
  off = align_up(off + sizeof(float) * (size_t)L.slices * J * I, 256);
  L.scratch = off;
  if (J > kWJT) off = align_up(off + sizeof(double) * (size_t)B * I * K_pre, 256);
  L.total = off;
  return L;
}

int post_prep(hipStream_t st, int BJ, int K, const int *n, const float *t, const float *a,
              const float *x, const float *z, float *errors, float c, float tau, float dtt,
              int *n_eff, double *S1, double *S2) {
  if (BJ <= 0) return 0;
  hipLaunchKernelGGL(post_prep_kernel, dim3((BJ + kBwdThreads - 1) / kBwdThreads),
                     dim3(kBwdThreads), 0, st, BJ, K, n, t, a, x, z, errors, (double)c,
                     (double)tau, (double)dtt, n_eff, S1, S2);
  return check_launch("post_prep_kernel");
}

}  // namespace bats

using namespace bats;

extern "C" size_t bats_lif_backward_workspace_size(int B, int I, int K_pre, int J, int K_post) {
  if (B < 0 || I < 0 || K_pre < 0 || J < 0 || K_post < 0) return 0;
  return bwd_layout(B, I, K_pre, J, K_post).total;
}

// Largest post-spike buffer the backward's LDS row stage supports.
constexpr int kMaxBackwardK = 384;

extern "C" int bats_lif_backward(void *stream, int B, int I, int K_pre, int J, int K_post,
                                 const int *n, const float *t, const float *a, const float *x,
                                 const float *z, float *errors, const float *pre_times,
                                 const float *pre_exp_tau_s, const float *pre_exp_tau,
                                 const float *weights, float tau_s, float tau, float c,
                                 float delta_theta_tau, float *grad, int grad_mode,
                                 float *pre_errors, void *workspace, size_t workspace_bytes) {
  BATS_REQUIRE(B >= 0 && I >= 0 && K_pre >= 0 && J >= 0 && K_post >= 1, "bad sizes");
  BATS_REQUIRE(K_post <= kMaxBackwardK, "max_n_spike above 384 is not supported by the backward");
  BATS_REQUIRE(grad_mode >= 0 && grad_mode <= 2,
               "grad_mode must be 0 (batch mean), 1 (per sample) or 2 (batch sum)");
  hipStream_t st = as_stream(stream);
  const int64_t JI = (int64_t)J * I;
  if (B == 0) {
    if (grad_mode != 1 && JI > 0) {
      BATS_REQUIRE(grad, "null grad");
      hipLaunchKernelGGL(fill_kernel<float>, dim3((unsigned)std::min<int64_t>((JI + 255) / 256, 65535)),
                         dim3(256), 0, st, grad, JI, 0.f);
      return check_launch("fill grad");
    }
    return 0;
  }
  BATS_REQUIRE(n && t && a && x && z && errors, "null post-layer pointer");
  BATS_REQUIRE(grad || JI == 0, "null grad");
  const BwdLayout L = bwd_layout(B, I, K_pre, J, K_post);
  BATS_REQUIRE(workspace != nullptr || L.total == 0, "null workspace");
  BATS_REQUIRE(workspace_bytes >= L.total, "workspace too small");
  char *ws = static_cast<char *>(workspace);
  int *n_eff = reinterpret_cast<int *>(ws + L.neff);
  double *S1 = reinterpret_cast<double *>(ws + L.s1);
  double *S2 = reinterpret_cast<double *>(ws + L.s2);
  float *partial = reinterpret_cast<float *>(ws + L.partial);
  double *scratch = reinterpret_cast<double *>(ws + L.scratch);
  // samples staged per round: as many as fit 48 KB of LDS (at least one)
  const int SB = (int)std::max<size_t>(1, std::min<size_t>(4, (48 * 1024) / row_stage_bytes(1, kWJT, K_post)));
  const size_t lds_w = row_stage_bytes(SB, kWJT, K_post);
  const size_t lds_p = row_stage_bytes(1, kWJT, K_post);

  const int BJ = B * J;
  if (BJ > 0) {
    hipLaunchKernelGGL(post_prep_kernel, dim3((BJ + kBwdThreads - 1) / kBwdThreads),
                       dim3(kBwdThreads), 0, st, BJ, K_post, n, t, a, x, z, errors, (double)c,
                       (double)tau, (double)delta_theta_tau, n_eff, S1, S2);
    if (int r = check_launch("post_prep_kernel")) return r;
  }
  if (JI > 0) {
    BATS_REQUIRE(K_pre == 0 || (pre_times && pre_exp_tau_s && pre_exp_tau), "null pre pointer");
    const int gj = (J + kWJT - 1) / kWJT;
    if (grad_mode == 1) {
      dim3 g(L.gi, gj, B);  // one sample per block
      hipLaunchKernelGGL(wgrad_kernel, g, dim3(L.lanes), lds_w, st, B, I, K_pre, J, K_post, 1,
                         SB, pre_times, pre_exp_tau_s, pre_exp_tau, t, n_eff, S1, S2, 1, grad,
                         partial);
      if (int r = check_launch("wgrad_kernel")) return r;
    } else {
      dim3 g(L.gi, gj, L.slices);
      BATS_TIMED("wgrad_kernel", st,
                 hipLaunchKernelGGL(wgrad_kernel, g, dim3(L.lanes), lds_w, st, B, I, K_pre, J,
                                    K_post, L.b_per_slice, SB, pre_times, pre_exp_tau_s,
                                    pre_exp_tau, t, n_eff, S1, S2, 0, grad, partial));
      if (int r = check_launch("wgrad_kernel")) return r;
      hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((unsigned)((JI + kBwdThreads - 1) / kBwdThreads)),
                         dim3(kBwdThreads), 0, st, L.slices, JI,
                         grad_mode == 2 ? 1.0 : 1.0 / (double)B, partial, grad);
      if (int r = check_launch("wgrad_reduce_kernel")) return r;
    }
  }
  if (pre_errors && I > 0 && K_pre > 0) {
    BATS_REQUIRE(weights, "null weights");
    dim3 g(L.gi, B);
    BATS_TIMED("pre_err_kernel", st,
               hipLaunchKernelGGL(pre_err_kernel, g, dim3(L.lanes), lds_p, st, I, K_pre, J, K_post,
                                  pre_times, pre_exp_tau_s, pre_exp_tau, weights, t, n_eff, S1,
                                  S2, (double)tau_s, (double)tau, scratch, pre_errors));
    if (int r = check_launch("pre_err_kernel")) return r;
  }
  return 0;
}
