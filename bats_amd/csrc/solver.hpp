// Device-side pieces of the exact spike-time solver shared by the dense and conv layers.
#pragma once

#include "common.hpp"

namespace bats {

// ------------------------------------------------------------------ crossing filter
// Necessary condition (with a relative margin) for the reference's smaller root
// z = 2a/(b + sqrt(b^2 - 4ac)) to be accepted in (t_q, min(t_next, t_max)]:
// a > 0 and b > 0, and either g(zl) = c zl^2 - b zl + a <= 0 (one root inside) or the
// vertex b/(2c) of g lies in [zq, zl] with a non-negative discriminant (two roots inside).
constexpr double kFilterEps = 1e-10;

__device__ __forceinline__ bool crossing_candidate(double a, double b, const Event &e,
                                                   double c4, double inv2c) {
  if (!(a > 0.0) || !(b > 0.0)) return false;
  const double bb_hi = b * b * (1.0 + kFilterEps);
  if (isinf(e.zl)) return bb_hi >= c4 * a;
  const double g = (a + e.czl2) - b * e.zl * (1.0 + kFilterEps);
  if (g <= 0.0) return true;
  const double v = b * inv2c;
  if (v < e.zq * (1.0 - kFilterEps) || v > e.zl * (1.0 + kFilterEps)) return false;
  return bb_hi >= c4 * a;
}

// The reference's compute_spikes loop (compute_spike_times.cu:53-83) in float64.
// Returns false when the neuron's buffer became full (its event loop ends).
__device__ __forceinline__ bool solve_crossings(double a, double &b, int &n, double last,
                                                double tl, double c, double dtt, double tau,
                                                int K, size_t row, float *__restrict__ out_t,
                                                float *__restrict__ out_a,
                                                float *__restrict__ out_x,
                                                float *__restrict__ out_z, bool writer = true) {
  while (true) {
    const double disc = b * b - (4.0 * a) * c;
    if (disc < 0.0) return true;
    const double x = sqrt(disc);
    const double den = x + b;
    if (den == 0.0) return true;
    const double z = (2.0 * a) / den;
    if (z < 0.0) return true;
    const double t = tau * log(z);
    if (t <= last || t > tl) return true;
    if (writer) {
      out_a[row + n] = (float)a;
      out_x[row + n] = (float)x;
      out_t[row + n] = (float)t;
      out_z[row + n] = (float)z;
    }
    last = t;
    b = b - dtt * z;
    ++n;
    if (n >= K) return false;
  }
}

// Out-of-line copy of the closed-form loop for the lane-per-neuron solver, whose event loop
// is fully unrolled per chunk (the loop is entered rarely: only when the filter passes).
struct SolveState {
  double b;
  int n;
  int live;
};

static __device__ __attribute__((noinline)) SolveState solve_crossings_ool(
    double a, double b, int n, double last, double tl, double c, double dtt, double tau, int K,
    size_t row, float *out_t, float *out_a, float *out_x, float *out_z) {
  SolveState r;
  r.live = solve_crossings(a, b, n, last, tl, c, dtt, tau, K, row, out_t, out_a, out_x, out_z);
  r.b = b;
  r.n = n;
  return r;
}


// The filter with the interval given explicitly: lower end zq = exp(t_last/tau), upper end
// zl = exp(min(t_next, t_max)/tau) and czl2 = c*zl*zl.
__device__ __forceinline__ bool crossing_candidate3(double a, double b, double zq, double zl,
                                                    double czl2, double c4, double inv2c) {
  if (!(a > 0.0) || !(b > 0.0)) return false;
  const double bb_hi = b * b * (1.0 + kFilterEps);
  if (isinf(zl)) return bb_hi >= c4 * a;
  const double g = (a + czl2) - b * zl * (1.0 + kFilterEps);
  if (g <= 0.0) return true;
  const double v = b * inv2c;
  if (v < zq * (1.0 - kFilterEps) || v > zl * (1.0 + kFilterEps)) return false;
  return bb_hi >= c4 * a;
}

}  // namespace bats
