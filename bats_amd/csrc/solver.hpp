// Device-side pieces of the exact spike-time solver shared by the dense and conv layers.
#pragma once

#include "common.hpp"

namespace bats {

// ------------------------------------------------------------------ crossing filter
// Necessary condition (with a relative margin) for the reference's smaller root
// z = 2a/(b + sqrt(b^2 - 4ac)) to be accepted in (t_q, min(t_next, t_max)]:
// a > 0 and b > 0, and either g(zl) = c zl^2 - b zl + a <= 0 (one root inside) or the
// vertex b/(2c) of g lies in [zq, zl] with a non-negative discriminant (two roots inside).
// Evaluated without branches (one divergent branch per event is left to the caller).
__device__ __forceinline__ bool crossing_filter(double a, double b, const Event &e, double c4) {
  const double g = fma(-b, e.zl1, a + e.czl2);
  const double dd = fma(b, b * (1.0 + kFilterEps), -(c4 * a));
  return (a > 0.0) & (b > 0.0) & ((g <= 0.0) | ((b >= e.blo) & (b <= e.bhi) & (dd >= 0.0)));
}

// The same necessary condition from the minimum of the convex quadratic: with a > 0 and
// c > 0 the accepted (smaller) root lies in [zq, zl] only if g(z) = c z^2 - b z + a is <= 0
// somewhere there, i.e. at the vertex b/(2c) clamped to the interval. The interval is widened
// by kFilterEps and g may exceed 0 by kFilterEps * a, far more than the rounding of g near a
// root (a few ulps of 2a there) or of the computed root. A larger root at the interval start
// also passes; the exact loop rejects it.
// Clamp of a lane value to wave-uniform bounds held in SGPRs. Written as the two
// instructions: fmin/fmax would first re-quiet the loaded bounds (two more float64 ops per
// event) to honour signalling NaNs, which the bounds never are (finite or +inf).
__device__ __forceinline__ double clamp_uniform(double x, double lo, double hi) {
  double r;
  asm("v_max_f64 %0, %1, %2\n\tv_min_f64 %0, %0, %3" : "=&v"(r) : "v"(x), "s"(lo), "s"(hi));
  return r;
}

__device__ __forceinline__ bool clamp_filter(double a, double b, const FastEvent &f, double c,
                                             double inv2c) {
  const double zv = clamp_uniform(b * inv2c, f.zq_lo, f.zl_hi);
  const double g = fma(fma(c, zv, -b), zv, a);
  // (a > 0) and (g <= eps a) as one comparison, max(g - eps a, -a) <= 0; a = 0 also passes,
  // and the closed form rejects it (its root is 0)
  return fmax(fma(-kFilterEps, a, g), -a) <= 0.0;
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


// ------------------------------------------------------------------ deferred-log form
// The closed-form loop of solve_crossings with the same accept/reject decisions, made in
// z = exp(t/tau) space: t = tau*log(z) is monotone in z, so a root z clearly above the lower
// end z_lo (clearly below the upper end z_hi) is above last (below tl) in t as well. "Clearly"
// is a relative band kZBand around the end; a root inside the band is decided with the
// float64 logarithm exactly as solve_crossings does (the band, 1e-12, is far above the few
// ulps by which tau*log(z) and the precomputed ends can round for any t < 700 tau).
//
// An accepted spike's a and x are stored at once; its time is left for
// finish_spike_times: the float64 z is kept across the float32 slots of t and z
// (high word in t, low word in z), so the logarithm of all spikes of a lane runs once,
// after its event walk, instead of once per candidate interval inside it.
constexpr double kZBand = 1e-12;

// Square root and quotient for the candidate root: the hardware reciprocal (square root)
// estimate refined by two Newton steps (each about doubles the correct bits), without the
// rescaling that guards denormal and overflowing operands (the solver's discriminants and
// denominators are normal numbers). Results are within an ulp or two of the correctly
// rounded value, far inside kZBand, where every accept/reject decision is made.
__device__ __forceinline__ double root_sqrt(double d) {  // d >= 0
  if (d == 0.0) return 0.0;
  const double y = __builtin_amdgcn_rsq(d);
  double x = d * y, h = 0.5 * y;
  const double r = fma(-x, h, 0.5);
  x = fma(x, r, x);
  h = fma(h, r, h);
  const double e = fma(-x, x, d);
  return fma(e, h, x);
}

__device__ __forceinline__ double root_div(double num, double den) {
  double r = __builtin_amdgcn_rcp(den);
  double e = fma(-den, r, 1.0);
  r = fma(r, e, r);
  e = fma(-den, r, 1.0);
  r = fma(r, e, r);
  const double q = num * r;
  return fma(fma(-den, q, num), r, q);
}

struct SolveStateZ {
  double b;
  int n;
  int live;
};

__device__ __forceinline__ void stash_z(size_t slot, double z, float *__restrict__ out_t,
                                        float *__restrict__ out_z) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(z);
  out_t[slot] = __uint_as_float((unsigned)(bits >> 32));
  out_z[slot] = __uint_as_float((unsigned)bits);
}

// Where an accepted spike's float64 z waits for its logarithm. GlobalStash parks it in the
// spike's own float32 t and z slots; LaneStash keeps a lane's first kLaneStash spikes in LDS
// (one float64 per slot, [slot][lane]) and parks the later ones like GlobalStash.
// An accepted spike's a and x go straight to its slots (streamed past the caches).
__device__ __forceinline__ void store_ax(size_t slot, double a, double x, float *out_a,
                                         float *out_x) {
  __builtin_nontemporal_store((float)a, out_a + slot);
  __builtin_nontemporal_store((float)x, out_x + slot);
}

struct GlobalStash {
  float *out_t, *out_z;
  __device__ __forceinline__ void operator()(size_t slot, int, double z) const {
    stash_z(slot, z, out_t, out_z);
  }
  __device__ __forceinline__ void ax(size_t slot, double a, double x, float *out_a,
                                     float *out_x) const {
    store_ax(slot, a, x, out_a, out_x);
  }
};

constexpr int kLaneStash = 4;
struct LaneStash {
  double *lds;  // [kLaneStash][stride]
  int lane, stride;
  float *out_t, *out_z;
  __device__ __forceinline__ void operator()(size_t slot, int k, double z) const {
    if (k < kLaneStash)
      lds[k * stride + lane] = z;
    else
      stash_z(slot, z, out_t, out_z);
  }
  __device__ __forceinline__ void ax(size_t slot, double a, double x, float *out_a,
                                     float *out_x) const {
    store_ax(slot, a, x, out_a, out_x);
  }
};

// A segment lane's one spike (layers with one spike per neuron): z, a and x in its LDS slots
// until the lanes of the neuron's other segments are known.
struct SegStash {
  double *z;
  float *a, *x;
  __device__ __forceinline__ void operator()(size_t, int, double zv) const { *z = zv; }
  __device__ __forceinline__ void ax(size_t, double av, double xv, float *, float *) const {
    *a = (float)av;
    *x = (float)xv;
  }
};

__device__ __forceinline__ double tau_log(double tau, double z) { return tau * log(z); }

// The interval is (t_lo, t_hi] with z_lo = exp(t_lo/tau), z_hi = exp(t_hi/tau); t_lo is an
// input event time, t_hi the next event time or t_max. The ends come in widened by
// kFilterEps (as the filter uses them) and are narrowed back here to within a few ulps,
// far inside kZBand; the event record is read only for a root inside the band. After an
// accepted spike the clamp filter is applied to (z, t_hi] with the reset b, so that the
// closed form runs again only when another root can be accepted.
// The interval's end times (t_lo, t_hi) are read from `ends` (lo(), hi()) only for a root
// inside the band: EventEnds reads them from the event record of the dense event walk.
struct EventEnds {
  const Event *ev;
  __device__ float lo() const { return ev->t; }
  __device__ float hi() const { return ev->tl; }
};

template <class Ends, class Stash>
__device__ __forceinline__ SolveStateZ solve_crossings_deferred_stash(
    double a, double b, int n, double zq_lo, double zl_hi, const Ends ends, double c,
    double inv2c, double dtt, double tau, int K, size_t row, float *out_a, float *out_x,
    const Stash &stash) {
  const double z_lo = zq_lo * (1.0 / (1.0 - kFilterEps));
  const double z_hi = zl_hi * (1.0 / (1.0 + kFilterEps));
  const bool hi_finite = z_hi < 1e299;  // an end of +inf is stored as 1e300
  SolveStateZ r;
  r.live = 1;
  double zlast = z_lo;
  bool last_is_event = true;
  while (true) {
    // the candidate root, formed without branches (a negative discriminant or a zero
    // denominator makes it invalid below, exactly where the reference's loop stops)
    const double disc = b * b - (4.0 * a) * c;
    const double x = root_sqrt(fmax(disc, 0.0));
    const double den = x + b;
    const double z = root_div(2.0 * a, den);
    const bool valid = (disc >= 0.0) & (den != 0.0) & (z >= 0.0);
    // t > last and t <= tl, decided in z space; only a root within kZBand of an end (rare)
    // takes the branch that compares float64 logarithms
    const bool lo_clear = z > zlast * (1.0 + kZBand), lo_rej = z < zlast * (1.0 - kZBand);
    const bool hi_clear = z < z_hi * (1.0 - kZBand);
    const bool hi_rej = hi_finite & (z > z_hi * (1.0 + kZBand));
    bool above = lo_clear, below = hi_clear;
    if (valid & ((!lo_clear & !lo_rej) | (!hi_clear & !hi_rej))) {
      const double tz = tau_log(tau, z);
      if (!lo_clear & !lo_rej) {
        const double last = last_is_event ? (double)ends.lo() : tau_log(tau, zlast);
        above = tz > last;
      }
      if (!hi_clear & !hi_rej) below = !(tz > (double)ends.hi());
    }
    if (!(valid & above & below)) break;
    const size_t slot = row + n;
    stash.ax(slot, a, x, out_a, out_x);
    stash(slot, n, z);
    zlast = z;
    last_is_event = false;
    b = b - dtt * z;
    ++n;
    if (n >= K) {
      r.live = 0;
      break;
    }
    const double zv = fmin(fmax(b * inv2c, z * (1.0 - kFilterEps)), zl_hi);
    if (!(fma(fma(c, zv, -b), zv, a) <= kFilterEps * a)) break;
  }
  r.b = b;
  r.n = n;
  return r;
}

template <class Ends>
__device__ __forceinline__ SolveStateZ solve_crossings_deferred_ends(
    double a, double b, int n, double zq_lo, double zl_hi, const Ends ends, double c,
    double inv2c, double dtt, double tau, int K, size_t row, float *out_t, float *out_a,
    float *out_x, float *out_z) {
  return solve_crossings_deferred_stash(a, b, n, zq_lo, zl_hi, ends, c, inv2c, dtt, tau, K, row,
                                        out_a, out_x, GlobalStash{out_t, out_z});
}

__device__ __forceinline__ SolveStateZ solve_crossings_deferred(
    double a, double b, int n, double zq_lo, double zl_hi, const Event *ev, double c,
    double inv2c, double dtt, double tau, int K, size_t row, float *out_t, float *out_a,
    float *out_x, float *out_z) {
  return solve_crossings_deferred_ends(a, b, n, zq_lo, zl_hi, EventEnds{ev}, c, inv2c, dtt, tau,
                                       K, row, out_t, out_a, out_x, out_z);
}

// Spike times of one lane's stashed spikes k0 .. n-1: t = tau*log(z) and z rounded to float32,
// as solve_crossings stores them. Four slots at a time, their parked words loaded in one round.
__device__ __forceinline__ void finish_spike_times(size_t row, int n, double tau,
                                                   float *__restrict__ out_t,
                                                   float *__restrict__ out_z, int k0 = 0) {
  for (int kb = k0; kb < n; kb += 4) {
    unsigned hi[4], lo[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (kb + u < n) {
        hi[u] = __float_as_uint(out_t[row + kb + u]);
        lo[u] = __float_as_uint(out_z[row + kb + u]);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (kb + u < n) {
        const double z =
            __longlong_as_double((long long)(((unsigned long long)hi[u] << 32) | lo[u]));
        out_t[row + kb + u] = (float)(tau * log(z));
        out_z[row + kb + u] = (float)z;
      }
    }
  }
}

}  // namespace bats
