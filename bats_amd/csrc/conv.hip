// Convolutional LIF layer (valid, stride 1) and 2x2 spike pooling for gfx950.
//
// Forward: the per-sample event lists of the dense layer (sort_events_kernel) are shared;
// conv_solve_kernel gives one wave a small tile of output positions and channels, keeps the
// events inside the tile's joint receptive field with a ballot, and runs the dense solver's
// per-event arithmetic (clamp filter, deferred closed form) on them, each lane using only the
// events inside its own field (compute_spike_times_conv.cu:155-173).
//
// Backward: post_prep_kernel of the dense layer prepares the per-spike suffix sums; then
// conv_wgrad_kernel (one block per filter tap and sample, lanes over output positions, block
// reduction) and conv_pre_err_kernel (lanes over input neurons, loop over the output neurons
// whose field contains them) -- the gradient regrouping of lif_backward.hip applied to the
// receptive-field structure of compute_weights_gradient_conv.cu and
// propagate_errors_to_pre_spikes_conv.cu.
//
// Pooling: pool_forward_kernel merges the four children's ascending lists per output neuron
// (aggregate_spikes_conv.cu + the argsort of its wrapper), recording the source slot of every
// merged slot; pool_backward_kernel scatters errors back (bats/Layers/PoolingLayer.py:55-66).
#include "solver.hpp"

#include <algorithm>

namespace bats {

struct ConvGeom {
  int X, Y, Ci;     // input
  int Xo, Yo, Co;   // output
  int fx, fy;       // filter
};

// [Co, fx, fy, Ci] -> [fx, fy, Ci, Co]: consecutive output channels (consecutive lanes)
// read consecutive weights.
__global__ void conv_filter_transpose_kernel(const float *__restrict__ W, float *__restrict__ WT,
                                             int Co, int taps) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Co * taps) return;
  const int c = e / taps, r = e - c * taps;
  WT[(size_t)r * Co + c] = W[e];
}

// ---------------------------------------------------------------- forward solver
// One wave per (sample, tile of PX x PY output positions, group of CW output channels);
// lane = (position in the tile, channel). The sample's time-sorted event list is read 64
// pre-neuron indices at a time and a ballot keeps the events inside the tile's joint
// receptive field; each kept event is processed by the whole wave with wave-uniform event
// data (scalar loads), kConvBatch events at a time so that their weight loads overlap.
// A lane uses only the events inside its own receptive field ("own" events) and checks the
// interval (t_last, min(t_u, t_max)] since its previous own event when it reaches the next
// own event u, and (t_last, t_max] at the end: an event outside the field has weight zero and
// leaves a and b unchanged, so this is the decision sequence of the reference's kernel, which
// checks after every event of the sample (compute_spike_times_conv.cu:155-173). The check is
// the dense solver's: the clamp filter, then the closed form in z = exp(t/tau) space with
// the spike times formed after the walk.
constexpr int kConvWaves = 4;
constexpr int kConvBatch = 8;

struct ConvTiling {
  int PX, PY, CW;        // output positions per wave (PX x PY) and channels per wave
  int tilesX, tilesY, cgroups;
  float inv_ci, inv_y;   // 1/Ci and 1/Y (float32) for the pre-neuron index decomposition
};

// Interval ends of the conv solver: the lane's previous own event and the event closing the
// interval (null: t_max).
struct ConvEnds {
  const Event *lo_ev;
  const Event *hi_ev;
  float tmax;
  __device__ float lo() const { return lo_ev->t; }
  __device__ float hi() const { return hi_ev ? fminf(hi_ev->t, tmax) : tmax; }
};

// n / d for 0 <= n < 2^24 and d >= 1 with inv = 1.0f / d: the float32 quotient is within one
// of the exact one (exact for d = 1), and one correction step fixes it.
__device__ __forceinline__ int div_small(int n, int d, float inv) {
  int q = (int)((float)n * inv);
  const int r = n - q * d;
  q += (r >= d) ? 1 : 0;
  q -= (r < 0) ? 1 : 0;
  return q;
}

// kAlwaysExact: c <= 0 (no threshold quadratic to filter with), every own interval is solved
// with the reference's loop in t space.
template <bool kAlwaysExact>
__global__ __launch_bounds__(64 * kConvWaves) void conv_solve_kernel(
    const Event *__restrict__ events, const FastEvent *__restrict__ fast,
    const int *__restrict__ evidx, const int *__restrict__ counts, int Pcap, ConvGeom g,
    ConvTiling tl, int units, const float *__restrict__ WT, int K, double c, double dtt,
    double tau, float tmax, int *__restrict__ out_n, float *__restrict__ out_t,
    float *__restrict__ out_a, float *__restrict__ out_x, float *__restrict__ out_z) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int unit = xcd_logical_block(blockIdx.x, gridDim.x) * kConvWaves + wave;
  if (unit >= units) return;  // whole wave; no workgroup barrier follows
  const int lane = threadIdx.x & 63;
  const int tiles = tl.tilesX * tl.tilesY;
  const int cg = unit % tl.cgroups;
  const int tile = (unit / tl.cgroups) % tiles;
  const int b = unit / (tl.cgroups * tiles);
  const int x0 = (tile / tl.tilesY) * tl.PX, y0 = (tile % tl.tilesY) * tl.PY;
  const int pos = lane / tl.CW, cl = lane - pos * tl.CW;
  const int ox = pos / tl.PY, oy = pos - ox * tl.PY;
  const int cc = cg * tl.CW + cl;
  const bool active =
      pos < tl.PX * tl.PY && x0 + ox < g.Xo && y0 + oy < g.Yo && cc < g.Co;
  const int j = active ? ((x0 + ox) * g.Yo + (y0 + oy)) * g.Co + cc : 0;
  const size_t row = ((size_t)b * g.Xo * g.Yo * g.Co + j) * K;
  // weight of an own event: WT[ebase - loff], ebase = ((dx0 fy + dy0) Ci + pc) Co with
  // (dx0, dy0) the event's offset from the tile origin
  const int loff = (ox * g.fy + oy) * g.Ci * g.Co - cc;
  const int ux = tl.PX + g.fx - 1, uy = tl.PY + g.fy - 1;  // joint field of the tile
  const int P = counts[b];
  const Event *ev = events + (size_t)b * Pcap;
  const FastEvent *fv = fast + (size_t)b * Pcap;
  const int *ix = evidx + (size_t)b * Pcap;
  const double inv2c = 0.5 / c;
  const double zmax_hi = fmin(exp((double)tmax / tau), 1e300) * (1.0 + kFilterEps);
  constexpr double kLoToHi = (1.0 + kFilterEps) / (1.0 - kFilterEps);
  // an inactive lane, or one whose buffer is full, carries a = -inf (it fails a > 0); the
  // always-exact form uses `live`
  double a = (kAlwaysExact || active) ? 0.0 : -kInfD, bb = 0.0;
  double zq_last = 0.0;  // zq_lo of the lane's previous own event
  int last_q = 0, n = 0;
  bool live = active, seen = false;
  int qn = (lane < P) ? ix[lane] : -1;
  for (int base = 0; base < P; base += 64) {
    const int qi = qn;
    if (base + 64 < P) qn = (base + 64 + lane < P) ? ix[base + 64 + lane] : -1;
    int ebase = 0, dpk = 0;
    bool inu = false;
    if (qi >= 0) {
      const int r = div_small(qi, g.Ci, tl.inv_ci);
      const int pc = qi - r * g.Ci;
      const int px = div_small(r, g.Y, tl.inv_y);
      const int dx0 = px - x0, dy0 = r - px * g.Y - y0;
      inu = ((unsigned)dx0 < (unsigned)ux) & ((unsigned)dy0 < (unsigned)uy);
      ebase = ((dx0 * g.fy + dy0) * g.Ci + pc) * g.Co;
      dpk = (dx0 << 16) | dy0;
    }
    unsigned long long m = __ballot(inu);
    while (m) {
      int ql[kConvBatch];
#pragma unroll
      for (int k = 0; k < kConvBatch; ++k) {
        ql[k] = m ? (int)__builtin_ctzll(m) : -1;
        m &= m - 1;
      }
      float w[kConvBatch];
      bool own[kConvBatch];
#pragma unroll
      for (int k = 0; k < kConvBatch; ++k) {
        own[k] = false;
        w[k] = 0.f;
        if (ql[k] >= 0) {
          const int eb = __builtin_amdgcn_readlane(ebase, ql[k]);
          const int dp = __builtin_amdgcn_readlane(dpk, ql[k]);
          own[k] = active & ((unsigned)((dp >> 16) - ox) < (unsigned)g.fx) &
                   ((unsigned)((dp & 0xffff) - oy) < (unsigned)g.fy);
          if (own[k]) w[k] = WT[eb - loff];
        }
      }
#pragma unroll
      for (int k = 0; k < kConvBatch; ++k) {
        if (ql[k] < 0) break;
        const int q = base + ql[k];
        const FastEvent f = fv[q];
        const double zl_hi = fmin(f.zq_lo * kLoToHi, zmax_hi);
        if (kAlwaysExact) {
          if (own[k] & live & seen) {
            const SolveState r =
                solve_crossings_ool(a, bb, n, (double)ev[last_q].t,
                                    (double)fminf(ev[q].t, tmax), c, dtt, tau, K, row, out_t,
                                    out_a, out_x, out_z);
            bb = r.b;
            n = r.n;
            live = r.live != 0;
          }
        } else {
          const double zv = fmin(fmax(bb * inv2c, zq_last), zl_hi);
          const double gq = fma(fma(c, zv, -bb), zv, a);
          if (own[k] & (a > 0.0) & (gq <= kFilterEps * a)) {
            const SolveStateZ r = solve_crossings_deferred_ends(
                a, bb, n, zq_last, zl_hi, ConvEnds{ev + last_q, ev + q, tmax}, c, inv2c, dtt,
                tau, K, row, out_t, out_a, out_x, out_z);
            bb = r.b;
            n = r.n;
            if (!r.live) a = -kInfD;
          }
        }
        const double w64 = (double)w[k];
        a = fma(w64, f.es, a);
        bb = fma(w64, f.e, bb);
        if (own[k]) {
          zq_last = f.zq_lo;
          last_q = q;
          seen = true;
        }
      }
    }
  }
  // the last interval, up to t_max
  if (kAlwaysExact) {
    if (live & seen) {
      const SolveState r = solve_crossings_ool(a, bb, n, (double)ev[last_q].t, (double)tmax, c,
                                               dtt, tau, K, row, out_t, out_a, out_x, out_z);
      n = r.n;
    }
  } else {
    const double zv = fmin(fmax(bb * inv2c, zq_last), zmax_hi);
    const double gq = fma(fma(c, zv, -bb), zv, a);
    if ((a > 0.0) & (gq <= kFilterEps * a)) {
      const SolveStateZ r = solve_crossings_deferred_ends(
          a, bb, n, zq_last, zmax_hi, ConvEnds{ev + last_q, nullptr, tmax}, c, inv2c, dtt, tau,
          K, row, out_t, out_a, out_x, out_z);
      n = r.n;
    }
  }
  if (active) {
    if (!kAlwaysExact) finish_spike_times(row, n, tau, out_t, out_z);
    out_n[(size_t)b * g.Xo * g.Yo * g.Co + j] = n;
    for (int k = n; k < K; ++k) out_t[row + k] = kInfF;
  }
}

// ---------------------------------------------------------------- pooling
// One thread per pooled neuron merges the four children's finite spikes (ascending, ties in
// the children's order: aggregate_spikes_conv.cu:13-41 followed by the wrapper's argsort,
// aggregate_spikes_conv.py:24-26) with the four list heads held in registers. The +inf tail
// of every pooled row (child 0's empty slots, then child 1's, ...) is written afterwards by
// the whole block in one coalesced pass over the block's contiguous output rows, so that
// out_src stays a permutation of the children's slots.
constexpr int kPoolThreads = 256;

__global__ __launch_bounds__(kPoolThreads) void pool_forward_kernel(
    int B, int X, int Y, int C, int K, const int *__restrict__ n, const float *__restrict__ t,
    int *__restrict__ out_n, float *__restrict__ out_t, int *__restrict__ out_src) {
  __shared__ int s_cnt[kPoolThreads][5];  // finite count, then the children's counts
  const int Xo = X / 2, Yo = Y / 2;
  const int No = Xo * Yo * C;
  const int N = X * Y * C;
  const int64_t total = (int64_t)B * No;
  const int64_t g0 = (int64_t)blockIdx.x * kPoolThreads;
  const int64_t gid = g0 + threadIdx.x;
  if (gid < total) {
    const int b = (int)(gid / No), jo = (int)(gid - (int64_t)b * No);
    const int x = jo / (Yo * C), r = jo - x * (Yo * C), y = r / C, c = r - y * C;
    size_t base[4];
    int nd[4];
    int cnt = 0;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const int i = ((2 * x + (d >> 1)) * Y + (2 * y + (d & 1))) * C + c;
      base[d] = ((size_t)b * N + i) * K;
      nd[d] = min(max(n[(size_t)b * N + i], 0), K);
      cnt += nd[d];
    }
    out_n[gid] = cnt;
    s_cnt[threadIdx.x][0] = cnt;
#pragma unroll
    for (int d = 0; d < 4; ++d) s_cnt[threadIdx.x][1 + d] = nd[d];
    int p[4] = {0, 0, 0, 0};
    float head[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) head[d] = nd[d] > 0 ? t[base[d]] : kInfF;
    const size_t obase = (size_t)gid * 4 * K;
    for (int s = 0; s < cnt; ++s) {
      int best = 0;
      float bv = head[0];
      bool live = p[0] < nd[0];
#pragma unroll
      for (int d = 1; d < 4; ++d) {
        const bool ok = p[d] < nd[d];
        if (ok && (!live || head[d] < bv)) {
          best = d;
          bv = head[d];
          live = true;
        }
      }
      out_t[obase + s] = bv;
      size_t srcs = 0;
#pragma unroll
      for (int d = 0; d < 4; ++d)
        if (d == best) {
          srcs = base[d] + p[d];
          ++p[d];
          head[d] = p[d] < nd[d] ? t[base[d] + p[d]] : kInfF;
        }
      out_src[obase + s] = (int)srcs;
    }
  }
  __syncthreads();
  // the +inf tails of the block's rows, coalesced
  const int rows = (int)min<int64_t>(kPoolThreads, total - g0);
  const int slots = 4 * K;
  for (int e = threadIdx.x; e < rows * slots; e += kPoolThreads) {
    const int rr = e / slots, s = e - rr * slots;
    const int cnt = s_cnt[rr][0];
    if (s < cnt) continue;
    const int64_t g = g0 + rr;
    const int b = (int)(g / No), jo = (int)(g - (int64_t)b * No);
    const int x = jo / (Yo * C), r = jo - x * (Yo * C), y = r / C, c = r - y * C;
    int q = s - cnt;
    int d = 0;
    while (d < 3 && q >= K - s_cnt[rr][1 + d]) {
      q -= K - s_cnt[rr][1 + d];
      ++d;
    }
    const int i = ((2 * x + (d >> 1)) * Y + (2 * y + (d & 1))) * C + c;
    const size_t o = (size_t)g * slots + s;
    out_t[o] = kInfF;
    out_src[o] = (int)(((size_t)b * N + i) * K + s_cnt[rr][1 + d] + q);
  }
}

__global__ void pool_backward_kernel(int64_t total, const float *__restrict__ err,
                                     const int *__restrict__ src, float *__restrict__ pre_err) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < total) pre_err[src[e]] = err[e];
}

// ---------------------------------------------------------------- backward
__device__ __forceinline__ void lookup_row(const float *__restrict__ tr,
                                           const double *__restrict__ s1,
                                           const double *__restrict__ s2, int nj, float tp,
                                           double &v1, double &v2) {
  v1 = 0.0;
  v2 = 0.0;
  for (int k = nj - 1; k >= 0; --k) {
    if (tr[k] > tp) {
      v1 = s1[k];
      v2 = s2[k];
    }
  }
}

// One block per (filter tap, sample); lanes over output positions.
__global__ void conv_wgrad_kernel(ConvGeom g, int K, int K_pre, const float *__restrict__ pre_t,
                                  const float *__restrict__ pre_es,
                                  const float *__restrict__ pre_e,
                                  const float *__restrict__ post_t, const int *__restrict__ n_eff,
                                  const double *__restrict__ S1, const double *__restrict__ S2,
                                  double *__restrict__ per_sample) {
  __shared__ double s_red[16];
  const int taps = g.fx * g.fy * g.Ci;
  const int tap = blockIdx.x;  // over Co * taps, layout [Co, fx, fy, Ci]
  const int b = blockIdx.y;
  const int cc = tap / taps;
  const int r = tap - cc * taps;
  const int dx = r / (g.fy * g.Ci);
  const int r2 = r - dx * (g.fy * g.Ci);
  const int dy = r2 / g.Ci;
  const int ci = r2 - dy * g.Ci;
  const int Npost = g.Xo * g.Yo * g.Co;
  const int Npre = g.X * g.Y * g.Ci;
  double acc = 0.0;
  for (int p = threadIdx.x; p < g.Xo * g.Yo; p += blockDim.x) {
    const int x = p / g.Yo, y = p - (p / g.Yo) * g.Yo;
    const int j = (x * g.Yo + y) * g.Co + cc;
    const size_t prow = (size_t)b * Npost + j;
    const int nj = n_eff[prow];
    if (nj == 0) continue;
    const int i = ((x + dx) * g.Y + (y + dy)) * g.Ci + ci;
    const size_t pbase = ((size_t)b * Npre + i) * K_pre;
    for (int m = 0; m < K_pre; ++m) {
      const float tp = pre_t[pbase + m];
      if (!isfinite(tp)) break;
      double v1, v2;
      lookup_row(post_t + prow * K, S1 + prow * K, S2 + prow * K, nj, tp, v1, v2);
      acc = fma((double)pre_es[pbase + m], v1, acc);
      acc = fma(-(double)pre_e[pbase + m], v2, acc);
    }
  }
  // block reduction in a fixed order
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) s_red[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) s += s_red[w];
    per_sample[(size_t)b * gridDim.x + tap] = s;
  }
}

__global__ void conv_wgrad_finish_kernel(int B, int taps_all, int mode,
                                         const double *__restrict__ per_sample,
                                         float *__restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (mode == 1) {
    if (e < B * taps_all) grad[e] = (float)per_sample[e];
    return;
  }
  if (e >= taps_all) return;
  double s = 0.0;
  for (int b = 0; b < B; ++b) s += per_sample[(size_t)b * taps_all + e];
  grad[e] = (float)(mode == 2 ? s : s / (double)B);
}

// Lanes over the input neurons of one sample; every output neuron whose field holds the
// lane's neuron contributes W * (V1 es/tau_s - V2 e/tau) to each of its spikes.
__global__ void conv_pre_err_kernel(ConvGeom g, int K, int K_pre, const float *__restrict__ pre_t,
                                    const float *__restrict__ pre_es,
                                    const float *__restrict__ pre_e,
                                    const float *__restrict__ W, const float *__restrict__ post_t,
                                    const int *__restrict__ n_eff, const double *__restrict__ S1,
                                    const double *__restrict__ S2, double tau_s, double tau,
                                    float *__restrict__ pre_err) {
  const int Npre = g.X * g.Y * g.Ci;
  const int Npost = g.Xo * g.Yo * g.Co;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= Npre) return;
  const int px = i / (g.Y * g.Ci), r = i - px * (g.Y * g.Ci), py = r / g.Ci, pc = r - py * g.Ci;
  const size_t pbase = ((size_t)b * Npre + i) * K_pre;
  for (int m = 0; m < K_pre; ++m) {
    const float tp = pre_t[pbase + m];
    if (!isfinite(tp)) {
      pre_err[pbase + m] = 0.f;
      continue;
    }
    const double es = (double)pre_es[pbase + m], e = (double)pre_e[pbase + m];
    double acc = 0.0;
    for (int dx = 0; dx < g.fx; ++dx) {
      const int x = px - dx;
      if (x < 0 || x >= g.Xo) continue;
      for (int dy = 0; dy < g.fy; ++dy) {
        const int y = py - dy;
        if (y < 0 || y >= g.Yo) continue;
        for (int cc = 0; cc < g.Co; ++cc) {
          const int j = (x * g.Yo + y) * g.Co + cc;
          const size_t prow = (size_t)b * Npost + j;
          const int nj = n_eff[prow];
          if (nj == 0) continue;
          double v1, v2;
          lookup_row(post_t + prow * K, S1 + prow * K, S2 + prow * K, nj, tp, v1, v2);
          const double w = (double)W[(((size_t)cc * g.fx + dx) * g.fy + dy) * g.Ci + pc];
          acc += w * ((v1 * es) / tau_s - (v2 * e) / tau);
        }
      }
    }
    pre_err[pbase + m] = (float)acc;
  }
}

// defined in lif_backward.hip
int post_prep(hipStream_t st, int BJ, int K, const int *n, const float *t, const float *a,
              const float *x, const float *z, float *errors, float c, float tau, float dtt,
              int *n_eff, double *S1, double *S2);

// Wave tiling of the conv solver. A wave reads the whole event list of its sample once (cost
// ~ 1/64 per event) and processes the events inside its joint field (cost ~ 1 per event), so
// the work of a layer is ~ waves x (1/64 + joint field / input plane); the tiling that
// minimises it is taken (more positions per wave share the scan, but widen the joint field).
static ConvTiling conv_tiling(const ConvGeom &g) {
  ConvTiling best{};
  double best_cost = 1e300;
  const int g0 = (g.Co + 63) / 64;
  for (int groups = g0; groups <= std::min(g.Co, 4 * g0); ++groups) {
    const int CW = (g.Co + groups - 1) / groups;
    if (CW > 64 || (groups > g0 && (g.Co + CW - 1) / CW != groups)) continue;
    const int npos = 64 / CW;
    for (int PX = 1; PX <= std::min(npos, g.Xo); ++PX) {
      const int PY = std::min(npos / PX, g.Yo);
      const int tx = (g.Xo + PX - 1) / PX, ty = (g.Yo + PY - 1) / PY;
      const double waves = (double)groups * tx * ty;
      const double field = (double)(PX + g.fx - 1) * (PY + g.fy - 1) / ((double)g.X * g.Y);
      const double cost = waves * (1.0 / 64.0 + field);
      if (cost < best_cost) {
        best_cost = cost;
        best.PX = PX;
        best.PY = PY;
        best.CW = CW;
        best.tilesX = tx;
        best.tilesY = ty;
        best.cgroups = groups;
      }
    }
  }
  best.inv_ci = 1.0f / (float)g.Ci;
  best.inv_y = 1.0f / (float)g.Y;
  return best;
}

}  // namespace bats

using namespace bats;

static bool conv_geom(const int *pre_shape, const int *filters_shape, ConvGeom &g) {
  g.X = pre_shape[0];
  g.Y = pre_shape[1];
  g.Ci = pre_shape[2];
  g.fx = filters_shape[0];
  g.fy = filters_shape[1];
  g.Co = filters_shape[2];
  g.Xo = g.X - g.fx + 1;
  g.Yo = g.Y - g.fy + 1;
  return g.X > 0 && g.Y > 0 && g.Ci > 0 && g.fx > 0 && g.fy > 0 && g.Co > 0 && g.Xo > 0 &&
         g.Yo > 0 && g.Co <= 1024 && (long long)g.X * g.Y * g.Ci < (1LL << 24) &&
         (long long)(g.fx + 64) * (g.fy + 64) * g.Ci * g.Co < INT32_MAX;
}

extern "C" size_t bats_conv_lif_forward_workspace_size(int B, const int *pre_shape, int K_pre,
                                                       const int *filters_shape) {
  ConvGeom g;
  if (B < 0 || !pre_shape || !filters_shape || !conv_geom(pre_shape, filters_shape, g)) return 0;
  const int S = g.X * g.Y * g.Ci * K_pre;
  return align_up(sizeof(float) * (size_t)g.Co * g.fx * g.fy * g.Ci, 256) +
         event_lists_bytes(B, S);
}

extern "C" int bats_conv_lif_forward(void *stream, int B, const int *pre_shape, int K_pre,
                                     const int *filters_shape, int K_post,
                                     const float *pre_times, const float *weights, float tau_s,
                                     float tau, float c, float delta_theta_tau,
                                     float max_simulation, int *out_n, float *out_t,
                                     float *out_a, float *out_x, float *out_z,
                                     float *pre_exp_tau_s, float *pre_exp_tau, void *workspace,
                                     size_t workspace_bytes) {
  ConvGeom g;
  BATS_REQUIRE(pre_shape && filters_shape && conv_geom(pre_shape, filters_shape, g),
               "bad conv shapes (pre_shape = {X, Y, C_in}, filters_shape = {fx, fy, C_out}; "
               "C_out <= 1024, X * Y * C_in < 2^24)");
  BATS_REQUIRE(B >= 0 && K_pre >= 0 && K_post >= 1, "bad sizes");
  BATS_REQUIRE(B <= 65535, "batch above 65535");
  if (B == 0) return 0;
  BATS_REQUIRE(out_n && out_t && out_a && out_x && out_z, "null output pointer");
  const size_t need = bats_conv_lif_forward_workspace_size(B, pre_shape, K_pre, filters_shape);
  BATS_REQUIRE(workspace && workspace_bytes >= need, "workspace too small");
  hipStream_t st = as_stream(stream);
  const int taps = g.fx * g.fy * g.Ci;
  char *ws = static_cast<char *>(workspace);
  float *WT = reinterpret_cast<float *>(ws);
  void *ev_ws = ws + align_up(sizeof(float) * (size_t)g.Co * taps, 256);
  const int S = g.X * g.Y * g.Ci * K_pre;
  if (S > 0) BATS_REQUIRE(pre_times && weights && pre_exp_tau_s && pre_exp_tau, "null input");
  hipLaunchKernelGGL(conv_filter_transpose_kernel, dim3((g.Co * taps + 255) / 256), dim3(256), 0,
                     st, weights, WT, g.Co, taps);
  if (int r = check_launch("conv_filter_transpose_kernel")) return r;
  EventLists el;
  if (int r = build_event_lists(st, B, S, K_pre, pre_times, tau_s, tau, c, max_simulation,
                                pre_exp_tau_s, pre_exp_tau, ev_ws, &el))
    return r;
  const ConvTiling tl = conv_tiling(g);
  const long long units = (long long)B * tl.tilesX * tl.tilesY * tl.cgroups;
  BATS_REQUIRE(units <= (long long)INT32_MAX - 64, "conv layer too large");
  const int blocks = (int)((units + kConvWaves - 1) / kConvWaves);
  if (c > 0.f) {
    BATS_TIMED("conv_solve_kernel", st,
               hipLaunchKernelGGL(conv_solve_kernel<false>, dim3(blocks), dim3(64 * kConvWaves),
                                  0, st, el.events, el.fast, el.idx, el.counts, el.Pcap, g, tl,
                                  (int)units, WT, K_post, (double)c, (double)delta_theta_tau,
                                  (double)tau, max_simulation, out_n, out_t, out_a, out_x,
                                  out_z));
  } else {
    BATS_TIMED("conv_solve_kernel", st,
               hipLaunchKernelGGL(conv_solve_kernel<true>, dim3(blocks), dim3(64 * kConvWaves),
                                  0, st, el.events, el.fast, el.idx, el.counts, el.Pcap, g, tl,
                                  (int)units, WT, K_post, (double)c, (double)delta_theta_tau,
                                  (double)tau, max_simulation, out_n, out_t, out_a, out_x,
                                  out_z));
  }
  return check_launch("conv_solve_kernel");
}

extern "C" size_t bats_conv_lif_backward_workspace_size(int B, const int *pre_shape,
                                                        const int *filters_shape, int K_post) {
  ConvGeom g;
  if (B < 0 || !pre_shape || !filters_shape || !conv_geom(pre_shape, filters_shape, g)) return 0;
  const size_t BJ = (size_t)B * g.Xo * g.Yo * g.Co;
  size_t off = 0;
  off = align_up(off + sizeof(int) * BJ, 256);
  off = align_up(off + sizeof(double) * BJ * K_post, 256);
  off = align_up(off + sizeof(double) * BJ * K_post, 256);
  off = align_up(off + sizeof(double) * (size_t)B * g.Co * g.fx * g.fy * g.Ci, 256);
  return off;
}

extern "C" int bats_conv_lif_backward(void *stream, int B, const int *pre_shape, int K_pre,
                                      const int *filters_shape, int K_post, const int *n,
                                      const float *t, const float *a, const float *x,
                                      const float *z, float *errors, const float *pre_times,
                                      const float *pre_exp_tau_s, const float *pre_exp_tau,
                                      const float *weights, float tau_s, float tau, float c,
                                      float delta_theta_tau, float *grad, int grad_mode,
                                      float *pre_errors, void *workspace,
                                      size_t workspace_bytes) {
  ConvGeom g;
  BATS_REQUIRE(pre_shape && filters_shape && conv_geom(pre_shape, filters_shape, g),
               "bad conv shapes");
  BATS_REQUIRE(B >= 0 && K_pre >= 0 && K_post >= 1, "bad sizes");
  BATS_REQUIRE(grad_mode >= 0 && grad_mode <= 2, "grad_mode must be 0, 1 or 2");
  BATS_REQUIRE(B <= 65535, "batch above 65535");
  if (B == 0) return 0;
  const size_t need = bats_conv_lif_backward_workspace_size(B, pre_shape, filters_shape, K_post);
  BATS_REQUIRE(workspace && workspace_bytes >= need, "workspace too small");
  BATS_REQUIRE(n && t && a && x && z && errors && grad, "null pointer");
  hipStream_t st = as_stream(stream);
  const size_t BJ = (size_t)B * g.Xo * g.Yo * g.Co;
  char *ws = static_cast<char *>(workspace);
  size_t off = 0;
  int *n_eff = reinterpret_cast<int *>(ws + off);
  off = align_up(off + sizeof(int) * BJ, 256);
  double *S1 = reinterpret_cast<double *>(ws + off);
  off = align_up(off + sizeof(double) * BJ * K_post, 256);
  double *S2 = reinterpret_cast<double *>(ws + off);
  off = align_up(off + sizeof(double) * BJ * K_post, 256);
  double *per_sample = reinterpret_cast<double *>(ws + off);
  if (int r = post_prep(st, (int)BJ, K_post, n, t, a, x, z, errors, c, tau, delta_theta_tau,
                        n_eff, S1, S2))
    return r;
  const int taps_all = g.Co * g.fx * g.fy * g.Ci;
  if (K_pre > 0) BATS_REQUIRE(pre_times && pre_exp_tau_s && pre_exp_tau, "null pre pointer");
  const int threads = std::min(1024, std::max(64, (g.Xo * g.Yo + 63) / 64 * 64));
  BATS_TIMED("conv_wgrad_kernel", st,
             hipLaunchKernelGGL(conv_wgrad_kernel, dim3(taps_all, B), dim3(threads), 0, st, g,
                                K_post, K_pre, pre_times, pre_exp_tau_s, pre_exp_tau, t, n_eff,
                                S1, S2, per_sample));
  if (int r = check_launch("conv_wgrad_kernel")) return r;
  const int fin = grad_mode == 1 ? B * taps_all : taps_all;
  hipLaunchKernelGGL(conv_wgrad_finish_kernel, dim3((fin + 255) / 256), dim3(256), 0, st, B,
                     taps_all, grad_mode, per_sample, grad);
  if (int r = check_launch("conv_wgrad_finish_kernel")) return r;
  if (pre_errors && K_pre > 0) {
    BATS_REQUIRE(weights, "null weights");
    const int Npre = g.X * g.Y * g.Ci;
    BATS_TIMED("conv_pre_err_kernel", st,
               hipLaunchKernelGGL(conv_pre_err_kernel, dim3((Npre + 255) / 256, B), dim3(256), 0,
                                  st, g, K_post, K_pre, pre_times, pre_exp_tau_s, pre_exp_tau,
                                  weights, t, n_eff, S1, S2, (double)tau_s, (double)tau,
                                  pre_errors));
    if (int r = check_launch("conv_pre_err_kernel")) return r;
  }
  return 0;
}

extern "C" int bats_pool2x2_forward(void *stream, int B, const int *pre_shape, int K,
                                    const int *n, const float *t, int *out_n, float *out_t,
                                    int *out_src) {
  BATS_REQUIRE(pre_shape && B >= 0 && K >= 0, "bad sizes");
  const int X = pre_shape[0], Y = pre_shape[1], C = pre_shape[2];
  BATS_REQUIRE(X >= 2 && Y >= 2 && C >= 1, "pooling needs at least 2x2 positions");
  const long long total = (long long)B * (X / 2) * (Y / 2) * C;
  BATS_REQUIRE((double)B * X * Y * C * K < 2147483647.0, "pooling input above 2^31 slots");
  if (total == 0 || K == 0) return 0;
  BATS_REQUIRE(n && t && out_n && out_t && out_src, "null pointer");
  hipLaunchKernelGGL(pool_forward_kernel, dim3((unsigned)((total + kPoolThreads - 1) / kPoolThreads)), dim3(kPoolThreads), 0,
                     as_stream(stream), B, X, Y, C, K, n, t, out_n, out_t, out_src);
  return check_launch("pool_forward_kernel");
}

extern "C" int bats_pool2x2_backward(void *stream, int B, const int *pre_shape, int K,
                                     const float *errors, const int *src, float *pre_errors) {
  BATS_REQUIRE(pre_shape && B >= 0 && K >= 0, "bad sizes");
  const int X = pre_shape[0], Y = pre_shape[1], C = pre_shape[2];
  const int64_t npre = (int64_t)B * X * Y * C * K;
  const int64_t npost = (int64_t)B * (X / 2) * (Y / 2) * C * 4 * K;
  if (npre == 0) return 0;
  BATS_REQUIRE(errors && src && pre_errors, "null pointer");
  hipStream_t st = as_stream(stream);
  // slots outside every pooling window (odd input sizes) get zero
  hipLaunchKernelGGL(fill_kernel<float>, dim3((unsigned)std::min<int64_t>((npre + 255) / 256, 65535)),
                     dim3(256), 0, st, pre_errors, npre, 0.f);
  if (int r = check_launch("fill pre errors")) return r;
  if (npost == 0) return 0;
  hipLaunchKernelGGL(pool_backward_kernel, dim3((unsigned)((npost + 255) / 256)), dim3(256), 0, st,
                     npost, errors, src, pre_errors);
  return check_launch("pool_backward_kernel");
}
