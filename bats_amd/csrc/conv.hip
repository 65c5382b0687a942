// Convolutional LIF layer (valid, stride 1) and 2x2 spike pooling for gfx950.
//
// Forward: the per-sample event lists of the dense layer (sort_events_kernel) are shared;
// conv_solve_kernel gives one wave a small tile of output positions and channels, keeps the
// events inside the tile's joint receptive field with a ballot, and runs the dense solver's
// per-event arithmetic (clamp filter, deferred closed form) on them, each lane using only the
// events inside its own field (compute_spike_times_conv.cu:155-173).
//
// Backward: row_prep_kernel of the dense layer runs the reset chain and forms the per-spike
// terms d1, d2; conv_rec_kernel lists every output position's spikes (all channels) in time
// order; conv_wgrad_kernel (lane per filter tap, merge walk of the position's list against
// the tap's pre spikes) and conv_pre_err_kernel (lane per pre spike, walk of the lists of the
// output positions whose field holds it) apply the gradient regrouping of lif_backward.hip to
// the receptive-field structure of compute_weights_gradient_conv.cu and
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
constexpr int kConvWaves = 4;       // waves per workgroup of the backward kernels
constexpr int kConvSolveWaves = 8;  // waves per workgroup of the solver (they share its LDS)
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

// A lane value clamped to [lo, hi], written as the two instructions (fmax/fmin would first
// quiet both operands, two more float64 operations; the bounds are never NaN).
__device__ __forceinline__ double clamp_lane2(double x, double lo, double hi) {
  double r;
  asm("v_max_f64 %0, %1, %2\n\tv_min_f64 %0, %0, %3" : "=&v"(r) : "v"(x), "v"(lo), "v"(hi));
  return r;
}

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
// The decomposition of one pre-neuron index of the chunk: in the tile's joint field or not,
// the uniform part ebase of its weight index and its packed offset (dx0 << 16 | dy0) from
// the tile origin.
struct ConvChunkEv {
  int ebase, dpk;
  bool inu;
};

__device__ __forceinline__ ConvChunkEv conv_decompose(int qi, const ConvGeom &g,
                                                      const ConvTiling &tl, int x0, int y0,
                                                      int ux, int uy) {
  ConvChunkEv r{0, 0, false};
  if (qi >= 0) {
    const int rr = div_small(qi, g.Ci, tl.inv_ci);
    const int pc = qi - rr * g.Ci;
    const int px = div_small(rr, g.Y, tl.inv_y);
    const int dx0 = px - x0, dy0 = rr - px * g.Y - y0;
    r.inu = ((unsigned)dx0 < (unsigned)ux) & ((unsigned)dy0 < (unsigned)uy);
    r.ebase = ((dx0 * g.fy + dy0) * g.Ci + pc) * g.Co;
    r.dpk = (dx0 << 16) | dy0;
  }
  return r;
}

// The event data of an in-field event, as staged for the wave: es, e, the lower end zq_lo of
// the interval that the event opens and the upper end zl_hi of the interval that it closes.
struct ConvStageEv {
  double zq, zh;
  float es, e;  // float32 values (the event records hold them widened to float64)
};

__device__ __forceinline__ ConvStageEv conv_lane_event(bool inu, const FastEvent *fv, int q,
                                                       double zmax_hi) {
  constexpr double kLoToHi = (1.0 + kFilterEps) / (1.0 - kFilterEps);
  ConvStageEv r{0.0, 0.0, 0.f, 0.f};
  if (inu) {
    const FastEvent f = fv[q];
    r.es = (float)f.es;
    r.e = (float)f.e;
    r.zq = f.zq_lo;
    r.zh = fmin(f.zq_lo * kLoToHi, zmax_hi);
  }
  return r;
}

// Per-wave LDS stage of one 64-event chunk: its in-field events in order (slot = rank among
// the chunk's in-field events), with their weight-index base, field offset and list index.
struct ConvWaveStage {
  ConvStageEv ev[64];
  int eb[64];
  int dp[64];
  int q[64];
};

__host__ inline size_t conv_solve_lds_bytes(size_t wbytes, bool lds_w) {
  return (lds_w ? (wbytes + 15) / 16 * 16 + 16 : 0) + sizeof(ConvWaveStage) * kConvSolveWaves;
}

// kAlwaysExact: c <= 0 (no threshold quadratic to filter with), every own interval is solved
// with the reference's loop in t space. kLdsW: the transposed filter bank is staged in the
// workgroup's LDS, followed by one zero (the weight of an event outside a lane's field).
// kSingle: one output position per wave, so that every in-field event is in the field of
// every lane (no per-lane field test).
template <bool kAlwaysExact, bool kLdsW, bool kSingle>
__global__ __launch_bounds__(64 * kConvSolveWaves) void conv_solve_kernel(
    const Event *__restrict__ events, const FastEvent *__restrict__ fast,
    const int *__restrict__ evidx, const int *__restrict__ counts, int Pcap, ConvGeom g,
    ConvTiling tl, int units, const float *__restrict__ WT_g, int K, double c, double dtt,
    double tau, float tmax, int *__restrict__ out_n, float *__restrict__ out_t,
    float *__restrict__ out_a, float *__restrict__ out_x, float *__restrict__ out_z) {
  extern __shared__ __align__(16) char smem[];
  const int nw = g.fx * g.fy * g.Ci * g.Co;
  const float *WT = WT_g;
  ConvWaveStage *stages = reinterpret_cast<ConvWaveStage *>(smem);
  if (kLdsW) {
    float *sW = reinterpret_cast<float *>(smem);
    for (int e = threadIdx.x; e <= nw; e += blockDim.x) sW[e] = e < nw ? WT_g[e] : 0.f;
    __syncthreads();
    WT = sW;
    stages = reinterpret_cast<ConvWaveStage *>(smem + ((size_t)nw * 4 + 15) / 16 * 16 + 16);
  }
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int unit = xcd_logical_block(blockIdx.x, gridDim.x) * kConvSolveWaves + wave;
  if (unit >= units) return;  // whole wave; no workgroup barrier follows
  ConvWaveStage &stg = stages[wave];
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
  // (dx0, dy0) the event's offset from the tile origin; an inactive lane reads weight 0
  // (the padding zero in LDS, or WT[0] times a zero flag from global memory)
  const int loff = (ox * g.fy + oy) * g.Ci * g.Co - cc;
  const int ux = tl.PX + g.fx - 1, uy = tl.PY + g.fy - 1;  // joint field of the tile
  const int P = counts[b];
  const Event *ev = events + (size_t)b * Pcap;
  const FastEvent *fv = fast + (size_t)b * Pcap;
  const int *ix = evidx + (size_t)b * Pcap;
  const double inv2c = 0.5 / c;
  const double zmax_hi = fmin(exp((double)tmax / tau), 1e300) * (1.0 + kFilterEps);
  // an inactive lane, or one whose buffer is full, carries a = -inf (it fails a > 0); the
  // always-exact form uses `live`
  double a = (kAlwaysExact || active) ? 0.0 : -kInfD, bb = 0.0;
  double zq_last = 0.0;  // zq_lo of the lane's previous own event
  int last_q = 0, n = 0;
  bool live = active, seen = false;
  // weight of event (ebase eb, offset dp) for this lane, 0 outside its field
  auto weight = [&](int eb, int dp) -> float {
    bool own = active;
    if (!kSingle)
      own = own & ((unsigned)((dp >> 16) - ox) < (unsigned)g.fx) &
            ((unsigned)((dp & 0xffff) - oy) < (unsigned)g.fy);
    if (kLdsW) return WT[own ? eb - loff : nw];
    const float w = WT[own ? eb - loff : 0];
    return own ? w : 0.f;
  };
  // software pipeline over 64-event chunks: the next chunk's event data and the indices of
  // the one after it are requested before the current chunk is processed
  ConvChunkEv nx = conv_decompose((lane < P) ? ix[lane] : -1, g, tl, x0, y0, ux, uy);
  ConvStageEv nf = conv_lane_event(nx.inu, fv, lane, zmax_hi);
  int qn = (64 + lane < P) ? ix[64 + lane] : -1;
  for (int base = 0; base < P; base += 64) {
    const ConvChunkEv cu = nx;
    const unsigned long long m = __ballot(cu.inu);
    const int cnt = __popcll(m);
    wave_lds_fence();  // the previous chunk's stage reads are done
    if (cu.inu) {
      const int slot = (int)__popcll(m & ((1ull << lane) - 1ull));
      stg.ev[slot] = nf;
      stg.eb[slot] = cu.ebase;
      stg.dp[slot] = cu.dpk;
      stg.q[slot] = base + lane;
    }
    wave_lds_fence();
    if (base + 64 < P) {
      nx = conv_decompose(qn, g, tl, x0, y0, ux, uy);
      nf = conv_lane_event(nx.inu, fv, base + 64 + lane, zmax_hi);
      qn = (base + 128 + lane < P) ? ix[base + 128 + lane] : -1;
    }
    if (cnt == 0) continue;
    // the in-field events in order, each one's stage data and weight read one event ahead
    ConvStageEv E = stg.ev[0];
    int dpc = stg.dp[0], qc = stg.q[0];
    float w = weight(stg.eb[0], dpc);
    for (int k = 0; k < cnt; ++k) {
      const bool more = k + 1 < cnt;
      ConvStageEv En{};
      int dpn = 0, qnx = 0;
      float wn = 0.f;
      if (more) {
        En = stg.ev[k + 1];
        dpn = stg.dp[k + 1];
        qnx = stg.q[k + 1];
        wn = weight(stg.eb[k + 1], dpn);
      }
      const int q = qc;
      bool own = active;
      if (!kSingle)
        own = own & ((unsigned)((dpc >> 16) - ox) < (unsigned)g.fx) &
              ((unsigned)((dpc & 0xffff) - oy) < (unsigned)g.fy);
      if (kAlwaysExact) {
        if (own & live & seen) {
          const SolveState r =
              solve_crossings_ool(a, bb, n, (double)ev[last_q].t, (double)fminf(ev[q].t, tmax),
                                  c, dtt, tau, K, row, out_t, out_a, out_x, out_z);
          bb = r.b;
          n = r.n;
          live = r.live != 0;
        }
      } else {
        const double zv = clamp_lane2(bb * inv2c, zq_last, E.zh);
        const double gq = fma(fma(c, zv, -bb), zv, a);
        if (own & (a > 0.0) & (gq <= kFilterEps * a)) {
          const SolveStateZ r = solve_crossings_deferred_ends(
              a, bb, n, zq_last, E.zh, ConvEnds{ev + last_q, ev + q, tmax}, c, inv2c, dtt, tau,
              K, row, out_t, out_a, out_x, out_z);
          bb = r.b;
          n = r.n;
          if (!r.live) a = -kInfD;
        }
      }
      const double w64 = (double)w;
      a = fma(w64, (double)E.es, a);
      bb = fma(w64, (double)E.e, bb);
      if (kSingle) {
        zq_last = E.zq;
        last_q = q;
        seen = true;
      } else if (own) {
        zq_last = E.zq;
        last_q = q;
        seen = true;
      }
      E = En;
      dpc = dpn;
      qc = qnx;
      w = wn;
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
    const double zv = clamp_lane2(bb * inv2c, zq_last, zmax_hi);
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
  }
  // the +inf tails, written by the whole wave: the channels of one position are consecutive
  // rows, so each position of the tile is one contiguous block of CW x K slots
  wave_lds_fence();
  stg.eb[lane] = active ? n : K;
  wave_lds_fence();
  const int cw = min(tl.CW, g.Co - cg * tl.CW);
  for (int pt = 0; pt < tl.PX * tl.PY; ++pt) {
    const int px_ = x0 + pt / tl.PY, py_ = y0 + pt % tl.PY;
    if (px_ >= g.Xo || py_ >= g.Yo) continue;
    float *blk = out_t + ((size_t)b * g.Xo * g.Yo * g.Co +
                          ((size_t)px_ * g.Yo + py_) * g.Co + cg * tl.CW) * K;
    for (int e = lane; e < cw * K; e += 64) {
      const int r = e / K;
      if (e - r * K >= stg.eb[pt * tl.CW + r]) blk[e] = kInfF;
    }
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
  __shared__ int s_src[kPoolThreads][4];  // the children's first slot in the input
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
    for (int d = 0; d < 4; ++d) {
      s_cnt[threadIdx.x][1 + d] = nd[d];
      s_src[threadIdx.x][d] = (int)base[d];
    }
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
  // (row, slot) of the thread's element, advanced by kPoolThreads elements per step
  int rr = threadIdx.x / slots, s = threadIdx.x - (threadIdx.x / slots) * slots;
  const int step_r = kPoolThreads / slots, step_s = kPoolThreads - step_r * slots;
  float *ot = out_t + (size_t)g0 * slots;
  int *os = out_src + (size_t)g0 * slots;
  for (int e = threadIdx.x; e < rows * slots; e += kPoolThreads) {
    const int cnt = s_cnt[rr][0];
    if (s >= cnt) {
      int q = s - cnt;
      int d = 0;
      while (d < 3 && q >= K - s_cnt[rr][1 + d]) {
        q -= K - s_cnt[rr][1 + d];
        ++d;
      }
      ot[e] = kInfF;
      os[e] = s_src[rr][d] + s_cnt[rr][1 + d] + q;
    }
    s += step_s;
    rr += step_r;
    if (s >= slots) {
      s -= slots;
      ++rr;
    }
  }
}

__global__ void pool_backward_kernel(int64_t total, const float *__restrict__ err,
                                     const int *__restrict__ src, float *__restrict__ pre_err) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < total) pre_err[src[e]] = err[e];
}

// ---------------------------------------------------------------- backward
// ---------------------------------------------------------------- backward
// Post-spike records, one per (sample, output position): slot 0 holds the count T and slots
// 1..T the position's spikes over all its channels in ascending time (ties: channel, then
// spike order) as {t, d1, d2, channel}, with d1, d2 the per-spike weight-gradient terms of
// row_prep_kernel. One wave per position: the channels' spike times are staged in the wave's
// LDS in (channel, spike) order, and an entry's rank is the number of staged entries before it
// in (time, stage order).
__host__ __device__ inline int conv_rec_slots(int Co, int K) {
  return (1 + Co * K + 3) / 4 * 4;
}
__host__ __device__ inline int conv_rec_lds_per_wave(int Co, int K) {
  return ((Co + 1 + Co * K) * 4 + 15) / 16 * 16;
}

__global__ __launch_bounds__(64 * kConvWaves) void conv_rec_kernel(
    int units, int Co, int K, int waves_per_block, const int *__restrict__ n_eff,
    const float *__restrict__ post_t, const float2 *__restrict__ D, float4 *__restrict__ rec) {
  extern __shared__ __align__(16) char smem[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int unit = blockIdx.x * waves_per_block + wave;
  if (unit >= units) return;  // whole wave; no workgroup barrier follows
  const int lane = threadIdx.x & 63;
  int *off = reinterpret_cast<int *>(smem + (size_t)wave * conv_rec_lds_per_wave(Co, K));
  float *ts = reinterpret_cast<float *>(off + Co + 1);
  const size_t row0 = (size_t)unit * Co;  // rows of the position's channels
  int T = 0;
  for (int c0 = 0; c0 < Co; c0 += 64) {
    const int cc = c0 + lane;
    const int nj = cc < Co ? n_eff[row0 + cc] : 0;
    int incl = nj;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(incl, d, 64);
      if (lane >= d) incl += y;
    }
    const int ex = T + incl - nj;
    if (cc < Co) off[cc] = ex;
    for (int k = 0; k < nj; ++k) ts[ex + k] = post_t[(row0 + cc) * K + k];
    T += __shfl(incl, 63, 64);
  }
  if (lane == 0) off[Co] = T;
  wave_lds_fence();
  float4 *R = rec + (size_t)unit * conv_rec_slots(Co, K);
  if (lane == 0) R[0] = make_float4(__int_as_float(T), 0.f, 0.f, 0.f);
  for (int e0 = 0; e0 < T; e0 += 64) {
    const int e = e0 + lane;
    if (e >= T) break;
    // the entry's channel: the last cc with off[cc] <= e (a channel with spikes)
    int lo = 0, hi = Co - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= e)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int k = e - off[lo];
    const float te = ts[e];
    int rank = 0;
    for (int f = 0; f < T; ++f) {
      const float tf = ts[f];
      rank += ((tf < te) | ((tf == te) & (f < e))) ? 1 : 0;
    }
    const float2 d = D[(row0 + lo) * K + k];
    R[1 + rank] = make_float4(te, d.x, d.y, __int_as_float(lo));
  }
}

// Weight gradient: one workgroup per (sample, chunk of output positions, group of 64 filter
// taps (dx, dy, ci), group of CG output channels); lane = tap, and the workgroup's waves take
// the chunk's positions in turn. For output position p the lane's pre neuron is p + (dx, dy),
// channel ci. The position's record is walked in ascending time while the lane's pointer into
// the pre neuron's (ascending) spikes advances, keeping P1 = sum es_m and P2 = sum e_m over the
// pre spikes before the post spike; a post spike of channel cc then adds d1 P1 - d2 P2 to
// acc[cc][lane], the wave's float64 accumulators in LDS (a lane only touches its own
// addresses, so the order of its additions is fixed). At the end the waves' accumulators are
// summed in wave order into the workgroup's partial sums.
struct ConvWgradTiling {
  int tgroups, CG, cgroups, nchunks, ppc;  // ppc: positions per chunk
  int blocks;
};

__global__ __launch_bounds__(64 * kConvWaves) void conv_wgrad_kernel(
    ConvGeom g, int K_pre, const float *__restrict__ pre_t, const float *__restrict__ pre_es,
    const float *__restrict__ pre_e, int rslots, const float4 *__restrict__ rec,
    ConvWgradTiling wt, double *__restrict__ partial) {
  extern __shared__ __align__(16) char smem[];
  double *accs = reinterpret_cast<double *>(smem);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int blk = xcd_logical_block(blockIdx.x, gridDim.x);
  const int cg = blk % wt.cgroups;
  const int tg = (blk / wt.cgroups) % wt.tgroups;
  const int ch = (blk / (wt.cgroups * wt.tgroups)) % wt.nchunks;
  const int b = blk / (wt.cgroups * wt.tgroups * wt.nchunks);
  const int taps = g.fx * g.fy * g.Ci;
  const int tap = tg * 64 + lane;
  const bool active = tap < taps;
  const int tc = active ? tap : 0;
  const int dx = tc / (g.fy * g.Ci);
  const int r2 = tc - dx * (g.fy * g.Ci);
  const int dy = r2 / g.Ci;
  const int ci = r2 - dy * g.Ci;
  const int c_lo = cg * wt.CG;
  double *acc = accs + (size_t)wave * wt.CG * 64;
  for (int u = 0; u < wt.CG; ++u) acc[u * 64 + lane] = 0.0;
  wave_lds_fence();
  const int npos = g.Xo * g.Yo;
  const int Npre = g.X * g.Y * g.Ci;
  const int p_end = min(npos, (ch + 1) * wt.ppc);
  for (int p = ch * wt.ppc + wave; p < p_end; p += kConvWaves) {
    const int x = p / g.Yo, y = p - (p / g.Yo) * g.Yo;
    const size_t pb = ((size_t)b * Npre + ((x + dx) * g.Y + (y + dy)) * g.Ci + ci) * K_pre;
    const float4 *R = rec + ((size_t)b * npos + p) * rslots;
    const int T = __float_as_int(R[0].x);
    int m = 0;
    double P1 = 0.0, P2 = 0.0;
    float tn = (active && K_pre > 0) ? pre_t[pb] : kInfF;
    for (int k = 0; k < T; ++k) {
      const float4 E = R[1 + k];
      while (tn < E.x) {
        P1 += (double)pre_es[pb + m];
        P2 += (double)pre_e[pb + m];
        ++m;
        tn = (m < K_pre) ? pre_t[pb + m] : kInfF;
      }
      const int cc = __float_as_int(E.w) - c_lo;
      if ((unsigned)cc < (unsigned)wt.CG) {
        const double q = fma((double)E.y, P1, -((double)E.z * P2));
        __hip_atomic_fetch_add(acc + cc * 64 + lane, q, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WAVEFRONT);
      }
    }
  }
  __syncthreads();
  double *out = partial + (size_t)blk * wt.CG * 64;
  for (int e = threadIdx.x; e < wt.CG * 64; e += blockDim.x) {
    double v = 0.0;
    for (int w = 0; w < kConvWaves; ++w) v += accs[(size_t)w * wt.CG * 64 + e];
    out[e] = v;
  }
}

// grad[cc, tap] (per sample for mode 1, else summed over the batch and divided by B for
// mode 0) from the workgroups' partial sums, in two passes with a fixed order of additions.
// Pass 1: block (group g = (sample for mode 1, channel, tap group), split s) of 64 x 4
// threads; lane l = tap in the group, row r; the partials n = s*4 + r, s*4 + r + 4S, ... of
// the group are summed per thread, then over the 4 rows, into stage[g][s][l].
// Pass 2: one thread per output sums its S stage values.
__global__ __launch_bounds__(256) void conv_wgrad_reduce1_kernel(
    int B, ConvWgradTiling wt, int mode, int S, const double *__restrict__ partial,
    double *__restrict__ stage) {
  __shared__ double s_row[4][64];
  const int l = threadIdx.x & 63, r = threadIdx.x >> 6;
  const int g = blockIdx.x / S, sp = blockIdx.x - (blockIdx.x / S) * S;
  const int tg = g % wt.tgroups;
  const int cc = (g / wt.tgroups);  // channel, or sample * Co + channel for mode 1
  const int Co = wt.CG * wt.cgroups;  // channel capacity of the groups
  const int bo = mode == 1 ? cc / Co : 0;
  const int ch_c = mode == 1 ? cc - bo * Co : cc;
  const int cg = ch_c / wt.CG, u = ch_c - cg * wt.CG;
  const int N = mode == 1 ? wt.nchunks : B * wt.nchunks;
  double v = 0.0;
  for (int n = sp * 4 + r; n < N; n += 4 * S) {
    const int b = mode == 1 ? bo : n / wt.nchunks;
    const int ch = mode == 1 ? n : n - (n / wt.nchunks) * wt.nchunks;
    const int64_t blk = (((int64_t)b * wt.nchunks + ch) * wt.tgroups + tg) * wt.cgroups + cg;
    v += partial[(blk * wt.CG + u) * 64 + l];
  }
  s_row[r][l] = v;
  __syncthreads();
  if (r == 0)
    stage[(size_t)blockIdx.x * 64 + l] = ((s_row[0][l] + s_row[1][l]) + s_row[2][l]) + s_row[3][l];
}

__global__ void conv_wgrad_reduce2_kernel(int B, int Co_real, int taps, ConvWgradTiling wt,
                                          int mode, int S, const double *__restrict__ stage,
                                          float *__restrict__ grad) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)Co_real * taps;
  if (e >= (mode == 1 ? (int64_t)B * per : per)) return;
  const int bo = (int)(e / per);
  const int rr = (int)(e - (int64_t)bo * per);
  const int cc = rr / taps, tap = rr - cc * taps;
  const int tg = tap / 64, l = tap - tg * 64;
  const int Co = wt.CG * wt.cgroups;
  const int64_t g = ((int64_t)(mode == 1 ? bo * Co + cc : cc)) * wt.tgroups + tg;
  double s = 0.0;
  for (int sp = 0; sp < S; ++sp) s += stage[((size_t)g * S + sp) * 64 + l];
  grad[e] = (float)(mode == 0 ? s / (double)B : s);
}

// First compacted pre spike of every input position of a sample (compacted spikes are in
// neuron order, and the neurons of one position are consecutive): pos_off[b][pp] for
// pp = 0..npos, with pos_off[b][npos] = the sample's spike count.
__global__ void conv_spk_bounds_kernel(int npos, int Ci, int cap, const float4 *__restrict__ spk,
                                       const int *__restrict__ count, int *__restrict__ pos_off) {
  const int b = blockIdx.x;
  const int cnt = count[b];
  const float4 *sp = spk + (size_t)b * cap;
  int *po = pos_off + (size_t)b * (npos + 1);
  if (cnt == 0) {
    for (int pp = threadIdx.x; pp <= npos; pp += blockDim.x) po[pp] = 0;
    return;
  }
  for (int s = threadIdx.x; s < cnt; s += blockDim.x) {
    const int q = __float_as_int(sp[s].w) / Ci;
    const int qp = s > 0 ? __float_as_int(sp[s - 1].w) / Ci : -1;
    for (int pp = qp + 1; pp <= q; ++pp) po[pp] = s;
    if (s == cnt - 1)
      for (int pp = q + 1; pp <= npos; ++pp) po[pp] = cnt;
  }
}

// Pre-spike errors: one wave per (sample, input position); lanes take the position's compacted
// pre spikes (all input channels), 64 at a time. For every output position whose field holds
// the input position (at offset (dx, dy)), the position's record is walked from its latest
// spike down to the earliest pre spike of the chunk, and every post spike (channel cc) later
// than the lane's pre spike adds W[cc, dx, dy, ci] (d1 es/tau_s - d2 e/tau): summed in
// float32 per output position and in float64 over the positions.
constexpr int kPreGroup = 4;

__global__ __launch_bounds__(64 * kConvWaves) void conv_pre_err_kernel(
    ConvGeom g, int B, int K_pre, int cap, const float4 *__restrict__ spk,
    const int *__restrict__ slot, const int *__restrict__ pos_off, int rslots,
    const float4 *__restrict__ rec, const float *__restrict__ W, float inv_tau_s, float inv_tau,
    float *__restrict__ pre_errors) {
  const int npos_in = g.X * g.Y;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int unit = xcd_logical_block(blockIdx.x, gridDim.x) * kConvWaves + wave;
  if (unit >= B * npos_in) return;  // whole wave; no workgroup barrier follows
  const int lane = threadIdx.x & 63;
  const int b = unit / npos_in, pp = unit - b * npos_in;
  const int px = pp / g.Y, py = pp - (pp / g.Y) * g.Y;
  const int s0 = pos_off[(size_t)b * (npos_in + 1) + pp];
  const int s1 = pos_off[(size_t)b * (npos_in + 1) + pp + 1];
  const int x_lo = max(0, px - g.fx + 1), x_hi = min(g.Xo - 1, px);
  const int y_lo = max(0, py - g.fy + 1), y_hi = min(g.Yo - 1, py);
  const int taps = g.fx * g.fy * g.Ci;
  const int npos_out = g.Xo * g.Yo;
  const float4 *sp = spk + (size_t)b * cap;
  for (int c0 = s0; c0 < s1; c0 += 64) {
    const int s = c0 + lane;
    const bool act = s < s1;
    const float4 P = act ? sp[s] : make_float4(kInfF, 0.f, 0.f, 0.f);
    const float tp = P.x, esd = P.y * inv_tau_s, ed = P.z * inv_tau;
    const int ci = act ? __float_as_int(P.w) - pp * g.Ci : 0;
    float tmin = tp;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) tmin = fminf(tmin, __shfl_xor(tmin, d, 64));
    tmin = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(tmin)));
    double pe = 0.0;
    for (int x = x_lo; x <= x_hi; ++x) {
      for (int y = y_lo; y <= y_hi; ++y) {
        const int wb = ((px - x) * g.fy + (py - y)) * g.Ci + ci;
        const float4 *R = rec + ((size_t)b * npos_out + x * g.Yo + y) * rslots;
        const int T = __float_as_int(R[0].x);
        float pt = 0.f;
        for (int k = T; k > 0; k -= kPreGroup) {
          float4 E[kPreGroup];
          float w[kPreGroup];
#pragma unroll
          for (int u = 0; u < kPreGroup; ++u)
            E[u] = (k - u > 0) ? R[k - u] : make_float4(-kInfF, 0.f, 0.f, __int_as_float(0));
#pragma unroll
          for (int u = 0; u < kPreGroup; ++u) w[u] = W[(size_t)__float_as_int(E[u].w) * taps + wb];
#pragma unroll
          for (int u = 0; u < kPreGroup; ++u) {
            const float q = fmaf(esd, E[u].y, -(ed * E[u].z));
            pt = fmaf((tp < E[u].x) ? w[u] : 0.f, q, pt);
          }
          if (!(E[kPreGroup - 1].x > tmin)) break;
        }
        pe += (double)pt;
      }
    }
    if (act) pre_errors[(size_t)b * cap + slot[(size_t)b * cap + s]] = (float)pe;
  }
}

// defined in lif_backward.hip
int post_prep(hipStream_t st, int BJ, int K, const int *n, const float *t, const float *a,
              const float *x, const float *z, float *errors, float c, float tau, float dtt,
              int *n_eff, float2 *D);
int compact_pre_spikes(hipStream_t st, int B, int I, int K_pre, const float *pre_t,
                       const float *pre_es, const float *pre_e, float4 *spk, int *slot,
                       int *count);

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
  const int blocks = (int)((units + kConvSolveWaves - 1) / kConvSolveWaves);
  const size_t wbytes = sizeof(float) * (size_t)taps * g.Co;
  const bool lds_w = conv_solve_lds_bytes(wbytes, true) <= 80 * 1024;
  const bool single = tl.PX * tl.PY == 1;
  const bool exact = !(c > 0.f);
  using KernT = void (*)(const Event *, const FastEvent *, const int *, const int *, int,
                         ConvGeom, ConvTiling, int, const float *, int, double, double, double,
                         float, int *, float *, float *, float *, float *);
  static const KernT kernels[8] = {
      conv_solve_kernel<false, false, false>, conv_solve_kernel<false, false, true>,
      conv_solve_kernel<false, true, false>,  conv_solve_kernel<false, true, true>,
      conv_solve_kernel<true, false, false>,  conv_solve_kernel<true, false, true>,
      conv_solve_kernel<true, true, false>,   conv_solve_kernel<true, true, true>};
  const KernT kern = kernels[(exact ? 4 : 0) + (lds_w ? 2 : 0) + (single ? 1 : 0)];
  const size_t lds = conv_solve_lds_bytes(wbytes, lds_w);
  if (lds > 64 * 1024)
    BATS_REQUIRE(hipFuncSetAttribute((const void *)kern,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds) == hipSuccess,
                 "conv_solve_kernel: LDS request above the device limit");
  BATS_TIMED("conv_solve_kernel", st,
             hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * kConvSolveWaves), lds, st,
                                el.events, el.fast, el.idx, el.counts, el.Pcap, g, tl,
                                (int)units, WT, K_post, (double)c, (double)delta_theta_tau,
                                (double)tau, max_simulation, out_n, out_t, out_a, out_x, out_z));
  return check_launch("conv_solve_kernel");
}

// Workspace of the conv backward (byte offsets) and the weight-gradient tiling.
struct ConvBwdLayout {
  size_t n_eff, D, rec, partial, stage, spk, slot, count, pos_off, total;
  int rslots, cap;
  int S1;  // splits of the first reduction pass
  ConvWgradTiling wt;
};

static ConvBwdLayout conv_bwd_layout(int B, const ConvGeom &g, int K_pre, int K_post) {
  ConvBwdLayout L{};
  const size_t BJ = (size_t)B * g.Xo * g.Yo * g.Co;
  const int npos = g.Xo * g.Yo;
  const int taps = g.fx * g.fy * g.Ci;
  ConvWgradTiling &wt = L.wt;
  wt.tgroups = (taps + 63) / 64;
  wt.cgroups = (g.Co + 63) / 64;
  wt.CG = (g.Co + wt.cgroups - 1) / wt.cgroups;
  // about 1536 workgroups (6144 waves), each wave with a few positions
  const long long per_chunk = (long long)std::max(B, 1) * wt.tgroups * wt.cgroups;
  int nchunks = (int)std::min<long long>((1536 + per_chunk - 1) / per_chunk,
                                         (npos + kConvWaves - 1) / kConvWaves);
  nchunks = std::max(nchunks, 1);
  wt.ppc = (npos + nchunks - 1) / nchunks;
  wt.nchunks = (npos + wt.ppc - 1) / wt.ppc;
  wt.blocks = (int)(per_chunk * wt.nchunks);
  L.rslots = conv_rec_slots(g.Co, K_post);
  L.cap = g.X * g.Y * g.Ci * K_pre;
  size_t off = 0;
  L.n_eff = off;
  off = align_up(off + sizeof(int) * BJ, 256);
  L.D = off;
  off = align_up(off + sizeof(float2) * BJ * K_post, 256);
  L.rec = off;
  off = align_up(off + sizeof(float4) * (size_t)B * npos * L.rslots, 256);
  L.partial = off;
  off = align_up(off + sizeof(double) * (size_t)wt.blocks * wt.CG * 64, 256);
  // first reduction pass: about 16 partials per thread
  const int Nred = wt.nchunks * std::max(B, 1);
  L.S1 = std::max(1, std::min(256, (Nred + 63) / 64));
  const int64_t groups = (int64_t)wt.cgroups * wt.CG * wt.tgroups;  // modes 0 and 2
  const int64_t groups1 = groups * std::max(B, 1);                 // mode 1 (S = 1)
  L.stage = off;
  off = align_up(off + sizeof(double) * 64 * (size_t)std::max(groups * L.S1, groups1), 256);
  L.spk = off;
  off = align_up(off + sizeof(float4) * (size_t)B * L.cap, 256);
  L.slot = off;
  off = align_up(off + sizeof(int) * (size_t)B * L.cap, 256);
  L.count = off;
  off = align_up(off + sizeof(int) * (size_t)B, 256);
  L.pos_off = off;
  off = align_up(off + sizeof(int) * (size_t)B * (g.X * g.Y + 1), 256);
  L.total = off;
  return L;
}

extern "C" size_t bats_conv_lif_backward_workspace_size(int B, const int *pre_shape, int K_pre,
                                                        const int *filters_shape, int K_post) {
  ConvGeom g;
  if (B < 0 || K_pre < 0 || K_post < 1 || !pre_shape || !filters_shape ||
      !conv_geom(pre_shape, filters_shape, g))
    return 0;
  return conv_bwd_layout(B, g, K_pre, K_post).total;
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
  BATS_REQUIRE(conv_rec_lds_per_wave(g.Co, K_post) <= 64 * 1024,
               "C_out * K_post above 16000 (post-spike record stage)");
  BATS_REQUIRE((long long)g.X * g.Y * g.Ci * K_pre < (1LL << 31), "pre spike buffer too large");
  if (B == 0) return 0;
  const ConvBwdLayout L = conv_bwd_layout(B, g, K_pre, K_post);
  BATS_REQUIRE(workspace && workspace_bytes >= L.total, "workspace too small");
  BATS_REQUIRE(n && t && a && x && z && errors && grad, "null pointer");
  if (K_pre > 0) BATS_REQUIRE(pre_times && pre_exp_tau_s && pre_exp_tau, "null pre pointer");
  hipStream_t st = as_stream(stream);
  char *ws = static_cast<char *>(workspace);
  int *n_eff = reinterpret_cast<int *>(ws + L.n_eff);
  float2 *D = reinterpret_cast<float2 *>(ws + L.D);
  float4 *rec = reinterpret_cast<float4 *>(ws + L.rec);
  double *partial = reinterpret_cast<double *>(ws + L.partial);
  const int npos = g.Xo * g.Yo;
  const size_t BJ = (size_t)B * npos * g.Co;
  if (int r = post_prep(st, (int)BJ, K_post, n, t, a, x, z, errors, c, tau, delta_theta_tau,
                        n_eff, D))
    return r;
  {
    const int units = B * npos;
    const int lw = conv_rec_lds_per_wave(g.Co, K_post);
    const int wpb = std::max(1, std::min(kConvWaves, 64 * 1024 / lw));
    BATS_TIMED("conv_rec_kernel", st,
               hipLaunchKernelGGL(conv_rec_kernel, dim3((units + wpb - 1) / wpb),
                                  dim3(64 * wpb), (size_t)lw * wpb, st, units, g.Co, K_post, wpb,
                                  n_eff, t, D, rec));
    if (int r = check_launch("conv_rec_kernel")) return r;
  }
  const int taps = g.fx * g.fy * g.Ci;
  const ConvWgradTiling &wt = L.wt;
  const size_t wlds = sizeof(double) * kConvWaves * wt.CG * 64;
  if (wlds > 64 * 1024)
    BATS_REQUIRE(hipFuncSetAttribute((const void *)conv_wgrad_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)wlds) == hipSuccess,
                 "conv_wgrad_kernel: LDS request above the device limit");
  BATS_TIMED("conv_wgrad_kernel", st,
             hipLaunchKernelGGL(conv_wgrad_kernel, dim3(wt.blocks), dim3(64 * kConvWaves), wlds,
                                st, g, K_pre, pre_times, pre_exp_tau_s, pre_exp_tau, L.rslots,
                                rec, wt, partial));
  if (int r = check_launch("conv_wgrad_kernel")) return r;
  {
    double *stage = reinterpret_cast<double *>(ws + L.stage);
    const int S = grad_mode == 1 ? 1 : L.S1;
    const int64_t groups =
        (int64_t)(grad_mode == 1 ? B : 1) * wt.cgroups * wt.CG * wt.tgroups;
    BATS_TIMED("conv_wgrad_reduce_kernel", st,
               hipLaunchKernelGGL(conv_wgrad_reduce1_kernel, dim3((unsigned)(groups * S)),
                                  dim3(256), 0, st, B, wt, grad_mode, S, partial, stage));
    if (int r = check_launch("conv_wgrad_reduce1_kernel")) return r;
    const int64_t fin = (int64_t)(grad_mode == 1 ? B : 1) * g.Co * taps;
    hipLaunchKernelGGL(conv_wgrad_reduce2_kernel, dim3((unsigned)((fin + 255) / 256)),
                       dim3(256), 0, st, B, g.Co, taps, wt, grad_mode, S, stage, grad);
    if (int r = check_launch("conv_wgrad_reduce2_kernel")) return r;
  }
  if (pre_errors && K_pre > 0) {
    BATS_REQUIRE(weights, "null weights");
    float4 *spk = reinterpret_cast<float4 *>(ws + L.spk);
    int *slot = reinterpret_cast<int *>(ws + L.slot);
    int *count = reinterpret_cast<int *>(ws + L.count);
    int *pos_off = reinterpret_cast<int *>(ws + L.pos_off);
    const int npos_in = g.X * g.Y;
    const int64_t nerr = (int64_t)B * L.cap;
    hipLaunchKernelGGL(fill_kernel<float>, dim3((unsigned)std::min<int64_t>((nerr + 255) / 256, 65535)),
                       dim3(256), 0, st, pre_errors, nerr, 0.f);
    if (int r = check_launch("fill_kernel")) return r;
    if (int r = compact_pre_spikes(st, B, g.X * g.Y * g.Ci, K_pre, pre_times, pre_exp_tau_s,
                                   pre_exp_tau, spk, slot, count))
      return r;
    hipLaunchKernelGGL(conv_spk_bounds_kernel, dim3(B), dim3(256), 0, st, npos_in, g.Ci, L.cap,
                       spk, count, pos_off);
    if (int r = check_launch("conv_spk_bounds_kernel")) return r;
    const int64_t units = (int64_t)B * npos_in;
    BATS_TIMED("conv_pre_err_kernel", st,
               hipLaunchKernelGGL(conv_pre_err_kernel,
                                  dim3((unsigned)((units + kConvWaves - 1) / kConvWaves)),
                                  dim3(64 * kConvWaves), 0, st, g, B, K_pre, L.cap, spk, slot,
                                  pos_off, L.rslots, rec, weights, (float)(1.0 / (double)tau_s),
                                  (float)(1.0 / (double)tau), pre_errors));
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
