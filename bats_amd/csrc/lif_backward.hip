// Dense LIF layer BATS backward for gfx950.
//
// Both the weight gradient (reference compute_weights_gradient.cu:33-45) and the errors of
// the pre-synaptic spikes (propagate_errors_to_pre_spikes.cu:41-53) are functions of the
// same quantity: for a pre spike at time t_m and a post neuron j of the same sample,
//   V1_j(t_m) = sum_{k: t_jk > t_m} e_jk f1_jk,   V2_j(t_m) = sum_{k: t_jk > t_m} e_jk f2_jk.
// Regrouping the reference's per-post-spike sums gives
//   dW[j,i]      = sum_b sum_m  es_m V1_j(t_m) - e_m V2_j(t_m)       (m: spikes of pre i)
//   pre_err[b,i,m] = sum_j W[j,i] (V1_j(t_m) es_m / tau_s - V2_j(t_m) e_m / tau)
// so the backward of a wide layer is three kernels:
//   1. row_prep_kernel   one lane per (sample, post neuron): the reset-chain error scan
//      (propagate_recurrent_errors.cu:17-22, written back in place like the reference),
//      the factors f1, f2 (compute_factors.cu:20-30, rounded to float32 as stored there),
//      and per spike the terms d1 = e f1, d2 = e f2, written as per-tile spike records
//      (V_j(t) is the sum of the d of the spikes of row j later than t).
//   2. wgrad_tile_kernel lanes over pre neurons i, a tile of 16 post neurons, a slice of the
//      batch: the terms of the post spikes later than the lane's pre spike, float64
//      accumulation per row; a deterministic fixed-order reduction over batch slices
//      (wgrad_reduce_kernel) finishes the mean.
//   3. pre_err_tile_kernel lanes over pre neurons i of one sample, the same tile records
//      walked tile by tile with the tile's block of W in LDS.
// Layers of at most 16 post neurons (an output layer) run all of it in small_bwd_kernel.
#include "backward_bins.hpp"
#include "common.hpp"

#include <algorithm>

namespace bats {

constexpr int kBwdThreads = 256;

// One lane per post row (sample b, neuron j): the reset-chain error scan
// (propagate_recurrent_errors.cu:17-22, written back in place like the reference), the
// factors f1, f2 (compute_factors.cu:20-30, rounded to float32 as stored there) and the
// per-spike weight-gradient terms d1 = e f1, d2 = e f2. Outputs (each optional):
//   n_eff [B*J]      spikes before the row's first infinite time
//   D [B*J*K]        the per-spike terms {d1, d2} in float32 (read by the conv backward)
//   rec              one record of tile_rec_slots(K) float4 per (sample, tile of kWT rows):
//                      slot 0     x: number of spikes in the tile (int)
//                      slots 1..  the spikes of row 0, row 1, ... as {t, d1, d2, r} with r
//                                 the byte offset of the row's accumulators (row * 512)
// Lanes are (sample, tile, row in tile), so the kWT rows of a tile are kWT consecutive lanes
// of one wave and a record's layout comes from a 16-lane prefix sum.
// Rows of up to kPrepFast spikes (most rows of the hidden layers measured) load their slots
// with independent loads into registers, so that a row costs one memory round trip instead of
// one per spike; longer rows run the same arithmetic from global memory, kPrepFast slots per
// round. 4 rather than 8: the deep network's three hidden layers (about 1, 1.7 and 3 spikes
// per row) 51.5 -> 49 us per launch on average, the 784-800-10 hidden layer unchanged.
// apply_reset = 0: the errors already hold the reset-chain result.
constexpr int kPrepFast = 4;

__global__ __launch_bounds__(kBwdThreads) void row_prep_kernel(
    int B, int J, int K, const int *__restrict__ n, const float *__restrict__ t,
    const float *__restrict__ a, const float *__restrict__ x, const float *__restrict__ z,
    float *__restrict__ errors, double c, double tau, double dtt, int apply_reset,
    int *__restrict__ n_eff, float2 *__restrict__ D, float4 *__restrict__ rec) {
  const int gj = (J + kWT - 1) / kWT;
  const int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tiles = (int64_t)B * gj;  // lane < tiles * kWT for every launched lane group
  const int64_t tile = lane / kWT;
  const int u = (int)(lane % kWT);
  const int b = (int)(tile / gj);
  const int j = (int)(tile % gj) * kWT + u;
  const bool valid = tile < tiles;
  const bool active = valid && j < J;
  const size_t r = active ? (size_t)b * J + j : 0;
  const size_t row = r * K;
  const int nk = active ? min(n[r], K) : 0;
  const bool fast = nk <= kPrepFast;

  float tv[kPrepFast], av[kPrepFast], xv[kPrepFast], zv[kPrepFast], ev[kPrepFast];
#pragma unroll
  for (int k = 0; k < kPrepFast; ++k) {
    const bool in = fast && k < nk;
    tv[k] = in ? t[row + k] : kInfF;
    av[k] = in ? a[row + k] : 1.f;
    xv[k] = in ? x[row + k] : 1.f;
    zv[k] = in ? z[row + k] : kInfF;
    ev[k] = in ? errors[row + k] : 0.f;
  }
  int ne = 0;
  if (fast) {
    if (apply_reset) {
      double C = 0.0;
#pragma unroll
      for (int k = kPrepFast - 1; k >= 0; --k) {
        if (k < nk && !isinf(zv[k])) {
          const double ek = (double)ev[k] + C * (double)zv[k];
          ev[k] = (float)ek;
          C = C + (double)ev[k] * dtt / (double)xv[k];
          errors[row + k] = ev[k];
        }
      }
    }
    bool run = true;
#pragma unroll
    for (int k = 0; k < kPrepFast; ++k) {
      run = run & (k < nk) & !isinf(tv[k]);
      ne += run ? 1 : 0;
    }
  } else {
    // longer rows: the same arithmetic over chunks of kPrepFast slots, each chunk's slots
    // loaded in one round (the reset chain runs from the last chunk to the first)
    if (apply_reset) {
      double C = 0.0;
      for (int c1 = nk; c1 > 0; c1 -= kPrepFast) {
        const int c0 = max(0, c1 - kPrepFast);
        float zc[kPrepFast], ec[kPrepFast], xc[kPrepFast];
#pragma unroll
        for (int k = 0; k < kPrepFast; ++k) {
          const bool in = c0 + k < c1;
          zc[k] = in ? z[row + c0 + k] : kInfF;
          ec[k] = in ? errors[row + c0 + k] : 0.f;
          xc[k] = in ? x[row + c0 + k] : 1.f;
        }
#pragma unroll
        for (int k = kPrepFast - 1; k >= 0; --k) {
          if (c0 + k < c1 && !isinf(zc[k])) {
            const double ek = (double)ec[k] + C * (double)zc[k];
            ec[k] = (float)ek;
            C = C + (double)ec[k] * dtt / (double)xc[k];
            errors[row + c0 + k] = ec[k];
          }
        }
      }
    }
    bool run = true;
    for (int c0 = 0; c0 < nk && run; c0 += kPrepFast) {
      float tc[kPrepFast];
#pragma unroll
      for (int k = 0; k < kPrepFast; ++k) tc[k] = c0 + k < nk ? t[row + c0 + k] : kInfF;
#pragma unroll
      for (int k = 0; k < kPrepFast; ++k) {
        run = run & !isinf(tc[k]);
        ne += run ? 1 : 0;
      }
    }
  }
  if (active && n_eff) n_eff[r] = ne;
  // the row's spikes in its tile record
  int incl = ne;
#pragma unroll
  for (int d = 1; d < kWT; d <<= 1) {
    const int y = __shfl_up(incl, d, kWT);
    if (u >= d) incl += y;
  }
  const int excl = incl - ne;
  const int total = __shfl(incl, kWT - 1, kWT);
  float4 *tr = rec ? rec + (size_t)tile * tile_rec_slots(K) : nullptr;
  if (tr && valid && u == 0) tr[0] = make_float4(__int_as_float(total), 0.f, 0.f, 0.f);
  if (!active) return;
  float4 *ent = tr ? tr + kRecHead + excl : nullptr;
  const float roff = __int_as_float(u * 64 * (int)sizeof(double));
  if (fast) {
#pragma unroll
    for (int k = kPrepFast - 1; k >= 0; --k) {
      if (k < ne) {
        const double ak = (double)av[k], xk = (double)xv[k], zk = (double)zv[k];
        const float f1 = (float)(tau / ak * (1.0 + zk * c / xk));
        const float f2 = (float)(tau / xk);
        const double d1 = (double)ev[k] * (double)f1, d2 = (double)ev[k] * (double)f2;
        if (D) D[row + k] = make_float2((float)d1, (float)d2);
        if (ent) ent[k] = make_float4(tv[k], (float)d1, (float)d2, roff);
      }
    }
    return;
  }
  for (int c1 = ne; c1 > 0; c1 -= kPrepFast) {
    const int c0 = max(0, c1 - kPrepFast);
#pragma unroll
    for (int k = 0; k < kPrepFast; ++k) {
      const bool in = c0 + k < c1;
      tv[k] = in ? t[row + c0 + k] : kInfF;
      av[k] = in ? a[row + c0 + k] : 1.f;
      xv[k] = in ? x[row + c0 + k] : 1.f;
      zv[k] = in ? z[row + c0 + k] : 1.f;
      ev[k] = in ? errors[row + c0 + k] : 0.f;
    }
#pragma unroll
    for (int k = kPrepFast - 1; k >= 0; --k) {
      if (c0 + k < c1) {
        const double ak = (double)av[k], xk = (double)xv[k], zk = (double)zv[k];
        const float f1 = (float)(tau / ak * (1.0 + zk * c / xk));
        const float f2 = (float)(tau / xk);
        const double d1 = (double)ev[k] * (double)f1, d2 = (double)ev[k] * (double)f2;
        if (D) D[row + c0 + k] = make_float2((float)d1, (float)d2);
        if (ent) ent[c0 + k] = make_float4(tv[k], (float)d1, (float)d2, roff);
      }
    }
  }
}

static int launch_row_prep(hipStream_t st, int B, int J, int K, const int *n, const float *t,
                           const float *a, const float *x, const float *z, float *errors,
                           double c, double tau, double dtt, int *n_eff, float2 *D,
                           float4 *rec) {
  const int64_t lanes = (int64_t)B * ((J + kWT - 1) / kWT) * kWT;
  if (lanes <= 0) return 0;
  // small launches (an output layer's rows) in one-wave blocks, spread over many CUs
  const int threads = lanes < 256 * 256 ? 64 : kBwdThreads;
  hipLaunchKernelGGL(row_prep_kernel, dim3((unsigned)((lanes + threads - 1) / threads)),
                     dim3(threads), 0, st, B, J, K, n, t, a, x, z, errors, c, tau, dtt, 1,
                     n_eff, D, rec);
  return check_launch("row_prep_kernel");
}

// Number of leading finite slots of a spike row. The first 8 slots are read with independent
// loads in one round; the rest of the row only when all 8 are finite (spike rows are short:
// a hidden neuron of the 784-800-10 network fires about once per sample).
__device__ __forceinline__ int leading_finite(const float *__restrict__ row, int K) {
  constexpr int kHead = 8;
  int n = 0;
  bool run = true;
#pragma unroll
  for (int m = 0; m < kHead; ++m) {
    const float v = m < K ? row[m] : kInfF;  // independent of run: all loads in one round
    run = run & (bool)isfinite(v);
    n += run ? 1 : 0;
  }
  if (!run) return n;
  for (int m = kHead; m < K; ++m) {
    if (!isfinite(row[m])) break;
    ++n;
  }
  return n;
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

// Many slices (one per sample): a block sums 64 consecutive elements, its 4 thread groups
// take every 4th slice each with independent loads, and the 4 group sums are added in a
// fixed order, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void wgrad_reduce_tall_kernel(
    int slices, int64_t JI, double inv_b, const float *__restrict__ partial,
    float *__restrict__ grad) {
  __shared__ double s_part[4][64];
  const int col = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const int64_t e = (int64_t)blockIdx.x * 64 + col;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (e < JI) {
    int z = grp;
    for (; z + 12 < slices; z += 16) {
      acc[0] += (double)partial[(size_t)z * JI + e];
      acc[1] += (double)partial[(size_t)(z + 4) * JI + e];
      acc[2] += (double)partial[(size_t)(z + 8) * JI + e];
      acc[3] += (double)partial[(size_t)(z + 12) * JI + e];
    }
    for (; z < slices; z += 4) acc[0] += (double)partial[(size_t)z * JI + e];
  }
  s_part[grp][col] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  if (grp == 0 && e < JI)
    grad[e] = (float)(((s_part[0][col] + s_part[1][col]) + (s_part[2][col] + s_part[3][col])) *
                      inv_b);
}

// Sums the per-slice partial gradients into grad, scaled by inv_b.
static int reduce_slices(hipStream_t st, int slices, int64_t JI, double inv_b,
                         const float *partial, float *grad) {
  if (slices >= 32) {
    hipLaunchKernelGGL(wgrad_reduce_tall_kernel, dim3((unsigned)((JI + 63) / 64)), dim3(256), 0,
                       st, slices, JI, inv_b, partial, grad);
    return check_launch("wgrad_reduce_tall_kernel");
  }
  hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((unsigned)((JI + kBwdThreads - 1) / kBwdThreads)),
                     dim3(kBwdThreads), 0, st, slices, JI, inv_b, partial, grad);
  return check_launch("wgrad_reduce_kernel");
}

// ------------------------------------------------------------------ wide layers
// Weight gradient: one wave per (block of 64 pre neurons, tile of kWT post neurons, slice of
// the batch); lane = pre neuron i. For each sample of the slice and each pre spike (t_p, es,
// e) of the lane, every spike k of every row u of the tile adds
//   q = es d1_k - e d2_k   if t_p < t_k      (d = the row_prep_kernel terms)
// to the float64 accumulator of (row u, lane), summing V1(t_p) es - V2(t_p) e term by term.
// The accumulators live in the wave's own LDS array acc[u][lane] and each term is added
// with one LDS float64 atomic that returns nothing (a lane only ever touches its own
// addresses, so the order of its additions is fixed): the walk is one loop over the tile's
// spike list with no per-row code, and empty rows (more than half of a hidden layer's rows)
// cost nothing. q is formed in float32 (the terms are float32).
//
// The sample's tile record is fetched one sample ahead with one coalesced load of its first
// kRecWindow slots, parked in the wave's own LDS window and read back with wave-uniform
// (broadcast) LDS reads; spikes past the window come from global memory.
//
// A workgroup holds kTileWaves independent waves (consecutive logical tasks, each with its
// own LDS; no workgroup barrier). Block order: the waves of one (tile, slice) read the same
// records, so consecutive logical blocks are placed on one XCD (workgroups are dealt to the
// 8 XCDs round robin) and share its L2.
constexpr int kRecWindow = 64;
constexpr int kTileWaves = 4;


template <bool kSingle>
__global__ __launch_bounds__(64 * kTileWaves) void wgrad_tile_kernel(
    int B, int I, int K_pre, int J, int K, int gi, int gj, int b_per_slice, int ntasks,
    const float *__restrict__ pre_t, const float *__restrict__ pre_es,
    const float *__restrict__ pre_e, const float4 *__restrict__ rec, int per_sample,
    float *__restrict__ grad, float *__restrict__ partial) {
  __shared__ float4 wins[kTileWaves][kRecWindow];
  __shared__ double accs[kTileWaves][kWT * 64];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
  const int L = xcd_logical_block(blockIdx.x, gridDim.x) * kTileWaves + wv;
  if (L >= ntasks) return;  // whole wave; no workgroup barrier follows
  float4 *win = wins[wv];
  double *acc = accs[wv];
  const int bi = L % gi;
  const int tj = (L / gi) % gj;
  const int sl = L / (gi * gj);
  const int lane = threadIdx.x & 63;
  const int i = bi * 64 + lane;
  const bool active = i < I;
  const int ic = active ? i : 0;
  const int j0 = tj * kWT;
  const int jn = min(kWT, J - j0);
  const int b0 = sl * b_per_slice;
  const int b1 = min(B, b0 + b_per_slice);
  const int R = tile_rec_slots(K);
  const int wn = min(kRecWindow, R);  // slots of a record held in the window
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
  char *acc_lane = reinterpret_cast<char *>(acc + lane);

#pragma unroll
  for (int u = 0; u < kWT; ++u) acc[u * 64 + lane] = 0.0;

  // the next sample's record window and pre slot, loaded during the current walk
  float4 nrec = zero4;
  float ntp = kInfF, nes = 0.f, nee = 0.f;
  if (b0 < b1) {
    if (lane < wn) nrec = rec[((size_t)b0 * gj + tj) * R + lane];
    if (kSingle && active) {
      const size_t pb = (size_t)b0 * I + i;
      ntp = pre_t[pb];
      nes = pre_es[pb];
      nee = pre_e[pb];
    }
  }
  for (int b = b0; b < b1; ++b) {
    const float4 *grec = rec + ((size_t)b * gj + tj) * R;
    const int cnt = __builtin_amdgcn_readlane(__float_as_int(nrec.x), 0);
    wave_lds_fence();  // the previous sample's window reads are done
    win[lane] = nrec;
    wave_lds_fence();
    if (b + 1 < b1 && lane < wn) nrec = grec[(size_t)gj * R + lane];
    const int nw = min(cnt, wn - kRecHead);  // spikes inside the window
    auto walk = [&](float tp, float es, float e) {
      auto term = [&](const float4 E) {
        float q = fmaf(es, E.y, -(e * E.z));
        // formed for every lane: an opaque barrier keeps the compiler from turning the
        // select into a branch around the term
        asm volatile("" : "+v"(q));
        q = (tp < E.x) ? q : 0.f;
        double *a = reinterpret_cast<double *>(acc_lane + __float_as_int(E.w));
        __hip_atomic_fetch_add(a, (double)q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      };
      // window entries eight (then four) at a time: their LDS reads in one round
      int k = 0;
      for (; k + 8 <= nw; k += 8) {
        float4 E[8];
#pragma unroll
        for (int g = 0; g < 8; ++g) E[g] = win[kRecHead + k + g];
#pragma unroll
        for (int g = 0; g < 8; ++g) term(E[g]);
      }
      for (; k + 4 <= nw; k += 4) {
        const float4 E0 = win[kRecHead + k], E1 = win[kRecHead + k + 1];
        const float4 E2 = win[kRecHead + k + 2], E3 = win[kRecHead + k + 3];
        term(E0);
        term(E1);
        term(E2);
        term(E3);
      }
      for (; k < nw; ++k) term(win[kRecHead + k]);
      for (int k = nw; k < cnt; ++k) term(grec[kRecHead + k]);
    };
    if (kSingle) {
      const float tp = ntp, es = nes, e = nee;
      if (b + 1 < b1 && active) {
        const size_t pb = (size_t)(b + 1) * I + i;
        ntp = pre_t[pb];
        nes = pre_es[pb];
        nee = pre_e[pb];
      }
      // a slot without a spike has no defined exps: its contribution is forced to zero
      const bool has = isfinite(tp);
      walk(has ? tp : kInfF, has ? es : 0.f, has ? e : 0.f);
    } else {
      const size_t pbase = ((size_t)b * I + ic) * K_pre;
      const int np = active ? leading_finite(pre_t + pbase, K_pre) : 0;
      for (int m = 0; m < np; ++m) walk(pre_t[pbase + m], pre_es[pbase + m], pre_e[pbase + m]);
    }
    if (per_sample) {
      wave_lds_fence();
#pragma unroll
      for (int u = 0; u < kWT; ++u) {
        if (u < jn && active) grad[((size_t)b * J + j0 + u) * I + i] = (float)acc[u * 64 + lane];
        acc[u * 64 + lane] = 0.0;
      }
    }
  }
  if (!per_sample && active) {
    wave_lds_fence();
#pragma unroll
    for (int u = 0; u < kWT; ++u)
      if (u < jn) partial[((size_t)sl * J + j0 + u) * I + i] = (float)acc[u * 64 + lane];
  }
}

// Pre-spike errors of wide layers (the previous layer is trainable):
//   pre_err[b,i,m] = sum_j W[j,i] sum_{k: t_jk > t_m} (d1_k es_m / tau_s - d2_k e_m / tau)
// One wave per (block of 64 pre neurons, sample); lane = pre neuron i. The wave walks the
// sample's kWT-row tile records (row_prep_kernel) one tile at a time: the tile's 16 x 64
// block of W sits in the wave's own LDS array (one coalesced load per row, fetched a tile
// ahead, with the next record window), and each spike entry adds
//   w[row] * (es'_m d1 - e'_m d2)     if t_m < t_k       (es' = es / tau_s, e' = e / tau)
// for kPreM pre spikes of the lane at once, in float32 within the tile and float64 across
// tiles. Lanes with more pre spikes take further passes over the records.
constexpr int kPreM = 2;

template <int M>
__global__ __launch_bounds__(64 * kTileWaves) void pre_err_tile_kernel(
    int B, int I, int K_pre, int J, int K, int gi, int ntasks, const float *__restrict__ pre_t,
    const float *__restrict__ pre_es, const float *__restrict__ pre_e,
    const float *__restrict__ W, const float4 *__restrict__ rec, double inv_tau_s,
    double inv_tau, float *__restrict__ pre_errors) {
  __shared__ float4 wins[kTileWaves][kRecWindow];
  __shared__ float wts[kTileWaves][kWT * 64];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
  const int L = xcd_logical_block(blockIdx.x, gridDim.x) * kTileWaves + wv;
  if (L >= ntasks) return;  // whole wave; no workgroup barrier follows
  float4 *win = wins[wv];
  float *wt = wts[wv];
  const int bi = L % gi;
  const int b = L / gi;
  const int lane = threadIdx.x & 63;
  const int i = bi * 64 + lane;
  const bool active = i < I;
  const int ic = active ? i : 0;
  const int gj = (J + kWT - 1) / kWT;
  const int R = tile_rec_slots(K);
  const int wn = min(kRecWindow, R);
  const size_t pbase = ((size_t)b * I + ic) * K_pre;
  const int np = active ? leading_finite(pre_t + pbase, K_pre) : 0;
  // passes over the records: enough for the lane of the wave with the most pre spikes
  int npmax = np;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) npmax = max(npmax, __shfl_xor(npmax, d, 64));
  npmax = __builtin_amdgcn_readfirstlane(npmax);
  const float4 *grec0 = rec + (size_t)b * gj * R;
  const float zf = 0.f;
  for (int m0 = 0; m0 < npmax; m0 += M) {
    float tp[M], esd[M], ed[M];
    double pe[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const bool has = m0 + m < np;
      tp[m] = has ? pre_t[pbase + m0 + m] : kInfF;
      esd[m] = has ? (float)((double)pre_es[pbase + m0 + m] * inv_tau_s) : zf;
      ed[m] = has ? (float)((double)pre_e[pbase + m0 + m] * inv_tau) : zf;
      pe[m] = 0.0;
    }
    // tile 0's record window and W block
    float4 nrec = lane < wn ? grec0[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    float nw[kWT];
#pragma unroll
    for (int u = 0; u < kWT; ++u) nw[u] = (active && u < J) ? W[(size_t)u * I + i] : 0.f;
    for (int tj = 0; tj < gj; ++tj) {
      const float4 *grec = grec0 + (size_t)tj * R;
      const int cnt = __builtin_amdgcn_readlane(__float_as_int(nrec.x), 0);
      wave_lds_fence();  // the previous tile's reads are done
      win[lane] = nrec;
#pragma unroll
      for (int u = 0; u < kWT; ++u) wt[u * 64 + lane] = nw[u];
      wave_lds_fence();
      if (tj + 1 < gj) {
        if (lane < wn) nrec = grec[R + lane];
        const int j1 = (tj + 1) * kWT;
#pragma unroll
        for (int u = 0; u < kWT; ++u)
          nw[u] = (active && j1 + u < J) ? W[(size_t)(j1 + u) * I + i] : 0.f;
      }
      float pt[M];
#pragma unroll
      for (int m = 0; m < M; ++m) pt[m] = 0.f;
      const int nwin = min(cnt, wn - kRecHead);
      auto term = [&](const float4 E) {
        // the entry's row: its accumulator offset (row * 512 bytes) / 2 is the row's W offset
        const float w = wt[(__float_as_int(E.w) >> 3) + lane];
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const float q = fmaf(esd[m], E.y, -(ed[m] * E.z));
          pt[m] = fmaf((tp[m] < E.x) ? w : 0.f, q, pt[m]);
        }
      };
      for (int k = 0; k < nwin; ++k) term(win[kRecHead + k]);
      for (int k = nwin; k < cnt; ++k) term(grec[kRecHead + k]);
#pragma unroll
      for (int m = 0; m < M; ++m) pe[m] += (double)pt[m];
    }
#pragma unroll
    for (int m = 0; m < M; ++m)
      if (m0 + m < np) pre_errors[pbase + m0 + m] = (float)pe[m];
  }
  if (active)
    for (int m = np; m < K_pre; ++m) pre_errors[pbase + m] = 0.f;
}

// ------------------------------------------------------------------ multi-spike pre layers
// When the pre neurons fire several times (the previous layer is a LIF layer with K_pre > 1),
// a lane per pre neuron would walk a tile once per spike of the busiest neuron of its wave.
// These kernels give each lane one pre spike instead: compact_spikes_kernel lists, per
// sample, the finite pre slots in neuron order as {t, es, e, neuron} with their slot index.

constexpr int kCompThreads = 256;

__global__ __launch_bounds__(kCompThreads) void compact_spikes_kernel(
    int I, int K_pre, const float *__restrict__ pre_t, const float *__restrict__ pre_es,
    const float *__restrict__ pre_e, int cap, float4 *__restrict__ spk, int *__restrict__ slot,
    int *__restrict__ count) {
  __shared__ int s_wave[kCompThreads / 64];
  __shared__ int s_base;
  const int b = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) s_base = 0;
  for (int i0 = 0; i0 < I; i0 += kCompThreads) {
    const int i = i0 + tid;
    const size_t row = ((size_t)b * I + (i < I ? i : 0)) * K_pre;
    const int np = i < I ? leading_finite(pre_t + row, K_pre) : 0;
    int incl = np;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(incl, d, 64);
      if (lane >= d) incl += y;
    }
    __syncthreads();  // s_base of the previous chunk is final; s_wave is free
    if (lane == 63) s_wave[wv] = incl;
    __syncthreads();
    int off = s_base + incl - np;
    for (int w = 0; w < wv; ++w) off += s_wave[w];
    float4 *o = spk + (size_t)b * cap + off;
    int *os = slot + (size_t)b * cap + off;
    for (int m = 0; m < np; ++m) {
      o[m] = make_float4(pre_t[row + m], pre_es[row + m], pre_e[row + m], __int_as_float(i));
      os[m] = i * K_pre + m;
    }
    __syncthreads();
    if (tid == kCompThreads - 1) {
      int tot = 0;
      for (int w = 0; w < kCompThreads / 64; ++w) tot += s_wave[w];
      s_base += tot;
    }
  }
  __syncthreads();
  if (tid == 0) count[b] = s_base;
}

// The same list in rank-major order: every neuron's first spike (in neuron order), then every
// second spike, and so on, so that the 64 spikes of a wave's chunk belong to distinct neurons
// (the weight-gradient kernel below adds them to per-neuron LDS addresses: lanes that share an
// address serialise). n_i is kept in LDS (I ints), and each rank is one pass of
// wave-aggregated block scans over the neurons.
__global__ __launch_bounds__(kCompThreads) void compact_spikes_rank_kernel(
    int I, int K_pre, const float *__restrict__ pre_t, const float *__restrict__ pre_es,
    const float *__restrict__ pre_e, int cap, float4 *__restrict__ spk, int *__restrict__ slot,
    int *__restrict__ count) {
  extern __shared__ int s_n[];  // [I]
  __shared__ int s_wave[kCompThreads / 64];
  __shared__ int s_max;
  const int b = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) s_max = 0;
  __syncthreads();
  int mx = 0;
  for (int i = tid; i < I; i += kCompThreads) {
    const int n = leading_finite(pre_t + ((size_t)b * I + i) * K_pre, K_pre);
    s_n[i] = n;
    mx = max(mx, n);
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) mx = max(mx, __shfl_xor(mx, d, 64));
  if (lane == 0) atomicMax(&s_max, mx);
  __syncthreads();
  const int maxn = s_max;
  float4 *o = spk + (size_t)b * cap;
  int *os = slot + (size_t)b * cap;
  int base = 0;
  for (int m = 0; m < maxn; ++m) {
    for (int i0 = 0; i0 < I; i0 += kCompThreads) {
      const int i = i0 + tid;
      const bool flag = i < I && s_n[i] > m;
      const unsigned long long bal = __ballot(flag);
      if (lane == 0) s_wave[wv] = __popcll(bal);
      __syncthreads();
      int off = base + (int)__popcll(bal & ((1ull << lane) - 1ull)), total = 0;
#pragma unroll
      for (int w = 0; w < kCompThreads / 64; ++w) {
        off += w < wv ? s_wave[w] : 0;
        total += s_wave[w];
      }
      if (flag) {
        const size_t row = ((size_t)b * I + i) * K_pre + m;
        o[off] = make_float4(pre_t[row], pre_es[row], pre_e[row], __int_as_float(i));
        os[off] = i * K_pre + m;
      }
      base += total;
      __syncthreads();  // s_wave is read by every wave before the next chunk writes it
    }
  }
  if (tid == 0) count[b] = base;
}

// Weight gradient with a lane per pre spike: one workgroup per (tile of kWT post rows, batch
// slice) accumulates acc[row][pre neuron] in float64 in LDS. Per sample the tile's record
// (row_prep_kernel) is copied to LDS, and each wave takes 64-spike chunks of the sample's
// compacted pre spikes and walks the record, adding es d1 - e d2 (t_pre < t_post) with a
// no-return LDS atomic. With the rank-major list the lanes of a chunk belong to distinct
// neurons except where a rank ends inside the chunk; lanes of one neuron, and waves, may still
// add to one address, so the float64 sums are exact up to the order of those additions.
// 24-bit unsigned product (full rate; v_mul_lo_u32 runs at a quarter of it). Operands < 2^24.
__device__ __forceinline__ unsigned mul_u24(unsigned a, unsigned b) {
  unsigned r;
  asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "s"(b));
  return r;
}

__host__ __device__ inline size_t spk_wgrad_lds_bytes(int I, int K) {
  return (size_t)kWT * I * sizeof(double) + (size_t)tile_rec_slots(K) * sizeof(float4);
}

__global__ __launch_bounds__(1024) void wgrad_spk_kernel(
    int B, int I, int J, int K, int gj, int b_per_slice, int cap, const float4 *__restrict__ spk,
    const int *__restrict__ count, const float4 *__restrict__ rec,
    float *__restrict__ partial) {
  extern __shared__ __align__(16) char smem[];
  double *acc = reinterpret_cast<double *>(smem);  // [kWT][I]
  float4 *rb = reinterpret_cast<float4 *>(acc + (size_t)kWT * I);
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), nw = nth >> 6;
  const int tj = blockIdx.x, sl = blockIdx.y;
  const int j0 = tj * kWT, jn = min(kWT, J - j0);
  const int b0 = sl * b_per_slice, b1 = min(B, b0 + b_per_slice);
  const int R = tile_rec_slots(K);
  for (int e = tid; e < kWT * I; e += nth) acc[e] = 0.0;
  // the next sample's spike count, record (the thread's entry tid, up to nth entries), pre
  // spike count and the wave's first chunk of its pre spikes, in registers one sample ahead
  float4 rn = make_float4(0.f, 0.f, 0.f, 0.f), Sn = make_float4(kInfF, 0.f, 0.f, 0.f);
  int cn = 0, Pn = 0;
  auto fetch = [&](int b) {
    const float4 *grec = rec + ((size_t)b * gj + tj) * R;
    cn = __float_as_int(grec[0].x);
    if (tid < R - kRecHead) rn = grec[kRecHead + tid];
    Pn = count[b];
    const int sidx = wv * 64 + lane;
    Sn = sidx < Pn ? spk[(size_t)b * cap + sidx] : make_float4(kInfF, 0.f, 0.f, 0.f);
  };
  if (b0 < b1) fetch(b0);
  for (int b = b0; b < b1; ++b) {
    const float4 *grec = rec + ((size_t)b * gj + tj) * R;
    __syncthreads();  // the previous record is consumed (and acc is zeroed)
    const int cnt = cn, P = Pn;
    const float4 Sfirst = Sn;
    if (tid < cnt) rb[tid] = rn;
    for (int e = tid + nth; e < cnt; e += nth) rb[e] = grec[kRecHead + e];
    if (b + 1 < b1) fetch(b + 1);
    __syncthreads();
    for (int c = wv; c * 64 < P; c += nw) {
      const int sidx = c * 64 + lane;
      const bool valid = sidx < P;
      const float4 S = c == wv ? Sfirst
                               : (valid ? spk[(size_t)b * cap + sidx]
                                        : make_float4(kInfF, 0.f, 0.f, 0.f));
      double *al = acc + (valid ? __float_as_int(S.w) : 0);
      // entries eight, then four at a time (their LDS reads in one round). The record lists
      // the tile's spikes row by row and every lane reads the same entry, so a row change is
      // uniform: a lane sums its terms of one row in a register (a pair adds only where the
      // causal mask holds; a skipped term is an exact zero) and adds the row's sum to acc
      // with one no-return LDS atomic when the row ends, if it is not zero
      double racc = 0.0;
      int crow = __builtin_amdgcn_readfirstlane(cnt > 0 ? __float_as_int(rb[0].w) >> 9 : 0);
      auto flush = [&]() {
        if (racc != 0.0)
          __hip_atomic_fetch_add(al + mul_u24((unsigned)crow, (unsigned)I), racc,
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        racc = 0.0;
      };
      auto term = [&](const float4 E) {
        const int r = __builtin_amdgcn_readfirstlane(__float_as_int(E.w) >> 9);
        if (r != crow) {
          flush();
          crow = r;
        }
        const float q = fmaf(S.y, E.y, -(S.z * E.z));
        if (S.x < E.x) racc += (double)q;
      };
      int k = 0;
      for (; k + 8 <= cnt; k += 8) {
        float4 E[8];
#pragma unroll
        for (int g = 0; g < 8; ++g) E[g] = rb[k + g];
#pragma unroll
        for (int g = 0; g < 8; ++g) term(E[g]);
      }
      for (; k + 4 <= cnt; k += 4) {
        const float4 E0 = rb[k], E1 = rb[k + 1], E2 = rb[k + 2], E3 = rb[k + 3];
        term(E0);
        term(E1);
        term(E2);
        term(E3);
      }
      for (; k < cnt; ++k) term(rb[k]);
      flush();
    }
  }
  __syncthreads();
  for (int e = tid; e < jn * I; e += nth) {
    const int u = e / I, i = e - u * I;
    partial[((size_t)sl * J + j0 + u) * I + i] = (float)acc[e];
  }
}

// Pre-spike errors with a lane per pre spike: a workgroup of kPreSpkWaves waves per (512 pre
// spikes, sample) walks all tile records of the sample; per tile each lane stages its 16
// weights W[row][neuron] in the wave's LDS block, and every entry adds
//   w[row] * (es' d1 - e' d2)  if t_pre < t_post   (float32 in the tile, float64 across).
// pre_errors must be zero-filled beforehand (slots without a spike keep 0).
constexpr int kPreSpkWaves = 8;  // waves (64 pre spikes each) sharing one tile record

__global__ __launch_bounds__(64 * kPreSpkWaves) void pre_err_spk_kernel(
    int I, int K_pre, int J, int K, int gj, int cap, const float4 *__restrict__ spk,
    const int *__restrict__ slot, const int *__restrict__ count, int count_stride,
    const float *__restrict__ W,
    const float4 *__restrict__ rec, double inv_tau_s, double inv_tau,
    float *__restrict__ pre_errors) {
  constexpr int kNth = 64 * kPreSpkWaves;
  constexpr int kPf = 1;  // record entries per thread prefetched for the next tile
  extern __shared__ __align__(16) char smem[];
  float *wl = reinterpret_cast<float *>(smem);  // [kPreSpkWaves][kWT][64]
  float4 *rb = reinterpret_cast<float4 *>(wl + kPreSpkWaves * kWT * 64);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int b = blockIdx.y;
  const int P = count[(size_t)b * count_stride];
  const int s0 = blockIdx.x * kNth;
  if (s0 >= P) return;  // whole block
  const int sidx = s0 + tid;
  const bool valid = sidx < P;
  const float4 S = valid ? spk[(size_t)b * cap + sidx] : make_float4(kInfF, 0.f, 0.f, 0.f);
  // the neuron index is the low 24 bits (the binned list keeps the spike's bin above them)
  const int i = valid ? (__float_as_int(S.w) & 0xffffff) : 0;
  const float esd = (float)((double)S.y * inv_tau_s), ed = (float)((double)S.z * inv_tau);
  float *ww = wl + wv * kWT * 64;
  const int R = tile_rec_slots(K);
  // the next tile's spike count, record entries (tid, tid + kNth) and weights, requested one
  // tile ahead
  float4 rn[kPf];
  float wn[kWT];
  int cn = 0;
  auto prefetch = [&](int tj) {
    cn = __float_as_int(rec[((size_t)b * gj + tj) * R].x);
    const float4 *grec = rec + ((size_t)b * gj + tj) * R + kRecHead;
#pragma unroll
    for (int p = 0; p < kPf; ++p) {
      const int e = tid + p * kNth;
      rn[p] = e < R - kRecHead ? grec[e] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const int j0 = tj * kWT;
#pragma unroll
    for (int u = 0; u < kWT; ++u)
      wn[u] = (valid && j0 + u < J) ? W[(size_t)(j0 + u) * I + i] : 0.f;
  };
  prefetch(0);
  double pe = 0.0;
  for (int tj = 0; tj < gj; ++tj) {
    const float4 *grec = rec + ((size_t)b * gj + tj) * R;
    __syncthreads();  // the previous tile is consumed
    const int cnt = cn;
#pragma unroll
    for (int p = 0; p < kPf; ++p)
      if (tid + p * kNth < cnt) rb[tid + p * kNth] = rn[p];
    for (int e = tid + kPf * kNth; e < cnt; e += kNth) rb[e] = grec[kRecHead + e];
#pragma unroll
    for (int u = 0; u < kWT; ++u) ww[u * 64 + lane] = wn[u];
    if (tj + 1 < gj) prefetch(tj + 1);
    __syncthreads();
    float pt = 0.f;
    auto term = [&](const float4 E) {
      const float w = ww[(__float_as_int(E.w) >> 9) * 64 + lane];
      const float q = fmaf(esd, E.y, -(ed * E.z));
      pt = fmaf((S.x < E.x) ? w : 0.f, q, pt);
    };
    int k = 0;
    for (; k + 4 <= cnt; k += 4) {
      const float4 E0 = rb[k], E1 = rb[k + 1], E2 = rb[k + 2], E3 = rb[k + 3];
      term(E0);
      term(E1);
      term(E2);
      term(E3);
    }
    for (; k < cnt; ++k) term(rb[k]);
    pe += (double)pt;
  }
  if (valid) pre_errors[(size_t)b * I * K_pre + slot[(size_t)b * cap + sidx]] = (float)pe;
}

// ------------------------------------------------------------------ narrow layers
// Layers with at most kST post neurons (an output layer): one block per (pre-neuron block,
// sample). The sample's raw post rows are copied to LDS, one thread per row runs the reset
// chain, the factors and the suffix sums (the arithmetic of row_prep_kernel), and every
// lane then forms, for each of its pre spikes, the weight-gradient terms of all post neurons
// and -- from the same lookups -- the pre-spike error sum_j W (V1 es/tau_s - V2 e/tau).
// The per-sample weight-gradient rows are reduced over the batch by wgrad_reduce_kernel.
constexpr int kST = 16;

struct DenseStage {
  double *s1, *s2;            // [R][K+1]; slot ne holds 0
  float *t;                   // [R][KT] (KT = K rounded up to 4: 16-byte rows)
  float *a, *x, *z, *er;      // [R][K]
  int *nn, *ne;               // [R]
};

__host__ __device__ inline int stage_row_stride(int K) { return (K + 3) & ~3; }

__host__ __device__ inline size_t dense_stage_bytes(int R, int K) {
  return (size_t)R * (K + 1) * 16 + (size_t)R * stage_row_stride(K) * 4 +
         (size_t)R * K * 16 + (size_t)R * 8;
}

__device__ __forceinline__ DenseStage dense_stage(char *smem, int R, int K) {
  DenseStage d;
  d.s1 = reinterpret_cast<double *>(smem);
  d.s2 = d.s1 + (size_t)R * (K + 1);
  d.t = reinterpret_cast<float *>(d.s2 + (size_t)R * (K + 1));
  d.a = d.t + (size_t)R * stage_row_stride(K);
  d.x = d.a + (size_t)R * K;
  d.z = d.x + (size_t)R * K;
  d.er = d.z + (size_t)R * K;
  d.nn = reinterpret_cast<int *>(d.er + (size_t)R * K);
  d.ne = d.nn + R;
  return d;
}

// Number of the first nj times of an ascending LDS row (16-byte aligned) that are <= tp:
// wave-uniform float4 reads, independent of each other.
__device__ __forceinline__ int count_le(const float *__restrict__ row, int nj, float tp) {
  const float4 *r4 = reinterpret_cast<const float4 *>(row);
  int kk = 0;
#pragma unroll 2
  for (int q = 0; q < nj; q += 4) {
    const float4 v = r4[q >> 2];
    kk += (v.x <= tp) ? 1 : 0;
    kk += (q + 1 < nj && v.y <= tp) ? 1 : 0;
    kk += (q + 2 < nj && v.z <= tp) ? 1 : 0;
    kk += (q + 3 < nj && v.w <= tp) ? 1 : 0;
  }
  return kk;
}

// kList (the whole sample's pre neurons in one block): the sample's pre spikes are numbered
// by a block scan of the per-neuron counts and each thread takes spikes tid, tid + nth, ...
// (about one each), adding its weight-gradient terms to float64 accumulators [J][I] in LDS
// with no-return atomics; without it a thread takes every spike of its own pre neuron, and
// the busiest pre neuron of the sample sets the block's length.
template <bool kPre, bool kList>
__global__ __launch_bounds__(1024) void small_bwd_kernel(
    int I, int K_pre, int J, int K, const int *__restrict__ n, const float *__restrict__ t,
    const float *__restrict__ a, const float *__restrict__ x, const float *__restrict__ z,
    float *__restrict__ errors, int write_errors, double c, double tau, double dtt,
    double inv_tau_s, double inv_tau, const float *__restrict__ pre_t,
    const float *__restrict__ pre_es, const float *__restrict__ pre_e,
    const float *__restrict__ W, float *__restrict__ out, float *__restrict__ pre_errors) {
  extern __shared__ __align__(16) char smem[];
  const int tid = threadIdx.x, nth = blockDim.x;
  const int i = blockIdx.x * nth + tid;
  const bool active = i < I;
  const int ic = active ? i : 0;
  const int b = blockIdx.y;
  const DenseStage st = dense_stage(smem, J, K);
  // the lane's pre spikes: count, and the first one (loads in flight with the row copy)
  const size_t pbase = ((size_t)b * I + ic) * K_pre;
  const int np = active ? leading_finite(pre_t + pbase, K_pre) : 0;
  float tp0 = kInfF, es0 = 0.f, e0 = 0.f;
  if (np > 0) {
    tp0 = pre_t[pbase];
    es0 = pre_es[pbase];
    e0 = pre_e[pbase];
  }
  float wv[kST];
#pragma unroll
  for (int u = 0; u < kST; ++u) wv[u] = (kPre && active && u < J) ? W[(size_t)u * I + i] : 0.f;
  // the sample's pre-spike errors start at zero, written coalesced (a lane-per-row zero fill
  // would touch a separate cache line per lane and store); the barrier below orders these
  // stores before the values written at the end
  if (kPre) {
    const size_t pb0 = (size_t)b * I * K_pre + (size_t)blockIdx.x * nth * K_pre;
    const int span = min(nth, I - (int)blockIdx.x * nth) * K_pre;
    for (int e = tid; e < span; e += nth) pre_errors[pb0 + e] = 0.f;
  }
  // 1. the sample's rows
  const size_t rbase = (size_t)b * J * K;
  const int KT = stage_row_stride(K);
  for (int e = tid; e < J * K; e += nth) {
    const int r = e / K;
    st.t[r * KT + (e - r * K)] = t[rbase + e];
    st.a[e] = a[rbase + e];
    st.x[e] = x[rbase + e];
    st.z[e] = z[rbase + e];
    st.er[e] = errors[rbase + e];
  }
  if (tid < J) st.nn[tid] = n[(size_t)b * J + tid];
  __syncthreads();
  // 2. one thread per row
  if (tid < J) {
    const int r = tid;
    const int nk = min(st.nn[r], K);
    float *er = st.er + r * K;
    const float *tr = st.t + r * KT, *ar = st.a + r * K, *xr = st.x + r * K, *zr = st.z + r * K;
    double C = 0.0;
    for (int k = nk - 1; k >= 0; --k) {
      const float zk = zr[k];
      if (isinf(zk)) continue;
      const double ek = (double)er[k] + C * (double)zk;
      const float ek32 = (float)ek;
      er[k] = ek32;
      if (write_errors) errors[rbase + (size_t)r * K + k] = ek32;
      C = C + (double)ek32 * dtt / (double)xr[k];
    }
    int ne = 0;
    while (ne < nk && !isinf(tr[ne])) ++ne;
    double *s1 = st.s1 + (size_t)r * (K + 1);
    double *s2 = st.s2 + (size_t)r * (K + 1);
    s1[ne] = 0.0;
    s2[ne] = 0.0;
    double a1 = 0.0, a2 = 0.0;
    for (int k = ne - 1; k >= 0; --k) {
      const double ak = (double)ar[k], xk = (double)xr[k], zk = (double)zr[k];
      const float f1 = (float)(tau / ak * (1.0 + zk * c / xk));
      const float f2 = (float)(tau / xk);
      const double ek = (double)er[k];
      a1 += ek * (double)f1;
      a2 += ek * (double)f2;
      s1[k] = a1;
      s2[k] = a2;
    }
    st.ne[r] = ne;
  }
  if (kList) {
    // 3'. spikes numbered by a block scan of the per-neuron counts; one spike per thread pass
    double *accl = reinterpret_cast<double *>(smem + ((dense_stage_bytes(J, K) + 15) & ~(size_t)15));
    int *sbase = reinterpret_cast<int *>(accl + (size_t)J * I);  // [I + 1]
    int *swave = sbase + I + 1;                                  // [32]
    for (int e = tid; e < J * I; e += nth) accl[e] = 0.0;
    const int lane = tid & 63, wv = tid >> 6, nwv = nth >> 6;
    int incl = np;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(incl, d, 64);
      if (lane >= d) incl += y;
    }
    if (lane == 63) swave[wv] = incl;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wv; ++w) woff += swave[w];
    int total = 0;
    for (int w = 0; w < nwv; ++w) total += swave[w];
    if (active) sbase[i] = woff + incl - np;
    if (tid == 0) sbase[I] = total;
    __syncthreads();
    for (int e = tid; e < total; e += nth) {
      // the spike's neuron: the last i with sbase[i] <= e (neurons without spikes repeat a
      // base, the search lands on the one whose range holds e)
      int lo = 0, hi = I - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sbase[mid] <= e) lo = mid; else hi = mid - 1;
      }
      const int si = lo, m = e - sbase[si];
      const size_t sb = ((size_t)b * I + si) * K_pre + m;
      const float tp = pre_t[sb];
      const double es = (double)pre_es[sb], ee = (double)pre_e[sb];
      float ws[kST];
#pragma unroll
      for (int u = 0; u < kST; ++u) ws[u] = (kPre && u < J) ? W[(size_t)u * I + si] : 0.f;
      double pacc = 0.0;
#pragma unroll
      for (int u = 0; u < kST; ++u) {
        if (u < J) {
          const int nj = __builtin_amdgcn_readfirstlane(st.ne[u]);
          if (nj > 0) {
            const int kk = count_le(st.t + u * KT, nj, tp);
            const double v1 = st.s1[(size_t)u * (K + 1) + kk];
            const double v2 = st.s2[(size_t)u * (K + 1) + kk];
            __hip_atomic_fetch_add(accl + (size_t)u * I + si, fma(-ee, v2, es * v1),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (kPre) pacc += (double)ws[u] * (v1 * es * inv_tau_s - v2 * ee * inv_tau);
          }
        }
      }
      if (kPre) pre_errors[sb] = (float)pacc;
    }
    __syncthreads();
    for (int e = tid; e < J * I; e += nth) {
      const int u = e / I, ii = e - u * I;
      out[((size_t)b * J + u) * I + ii] = (float)accl[e];
    }
    return;
  }
  __syncthreads();
  // 3. lanes
  double acc[kST];
#pragma unroll
  for (int u = 0; u < kST; ++u) acc[u] = 0.0;
  // the lane's next pre spike is requested while the current one is processed
  float tpn = tp0, esn = es0, en = e0;
  for (int m = 0; m < np; ++m) {
    const float tp = tpn;
    const double es = (double)esn, e = (double)en;
    if (m + 1 < np) {
      tpn = pre_t[pbase + m + 1];
      esn = pre_es[pbase + m + 1];
      en = pre_e[pbase + m + 1];
    }
    double pacc = 0.0;
#pragma unroll
    for (int u = 0; u < kST; ++u) {
      if (u < J) {
        const int nj = __builtin_amdgcn_readfirstlane(st.ne[u]);
        if (nj > 0) {
          const int kk = count_le(st.t + u * KT, nj, tp);
          const double v1 = st.s1[(size_t)u * (K + 1) + kk];
          const double v2 = st.s2[(size_t)u * (K + 1) + kk];
          acc[u] = fma(es, v1, acc[u]);
          acc[u] = fma(-e, v2, acc[u]);
          if (kPre) pacc += (double)wv[u] * (v1 * es * inv_tau_s - v2 * e * inv_tau);
        }
      }
    }
    if (kPre) pre_errors[pbase + m] = (float)pacc;
  }
  if (active) {
#pragma unroll
    for (int u = 0; u < kST; ++u)
      if (u < J) out[((size_t)b * J + u) * I + i] = (float)acc[u];
  }
}

struct BwdLayout {
  size_t rec, spk, slot, count, partial, total;
  int slices, b_per_slice, lanes, gi;
  int spk_slices, spk_b_per_slice;  // the lane-per-pre-spike weight gradient's batch slices
  bool small, spk_path;
};

// LDS a workgroup may use (the spike-lane weight gradient needs kWT x I float64 accumulators)
constexpr size_t kMaxWorkgroupLds = 160 * 1024 - 1024;

// Waves the weight-gradient grid aims at: about ten per SIMD of the 256-CU device, so that
// the scalar-load latency of one wave is covered by the others and the last round is short.
// More slices cost a longer slice reduction; on the 784->800 layer at batch 256 the whole
// backward measured 131.7 us at six waves per SIMD, 123.2 us at ten and 130.1 us at 26.
constexpr int kWgradTargetWaves = 10240;

static BwdLayout bwd_layout(int B, int I, int K_pre, int J, int K, bool per_sample = false) {
  BwdLayout L{};
  L.gi = std::max(1, (I + 1023) / 1024);
  const int per_block = (I + L.gi - 1) / L.gi;
  L.lanes = std::max(64, (per_block + 63) / 64 * 64);
  L.small = J <= kST && dense_stage_bytes(J, K) <= 64 * 1024;
  if (L.small) {
    L.slices = std::max(B, 1);
    L.b_per_slice = 1;
  } else {
    const int waves = std::max(1, ((I + 63) / 64) * ((J + kWT - 1) / kWT));
    int slices = per_sample ? B : (kWgradTargetWaves + waves - 1) / waves;
    slices = std::max(1, std::min(std::max(B, 1), slices));
    L.b_per_slice = B > 0 ? (B + slices - 1) / slices : 1;
    L.slices = B > 0 ? (B + L.b_per_slice - 1) / L.b_per_slice : 1;
  }
  // multi-spike pre layers: compacted pre spikes, and (when its tile fits in LDS) the
  // spike-lane weight gradient over ~2 workgroups per CU
  const bool multi = !L.small && K_pre > 1;
  L.spk_path = multi && !per_sample && spk_wgrad_lds_bytes(I, K) <= kMaxWorkgroupLds;
  if (L.spk_path) {
    const int gj = (J + kWT - 1) / kWT;
    const int sl = std::max(1, std::min(std::max(B, 1), (512 + gj - 1) / gj));
    L.spk_b_per_slice = B > 0 ? (B + sl - 1) / sl : 1;
    L.spk_slices = B > 0 ? (B + L.spk_b_per_slice - 1) / L.spk_b_per_slice : 1;
  }
  const size_t cap = (size_t)I * K_pre;
  size_t off = 0;
  L.spk = off;
  if (multi) off = align_up(off + sizeof(float4) * (size_t)B * cap, 256);
  L.slot = off;
  if (multi) off = align_up(off + sizeof(int) * (size_t)B * cap, 256);
  L.count = off;
  if (multi) off = align_up(off + sizeof(int) * (size_t)B, 256);
  L.rec = off;
  if (!L.small)
    off = align_up(off + sizeof(float4) * (size_t)B * ((J + kWT - 1) / kWT) * tile_rec_slots(K),
                   256);
  L.partial = off;
  off = align_up(off + sizeof(float) * (size_t)std::max(L.slices, L.spk_path ? L.spk_slices : 0) *
                           J * I,
                 256);
  L.total = off;
  return L;
}

// Dynamic LDS above the default 64 KB needs the kernel's attribute raised first.
static int allow_lds(const void *kernel, size_t bytes) {
  if (bytes <= 64 * 1024) return 0;
  if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) !=
      hipSuccess)
    return fail("backward: LDS request above the device limit");
  return 0;
}

// Compacted pre spikes of every sample in neuron order (compact_spikes_kernel), for the conv
// backward.
int compact_pre_spikes(hipStream_t st, int B, int I, int K_pre, const float *pre_t,
                       const float *pre_es, const float *pre_e, float4 *spk, int *slot,
                       int *count) {
  hipLaunchKernelGGL(compact_spikes_kernel, dim3(B), dim3(kCompThreads), 0, st, I, K_pre, pre_t,
                     pre_es, pre_e, I * K_pre, spk, slot, count);
  return check_launch("compact_spikes_kernel");
}

int post_prep(hipStream_t st, int BJ, int K, const int *n, const float *t, const float *a,
              const float *x, const float *z, float *errors, float c, float tau, float dtt,
              int *n_eff, float2 *D) {
  return launch_row_prep(st, 1, BJ, K, n, t, a, x, z, errors, (double)c, (double)tau,
                         (double)dtt, n_eff, D, nullptr);
}

}  // namespace bats

using namespace bats;

extern "C" size_t bats_lif_backward_workspace_size(int B, int I, int K_pre, int J, int K_post) {
  if (B < 0 || I < 0 || K_pre < 0 || J < 0 || K_post < 0) return 0;
  size_t total = bwd_layout(B, I, K_pre, J, K_post).total;
  // the binned path (batch-reduced gradients of multi-spike pre layers) has its own layout
  if (K_post >= 1 && bin_path_applies(I, K_pre, J, K_post, 0))
    total = std::max(total, bin_layout(B, I, K_pre, J, K_post).total);
  return total;
}

// Largest post-spike buffer the backward supports (LDS row stages of the narrow-layer and
// pre-spike error kernels).
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
  BATS_REQUIRE(B <= 65535, "batch above 65535");
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
  float4 *rec = reinterpret_cast<float4 *>(ws + L.rec);
  float4 *spk = reinterpret_cast<float4 *>(ws + L.spk);
  int *spk_slot = reinterpret_cast<int *>(ws + L.slot);
  int *spk_count = reinterpret_cast<int *>(ws + L.count);
  float *partial = reinterpret_cast<float *>(ws + L.partial);
  // the reset chain alone, written back in place
  auto reset_only = [&]() -> int {
    return launch_row_prep(st, B, J, K_post, n, t, a, x, z, errors, (double)c, (double)tau,
                           (double)delta_theta_tau, nullptr, nullptr, nullptr);
  };
  if (JI == 0) return reset_only();  // no weights: only the reset chain of the errors
  const bool want_pre = pre_errors && K_pre > 0;
  if (want_pre) BATS_REQUIRE(weights, "null weights");
  BATS_REQUIRE(K_pre == 0 || (pre_times && pre_exp_tau_s && pre_exp_tau), "null pre pointer");
  const double inv_b = grad_mode == 2 ? 1.0 : 1.0 / (double)B;

  if (L.small) {
    float *out = grad_mode == 1 ? grad : partial;
    // one block per sample: number the sample's pre spikes and accumulate in LDS when the
    // float64 accumulators fit
    const size_t stage = dense_stage_bytes(J, K_post);
    const size_t lds_list = ((stage + 15) & ~(size_t)15) + (size_t)J * I * sizeof(double) +
                            (size_t)(I + 1 + 32) * sizeof(int);
    const bool list = L.gi == 1 && lds_list <= kMaxWorkgroupLds;
    const size_t lds = list ? lds_list : stage;
    auto kern = list ? (want_pre ? small_bwd_kernel<true, true> : small_bwd_kernel<false, true>)
                     : (want_pre ? small_bwd_kernel<true, false> : small_bwd_kernel<false, false>);
    if (int r = allow_lds((const void *)kern, lds)) return r;
    BATS_TIMED("small_bwd_kernel", st,
               hipLaunchKernelGGL(kern, dim3(L.gi, B), dim3(L.lanes), lds, st, I, K_pre, J,
                                  K_post, n, t, a, x, z, errors, L.gi == 1 ? 1 : 0, (double)c,
                                  (double)tau, (double)delta_theta_tau, 1.0 / (double)tau_s,
                                  1.0 / (double)tau, pre_times, pre_exp_tau_s, pre_exp_tau,
                                  weights, out, want_pre ? pre_errors : nullptr));
    if (int r = check_launch("small_bwd_kernel")) return r;
    if (grad_mode != 1)
      if (int r = reduce_slices(st, B, JI, inv_b, partial, grad)) return r;
    // with several blocks of pre neurons every block read the raw errors; write the reset
    // chain back now
    if (L.gi > 1) return reset_only();
    return 0;
  }

  int prep_rc = 0;
  if (bin_path_applies(I, K_pre, J, K_post, grad_mode)) {
    // multi-spike pre layer: binned GEMMs on the matrix cores plus the in-bin pairs
    const BinLayout Lb = bin_layout(B, I, K_pre, J, K_post);
    BATS_REQUIRE(workspace_bytes >= Lb.total, "workspace too small");
    BATS_REQUIRE((int64_t)I * K_pre < (int64_t)1 << 31, "pre spike buffer too large");
    BATS_TIMED("row_prep_kernel", st,
               prep_rc = launch_row_prep(st, B, J, K_post, n, t, a, x, z, errors, (double)c,
                                         (double)tau, (double)delta_theta_tau,
                                         reinterpret_cast<int *>(ws + Lb.n_eff),
                                         reinterpret_cast<float2 *>(ws + Lb.D), nullptr));
    if (prep_rc) return prep_rc;
    if (want_pre) {
      const int64_t nfill = (int64_t)B * I * K_pre;
      hipLaunchKernelGGL(fill_kernel<float>,
                         dim3((unsigned)std::min<int64_t>((nfill + 255) / 256, 65535)), dim3(256),
                         0, st, pre_errors, nfill, 0.f);
      if (int r = check_launch("fill pre errors")) return r;
    }
    if (int r = bins_backward(st, Lb, B, I, K_pre, J, K_post, t, pre_times, pre_exp_tau_s,
                              pre_exp_tau, weights, 1.0 / (double)tau_s, 1.0 / (double)tau,
                              want_pre ? pre_errors : nullptr, ws))
      return r;
    return reduce_slices(st, Lb.gemm_splits + Lb.walk_slices, JI, inv_b,
                         reinterpret_cast<const float *>(ws + Lb.partial), grad);
  }
  BATS_TIMED("row_prep_kernel", st,
             prep_rc = launch_row_prep(st, B, J, K_post, n, t, a, x, z, errors, (double)c,
                                       (double)tau, (double)delta_theta_tau, nullptr,
                                       nullptr, rec));
  if (prep_rc) return prep_rc;
  const bool multi = K_pre > 1;
  const int cap = I * K_pre;
  if (multi) {
    BATS_REQUIRE((int64_t)I * K_pre < (int64_t)1 << 31, "pre spike buffer too large");
    // rank-major order when the per-neuron counts fit in LDS (every caller here reads the
    // list through `slot` or per spike, so the order is free)
    const bool rank = (size_t)I * sizeof(int) <= 64 * 1024;
    BATS_TIMED("compact_spikes_kernel", st, {
      if (rank)
        hipLaunchKernelGGL(compact_spikes_rank_kernel, dim3(B), dim3(kCompThreads),
                           sizeof(int) * (size_t)I, st, I, K_pre, pre_times, pre_exp_tau_s,
                           pre_exp_tau, cap, spk, spk_slot, spk_count);
      else
        hipLaunchKernelGGL(compact_spikes_kernel, dim3(B), dim3(kCompThreads), 0, st, I, K_pre,
                           pre_times, pre_exp_tau_s, pre_exp_tau, cap, spk, spk_slot,
                           spk_count);
    });
    if (int r = check_launch("compact_spikes_kernel")) return r;
  }
  if (L.spk_path && grad_mode != 1) {
    const int gj = (J + kWT - 1) / kWT;
    const size_t lds = spk_wgrad_lds_bytes(I, K_post);
    if (int r = allow_lds((const void *)wgrad_spk_kernel, lds)) return r;
    BATS_TIMED("wgrad_spk_kernel", st,
               hipLaunchKernelGGL(wgrad_spk_kernel, dim3(gj, L.spk_slices), dim3(1024), lds, st,
                                  B, I, J, K_post, gj, L.spk_b_per_slice, cap, spk, spk_count,
                                  rec, partial));
    if (int r = check_launch("wgrad_spk_kernel")) return r;
    if (int r = reduce_slices(st, L.spk_slices, JI, inv_b, partial, grad)) return r;
  } else {
    const BwdLayout F = bwd_layout(B, I, K_pre, J, K_post, grad_mode == 1);
    const int gi = (I + 63) / 64, gj = (J + kWT - 1) / kWT;
    const int64_t ntasks = (int64_t)gi * gj * F.slices;
    BATS_REQUIRE(ntasks < (int64_t)1 << 31, "weight-gradient grid too large");
    const int64_t nblk = (ntasks + kTileWaves - 1) / kTileWaves;
    auto kern = K_pre == 1 ? wgrad_tile_kernel<true> : wgrad_tile_kernel<false>;
    BATS_TIMED("wgrad_tile_kernel", st,
               hipLaunchKernelGGL(kern, dim3((unsigned)nblk), dim3(64 * kTileWaves), 0, st, B,
                                  I, K_pre, J, K_post, gi, gj, F.b_per_slice, (int)ntasks,
                                  pre_times, pre_exp_tau_s,
                                  pre_exp_tau, rec, grad_mode == 1 ? 1 : 0, grad,
                                  partial));
    if (int r = check_launch("wgrad_tile_kernel")) return r;
    if (grad_mode != 1)
      if (int r = reduce_slices(st, F.slices, JI, inv_b, partial, grad)) return r;
  }
  if (want_pre && multi) {
    const int64_t nfill = (int64_t)B * I * K_pre;
    hipLaunchKernelGGL(fill_kernel<float>,
                       dim3((unsigned)std::min<int64_t>((nfill + 255) / 256, 65535)), dim3(256), 0,
                       st, pre_errors, nfill, 0.f);
    if (int r = check_launch("fill pre errors")) return r;
    const int gj = (J + kWT - 1) / kWT;
    const size_t lds = (size_t)kPreSpkWaves * kWT * 64 * sizeof(float) +
                       (size_t)tile_rec_slots(K_post) * sizeof(float4);
    if (int r = allow_lds((const void *)pre_err_spk_kernel, lds)) return r;
    const int groups = (cap + 64 * kPreSpkWaves - 1) / (64 * kPreSpkWaves);
    BATS_TIMED("pre_err_spk_kernel", st,
               hipLaunchKernelGGL(pre_err_spk_kernel, dim3(groups, B), dim3(64 * kPreSpkWaves), lds,
                                  st, I, K_pre, J, K_post, gj, cap, spk, spk_slot, spk_count, 1,
                                  weights, rec, 1.0 / (double)tau_s, 1.0 / (double)tau,
                                  pre_errors));
    if (int r = check_launch("pre_err_spk_kernel")) return r;
  } else if (want_pre) {
    const int gi = (I + 63) / 64;
    const int64_t ntasks = (int64_t)gi * B;
    BATS_REQUIRE(ntasks < (int64_t)1 << 31, "pre-spike error grid too large");
    const int64_t nblk = (ntasks + kTileWaves - 1) / kTileWaves;
    BATS_TIMED("pre_err_tile_kernel", st,
               hipLaunchKernelGGL(pre_err_tile_kernel<kPreM>, dim3((unsigned)nblk),
                                  dim3(64 * kTileWaves), 0, st, B, I, K_pre, J, K_post, gi,
                                  (int)ntasks, pre_times, pre_exp_tau_s, pre_exp_tau, weights,
                                  rec, 1.0 / (double)tau_s, 1.0 / (double)tau, pre_errors));
    if (int r = check_launch("pre_err_tile_kernel")) return r;
  }
  return 0;
}
