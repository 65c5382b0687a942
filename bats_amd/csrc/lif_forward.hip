// Dense LIF layer forward for gfx950: per-sample event sort + exact multi-spike solver.
//
// Data flow of one call (all on the caller's stream, no host synchronisation):
//   1. transpose_kernel    W [J,I] -> W^T [I,J] so that the lanes of a wave (consecutive
//                          post neurons) read one input's weights with one coalesced load.
//   2. sort_events_kernel  one workgroup per sample: compacts the finite input slots,
//                          sorts (time, slot) keys in LDS and writes the sample's
//                          time-ordered Event list (exps in float32 exactly as stored for
//                          the backward, exp(t/tau) in float64 for the crossing filter).
//                          It also writes the per-slot exps the backward reads.
//   3. lif_solve_seq_kernel  one lane per (sample, post neuron): walks the sample's events
//                          in order, accumulates a and b in float64 (products of two
//                          float32 are exact in float64, so the running sums equal a
//                          float64 evaluation in the reference's order), evaluates a
//                          conservative crossing filter per event and runs the closed-form
//                          root (the reference's compute_spikes loop) in float64 only when
//                          the filter cannot exclude a crossing before the next event.
//   4. lif_solve_wave_kernel one wave per (sample, post neuron) for layers with few post
//                          neurons: the same decision sequence with the event axis spread
//                          over the 64 lanes (prefix sums by wave scan).
//
// Reference: bats/Layers/LIFLayer.py:63-91, bats/CudaKernels/CuFiles/Inference/
// compute_pre_exps.cu, get_spike_weights.cu, compute_spike_times.cu:34-138,
// bats/CudaKernels/Wrappers/Inference/get_sorted_spikes_indices.py:7-16.
#include "solver.hpp"

#include <algorithm>
#include <mutex>

namespace bats {

// ------------------------------------------------------------------ transpose
// W^T rows have Jp >= J entries (a multiple of 4); the Jp - J padding entries are written as 0.
__global__ __launch_bounds__(256) void transpose_kernel(const float *__restrict__ W,
                                                        float *__restrict__ WT, int J, int I,
                                                        int Jp) {
  __shared__ float tile[32][33];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int j = j0 + r, i = i0 + tx;
    tile[r][tx] = (j < J && i < I) ? W[(size_t)j * I + i] : 0.f;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + r, j = j0 + tx;
    if (j < Jp && i < I) WT[(size_t)i * Jp + j] = tile[tx][r];
  }
}

// ------------------------------------------------------------------ per-sample sort
constexpr int kSortThreads = 512;

__device__ __forceinline__ void bitonic_sort(uint64_t *keys, int N, bool global_mem) {
  for (int k = 2; k <= N; k <<= 1) {
    for (int jstride = k >> 1; jstride > 0; jstride >>= 1) {
      for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const int ixj = i ^ jstride;
        if (ixj > i) {
          const uint64_t x = keys[i], y = keys[ixj];
          const bool up = (i & k) == 0;
          if ((x > y) == up) {
            keys[i] = y;
            keys[ixj] = x;
          }
        }
      }
      if (global_mem) __threadfence();
      __syncthreads();
    }
  }
}

// Bitonic network over the 128 keys of one wave, two per lane (keys 2l and 2l+1): every
// compare-exchange partner is reached with a lane shuffle (no LDS, no barrier). Ascending.
__device__ __forceinline__ void bitonic128_wave(uint64_t &x0, uint64_t &x1, int l) {
#pragma unroll
  for (int k = 2; k <= 128; k <<= 1) {
    const bool up = ((2 * l) & k) == 0;
#pragma unroll
    for (int j = k >> 1; j >= 1; j >>= 1) {
      if (j == 1) {
        const uint64_t lo = x0 < x1 ? x0 : x1, hi = x0 < x1 ? x1 : x0;
        x0 = up ? lo : hi;
        x1 = up ? hi : lo;
        continue;
      }
      const int pj = j >> 1;
      const uint64_t y0 = __shfl_xor(x0, pj, 64), y1 = __shfl_xor(x1, pj, 64);
      const bool keep_min = ((l & pj) == 0) == up;
      x0 = keep_min ? (x0 < y0 ? x0 : y0) : (x0 < y0 ? y0 : x0);
      x1 = keep_min ? (x1 < y1 ? x1 : y1) : (x1 < y1 ? y1 : x1);
    }
  }
}

// Merge passes over sorted runs of L0 keys (N / L0 a power of two, N >= 2 * nthr, nthr a power
// of two): each pass merges run pairs into runs of twice the length, each of the first nthr
// threads producing N / nthr consecutive outputs from its merge-path split (a binary search on
// its diagonal); every thread of the block takes part in the barriers. Keys are distinct except
// the ~0 padding. With `global` one of the two buffers is in global memory and each pass ends
// with a device fence before the barrier. Returns the buffer holding the sorted keys.
__device__ uint64_t *merge_runs(uint64_t *src, uint64_t *dst, int N, int L0, int nthr,
                                bool global = false) {
  const int E = N / nthr;
  for (int L = L0; L < N; L <<= 1) {
    if ((int)threadIdx.x < nthr) {
    const int d = threadIdx.x * E;
    const int pair = d / (2 * L), d0 = d - pair * 2 * L;
    const uint64_t *A = src + (size_t)pair * 2 * L, *Bv = A + L;
    int lo = max(0, d0 - L), hi = min(d0, L);
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (A[mid] < Bv[d0 - 1 - mid])
        lo = mid + 1;
      else
        hi = mid;
    }
    int ia = lo, ib = d0 - lo;
    uint64_t *out = dst + (size_t)pair * 2 * L + d0;
    for (int e = 0; e < E; ++e) {
      const uint64_t av = ia < L ? A[ia] : ~0ull, bv = ib < L ? Bv[ib] : ~0ull;
      const bool take_a = (ib >= L) || (ia < L && av < bv);
      out[e] = take_a ? av : bv;
      ia += take_a ? 1 : 0;
      ib += take_a ? 0 : 1;
    }
    }
    if (global) __threadfence();
    __syncthreads();
    uint64_t *tmp = src;
    src = dst;
    dst = tmp;
  }
  return src;
}

// Writes event q of a sample's time-ordered list (P events, then padding up to a multiple of
// kEventPad): the Event record, its 32-byte fast form and its input index. t is the event's
// time, tn the next event's (or +inf), idx its input neuron; es32 / e32 receive the float32
// exps (as stored for the backward) of a real event.
__device__ __forceinline__ void write_event(int q, int P, float t, float tn, int idx, double dtau_s,
                                            double dtau, double dc, float tmax, Event *ev,
                                            FastEvent *fv, int *ix, float &es32, float &e32) {
  Event e;
  if (q < P) {
    const float tl = fminf(tn, tmax);
    const double es64 = exp((double)t / dtau_s);
    const double zq = exp((double)t / dtau);
    const double zl = fmin(exp((double)tl / dtau), 1e300);
    es32 = (float)es64;
    e32 = (float)zq;
    e.es = (double)es32;
    e.e = (double)e32;
    e.czl2 = dc * zl * zl;
    e.zl1 = zl * (1.0 + kFilterEps);
    e.blo = 2.0 * dc * zq * (1.0 - kFilterEps);
    e.bhi = 2.0 * dc * zl * (1.0 + kFilterEps);
    e.zq = zq;
    e.zl = zl;
    e.idx = idx;
    e.t = t;
    e.tl = tl;
  } else {
    // padding: no contribution, never a candidate
    e.es = 0.0;
    e.e = 0.0;
    e.czl2 = kInfD;
    e.zl1 = 0.0;
    e.blo = kInfD;
    e.bhi = -kInfD;
    e.zq = kInfD;
    e.zl = kInfD;
    e.idx = 0;
    e.t = kInfF;
    e.tl = kInfF;
  }
  e.pad = 0.f;
  ev[q] = e;
  FastEvent f;
  f.es = e.es;
  f.e = e.e;
  f.zq_lo = (q < P) ? e.zq * (1.0 - kFilterEps) : kInfD;  // padding: never a candidate
  f.zl_hi = (q < P) ? e.zl * (1.0 + kFilterEps) : kInfD;
  fv[q] = f;
  ix[q] = e.idx;
}

// The event list of sample b from its P ascending (time, slot) keys, by the whole workgroup:
// the records (padded to a multiple of kEventPad), the input indices (slot / K_pre), the
// per-slot float32 exps at the events' slots and the count.
__device__ __forceinline__ void write_sorted_events(
    const uint64_t *keys, int P, int b, int S, int K_pre, float tau_s, float tau, float c,
    float tmax, int Pcap, Event *__restrict__ events, FastEvent *__restrict__ fast,
    int *__restrict__ evidx, int *__restrict__ counts, float *__restrict__ pre_exp_s,
    float *__restrict__ pre_exp) {
  Event *ev = events + (size_t)b * Pcap;
  FastEvent *fv = fast + (size_t)b * Pcap;
  int *ix = evidx + (size_t)b * Pcap;
  const double dtau_s = (double)tau_s, dtau = (double)tau, dc = (double)c;
  const int Ppad = (P + kEventPad - 1) / kEventPad * kEventPad;
  for (int q = threadIdx.x; q < Ppad; q += blockDim.x) {
    float t = kInfF, tn = kInfF, es32 = 0.f, e32 = 0.f;
    int s = 0;
    if (q < P) {
      const uint64_t key = keys[q];
      t = key_float((uint32_t)(key >> 32));
      s = (int)(uint32_t)key;
      tn = (q + 1 < P) ? key_float((uint32_t)(keys[q + 1] >> 32)) : kInfF;
    }
    write_event(q, P, t, tn, s / K_pre, dtau_s, dtau, dc, tmax, ev, fv, ix, es32, e32);
    if (q < P) {
      pre_exp_s[(size_t)b * S + s] = es32;
      pre_exp[(size_t)b * S + s] = e32;
    }
  }
  if (threadIdx.x == 0) counts[b] = P;
}

// One workgroup per sample: compact the finite slots, sort (time, slot) keys, write the
// sample's Event list padded to a multiple of kEventPad, the event indices, the count, and
// the per-slot float32 exps the backward reads. The waves sort 128-key runs in registers
// (bitonic128_wave) and merge_runs merges them: up to 1024 events one run per wave; longer
// lists that fit the LDS in runs taken wave by wave, with the merge passes' second buffer in
// the LDS when the list fits twice, else in the sample's global key region; lists beyond the
// LDS with the global bitonic network.
__global__ __launch_bounds__(kSortThreads) void sort_events_kernel(
    const float *__restrict__ pre_times, int S, int K_pre, float tau_s, float tau, float c,
    float tmax, int Pcap, Event *__restrict__ events, FastEvent *__restrict__ fast,
    int *__restrict__ evidx,
    int *__restrict__ counts, float *__restrict__ pre_exp_s, float *__restrict__ pre_exp,
    uint64_t *__restrict__ gkeys, int gcap, int lds_cap) {
  extern __shared__ uint64_t lkeys[];
  __shared__ int s_count;
  __shared__ int s_fill;
  const int b = blockIdx.x;
  const float *times = pre_times + (size_t)b * S;
  if (threadIdx.x == 0) {
    s_count = 0;
    s_fill = 0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  // compaction into LDS in one pass: kUnroll independent loads per thread in flight, and a
  // wave's finite slots take consecutive positions from one atomic. Slots past the LDS
  // capacity are only counted; such a sample is compacted again into global memory.
  constexpr int kUnroll = 8;
  const int nth = blockDim.x;
  auto take = [&](float tv, int s) {
    const bool fin = isfinite(tv);
    const unsigned long long m = __ballot(fin);
    int base = 0;
    if (lane == 0 && m) base = atomicAdd(&s_fill, __popcll(m));
    base = __shfl(base, 0, 64);
    const int pos = base + (int)__popcll(m & ((1ull << lane) - 1ull));
    if (fin && pos < lds_cap) lkeys[pos] = ((uint64_t)float_key(tv) << 32) | (uint32_t)s;
  };
  if ((S & 3) == 0 && (reinterpret_cast<size_t>(times) & 15) == 0) {
    // 16-byte loads: four slots per load, kUnroll loads in flight per thread
    const float4 *t4 = reinterpret_cast<const float4 *>(times);
    const int S4 = S >> 2;
    for (int s0 = 0; s0 < S4; s0 += kUnroll * nth) {
      float4 tv[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int q = s0 + u * nth + threadIdx.x;
        tv[u] = q < S4 ? t4[q] : make_float4(kInfF, kInfF, kInfF, kInfF);
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        // the four components' ballots share one LDS atomic per wave
        const int s = 4 * (s0 + u * nth + threadIdx.x);
        const float4 v = tv[u];
        const bool f0 = isfinite(v.x), f1 = isfinite(v.y), f2 = isfinite(v.z),
                   f3 = isfinite(v.w);
        const unsigned long long m0 = __ballot(f0), m1 = __ballot(f1), m2 = __ballot(f2),
                                 m3 = __ballot(f3);
        const int c0 = __popcll(m0), c1 = __popcll(m1), c2 = __popcll(m2), c3 = __popcll(m3);
        const int tot = c0 + c1 + c2 + c3;
        if (tot == 0) continue;
        int base = 0;
        if (lane == 0) base = atomicAdd(&s_fill, tot);
        base = __shfl(base, 0, 64);
        const unsigned long long lt = (1ull << lane) - 1ull;
        int pos = base + (int)__popcll(m0 & lt);
        if (f0 && pos < lds_cap) lkeys[pos] = ((uint64_t)float_key(v.x) << 32) | (uint32_t)s;
        pos = base + c0 + (int)__popcll(m1 & lt);
        if (f1 && pos < lds_cap) lkeys[pos] = ((uint64_t)float_key(v.y) << 32) | (uint32_t)(s + 1);
        pos = base + c0 + c1 + (int)__popcll(m2 & lt);
        if (f2 && pos < lds_cap) lkeys[pos] = ((uint64_t)float_key(v.z) << 32) | (uint32_t)(s + 2);
        pos = base + c0 + c1 + c2 + (int)__popcll(m3 & lt);
        if (f3 && pos < lds_cap) lkeys[pos] = ((uint64_t)float_key(v.w) << 32) | (uint32_t)(s + 3);
      }
    }
  } else {
    for (int s0 = 0; s0 < S; s0 += kUnroll * nth) {
      float tv[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int s = s0 + u * nth + threadIdx.x;
        tv[u] = s < S ? times[s] : kInfF;
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) take(tv[u], s0 + u * nth + threadIdx.x);
    }
  }
  __syncthreads();
  const int P = s_fill;
  int N = 1;
  while (N < P) N <<= 1;
  const bool regs = P <= 1024;
  const bool use_global = !regs && N > lds_cap;
  uint64_t *keys = use_global ? gkeys + (size_t)b * gcap : lkeys;
  if (use_global) {
    __syncthreads();
    if (threadIdx.x == 0) s_count = 0;
    __syncthreads();
    for (int s0 = 0; s0 < S; s0 += nth) {
      const int s = s0 + threadIdx.x;
      const float t = s < S ? times[s] : kInfF;
      const bool fin = isfinite(t);
      const unsigned long long m = __ballot(fin);
      int base = 0;
      if (lane == 0 && m) base = atomicAdd(&s_count, __popcll(m));
      base = __shfl(base, 0, 64);
      if (fin) keys[base + (int)__popcll(m & ((1ull << lane) - 1ull))] =
          ((uint64_t)float_key(t) << 32) | (uint32_t)s;
    }
    __syncthreads();
  }
  if (regs) {
    // 128-key runs sorted by each wave in registers, then merge-path passes through LDS
    // (runs of 128 -> Ns, Ns = max(512, N) keys padded with ~0)
    const int Ns = N > 512 ? N : 512;
    const int l = threadIdx.x & 63, seg = (threadIdx.x >> 6) * 128;
    if (P > 1) {
      uint64_t x0 = 0, x1 = 0;
      if (seg < Ns) {
        x0 = seg + 2 * l < P ? keys[seg + 2 * l] : ~0ull;
        x1 = seg + 2 * l + 1 < P ? keys[seg + 2 * l + 1] : ~0ull;
        bitonic128_wave(x0, x1, l);
      }
      __syncthreads();
      if (seg < Ns) {
        keys[seg + 2 * l] = x0;
        keys[seg + 2 * l + 1] = x1;
      }
      __syncthreads();
      keys = merge_runs(keys, keys + 1024, Ns, 128, (int)blockDim.x);
    }
  } else if (!use_global) {
    // 128-key runs sorted in registers by the waves independently (wave w takes runs w,
    // w + waves, ...), then merge-path passes between two LDS buffers, or between the LDS and
    // the sample's global key region when the LDS holds the keys once only (the host then
    // sizes that region for the list)
    const bool alt_global = 2 * N > lds_cap;
    uint64_t *alt = alt_global ? gkeys + (size_t)b * gcap : keys + N;
    for (int q = P + threadIdx.x; q < N; q += blockDim.x) keys[q] = ~0ull;
    __syncthreads();
    const int l = threadIdx.x & 63, nwv = blockDim.x >> 6;
    for (int r0 = (threadIdx.x >> 6) * 128; r0 < N; r0 += nwv * 128) {
      uint64_t x0 = keys[r0 + 2 * l], x1 = keys[r0 + 2 * l + 1];
      bitonic128_wave(x0, x1, l);
      keys[r0 + 2 * l] = x0;  // the wave's own run: its reads above are done
      keys[r0 + 2 * l + 1] = x1;
    }
    __syncthreads();
    keys = merge_runs(keys, alt, N, 128, (int)blockDim.x, alt_global);
  } else {
    for (int q = P + threadIdx.x; q < N; q += blockDim.x) keys[q] = ~0ull;
    if (use_global) __threadfence();
    __syncthreads();
    bitonic_sort(keys, N, use_global);
  }

  write_sorted_events(keys, P, b, S, K_pre, tau_s, tau, c, tmax, Pcap, events, fast, evidx, counts,
                      pre_exp_s, pre_exp);
}

// One workgroup per sample of a packed list: sample b's events are
// indices/times[row_offsets[b] .. row_offsets[b+1]) in time order, and an index outside
// [0, I) or a time that is not finite ends the list. They become the sample's Event list directly (no per-neuron buffer,
// no sort): events of equal time keep their list order.
__global__ __launch_bounds__(256) void packed_events_kernel(
    const int *__restrict__ row_offsets, const int *__restrict__ indices,
    const float *__restrict__ times, int I, float tau_s, float tau, float c, float tmax,
    int P_max, int Pcap, Event *__restrict__ events, FastEvent *__restrict__ fast,
    int *__restrict__ evidx, int *__restrict__ counts) {
  __shared__ int s_end;
  const int b = blockIdx.x;
  const int o0 = row_offsets[b];
  const int n = min(max(row_offsets[b + 1] - o0, 0), P_max);  // events past P_max are dropped
  if (threadIdx.x == 0) s_end = n;
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const int i = indices[o0 + q];
    if (i < 0 || i >= I || !isfinite(times[o0 + q])) atomicMin(&s_end, q);
  }
  __syncthreads();
  const int P = s_end;
  const double dtau_s = (double)tau_s, dtau = (double)tau, dc = (double)c;
  const int Ppad = (P + kEventPad - 1) / kEventPad * kEventPad;
  Event *ev = events + (size_t)b * Pcap;
  FastEvent *fv = fast + (size_t)b * Pcap;
  int *ix = evidx + (size_t)b * Pcap;
  for (int q = threadIdx.x; q < Ppad; q += blockDim.x) {
    float t = kInfF, tn = kInfF, es32, e32;
    int i = 0;
    if (q < P) {
      t = times[o0 + q];
      i = indices[o0 + q];
      tn = q + 1 < P ? times[o0 + q + 1] : kInfF;
    }
    write_event(q, P, t, tn, i, dtau_s, dtau, dc, tmax, ev, fv, ix, es32, e32);
  }
  if (threadIdx.x == 0) counts[b] = P;
}

// ------------------------------------------------------------------ solver, lane per neuron
// A workgroup is one sample and up to 1024 consecutive post neurons (one lane each; all of them
// when J <= 1024, so that every wave of the sample shares one CU's caches). The 32-byte
// FastEvent records are wave-uniform and read with scalar loads one event ahead; the weights
// of the next kGroup events are requested while the current group is processed, through a
// buffer resource over W^T (the lane's column is a VGPR offset, the event's row offset a
// scalar one: no 64-bit address arithmetic per load), with the group's input indices loaded a
// group earlier still. Per event: a and b accumulate in float64 and the clamp filter runs;
// the branch into the closed-form loop is wave-uniform (a ballot of the lanes whose filter
// passes), so the common event costs 11 VALU instructions, two waits, one scalar load and an
// untaken branch. The spike times of a lane are computed after its walk.
constexpr int kSolveMaxThreads = 1024;
constexpr int kGroup = kEventPad;
__device__ __forceinline__ float load_weight(__amdgpu_buffer_rsrc_t r, int lane_off,
                                             int row_off) {
  return buffer_load_f32(r, lane_off, row_off);
}

// The next dense layer's event lists, built by the lane solver from its own spikes when the
// whole of a sample's layer is one workgroup: the regions build_event_lists lays out for that
// layer's S = J * K slots (records, indices, counts, global sort keys of gcap per sample), its
// per-slot exps, and its constants. lds_keys (a power of two) is the LDS room for a sample's
// sort keys; 0 turns the chain off. A sample with more keys sorts them in its global region
// (the host guarantees gcap >= the next power of two of S then).
struct ChainOut {
  Event *events;
  FastEvent *fast;
  int *idx;
  int *counts;
  uint64_t *gkeys;
  float *pre_es, *pre_e;
  int Pcap, gcap, lds_keys;
  float tau_s, tau, c, tmax;
};

// kAlwaysExact: c <= 0 (no threshold quadratic to filter with), every interval is solved.
// kStaged: the event records come from a wave-private LDS stage filled one 64-event chunk
// ahead (for launches with few waves per SIMD, whose per-event scalar loads would otherwise
// be exposed); otherwise each record is a scalar load issued one event ahead.
// W^T rows have Jp entries; wt_bytes (< 2^31) bounds the buffer resource.
template <bool kAlwaysExact, bool kStaged>
__global__ __launch_bounds__(kSolveMaxThreads) void lif_solve_seq_kernel(
    const Event *__restrict__ events, const FastEvent *__restrict__ fast,
    const int *__restrict__ evidx,
    const int *__restrict__ counts, int Pcap, const float *__restrict__ WT, int J, int Jp,
    int wt_bytes, int K, double c, double dtt, double tau, int *__restrict__ out_n,
    float *__restrict__ out_t, float *__restrict__ out_a, float *__restrict__ out_x,
    float *__restrict__ out_z, const ChainOut chain) {
  // dynamic LDS (seq_solver_lds_bytes): the spike stash [kLaneStash][nt] float64, the row
  // heads (first spike times of each row) [nt][kLaneStash] float32, the counts [nt], in the
  // staged form the wave-private record stages [nt / 64][64] FastEvent, and with the chain
  // the sort keys [chain.lds_keys] and the wave totals [nt / 64]
  extern __shared__ __align__(16) unsigned char s_dyn[];
  const int nt = blockDim.x;
  double *s_z = reinterpret_cast<double *>(s_dyn);
  float *s_head = reinterpret_cast<float *>(s_z + kLaneStash * nt);
  int *s_n = reinterpret_cast<int *>(s_head + kLaneStash * nt);
  const int b = blockIdx.y;
  const int j0 = blockIdx.x * blockDim.x;
  const int j = j0 + threadIdx.x;
  const bool active = j < J;
  const int jc = active ? j : 0;
  const int P = counts[b];
  const int Ppad = (P + kGroup - 1) / kGroup * kGroup;
  const Event *ev = events + (size_t)b * Pcap;
  const FastEvent *fv = fast + (size_t)b * Pcap;
  const int *ix = evidx + (size_t)b * Pcap;
  const int lane_off = jc * (int)sizeof(float);
  const int row_bytes = Jp * (int)sizeof(float);
  const __amdgpu_buffer_rsrc_t wres =
      __builtin_amdgcn_make_buffer_rsrc((void *)WT, (short)0, wt_bytes, kBufferWord3);
  const double inv2c = 0.5 / c;
  const size_t row = ((size_t)b * J + jc) * K;
  const LaneStash stash{s_z, (int)threadIdx.x, nt, out_t, out_z};
  // With the filter, a lane that is inactive or whose buffer is full carries a = -inf (it
  // stays -inf and fails the filter's a > 0); in the always-exact form a flag does this.
  double a = (kAlwaysExact || active) ? 0.0 : -kInfD, bb = 0.0;
  int n = 0;
  bool live = active;
  float wn[kGroup];  // weights of the next group
  int ixn[kGroup];   // input indices of the group after it
  if (Ppad > 0) {
#pragma unroll
    for (int u = 0; u < kGroup; ++u) wn[u] = load_weight(wres, lane_off, ix[u] * row_bytes);
  }
#pragma unroll
  for (int u = 0; u < kGroup; ++u) ixn[u] = kGroup < Ppad ? ix[kGroup + u] : 0;
  // staged form: each wave keeps the records of the next 64 events in its own LDS slice,
  // filled with one coalesced load per chunk issued a whole chunk ahead
  FastEvent *stage = reinterpret_cast<FastEvent *>(s_n + nt) + (kStaged ? (threadIdx.x & ~63) : 0);
  const int lane = threadIdx.x & 63;
  FastEvent reg{};
  if (kStaged && lane < Ppad) reg = fv[lane];
  // the record of the event being processed; the next one is always requested during the
  // current event's arithmetic (across group boundaries too)
  FastEvent cur{};
  if (!kStaged && Ppad > 0) cur = fv[0];
  for (int c0 = 0; c0 < Ppad; c0 += 64) {
    const int cend = min(c0 + 64, Ppad);
    if (kStaged) {
      stage[lane] = reg;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (c0 + 64 + lane < Ppad) reg = fv[c0 + 64 + lane];
      cur = stage[0];
    }
    const int qend = kStaged ? cend : Ppad;
    for (int q0 = c0; q0 < cend; q0 += kGroup) {
      float w[kGroup];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) w[u] = wn[u];
      if (q0 + kGroup < Ppad) {
#pragma unroll
        for (int u = 0; u < kGroup; ++u) wn[u] = load_weight(wres, lane_off, ixn[u] * row_bytes);
        if (q0 + 2 * kGroup < Ppad) {
#pragma unroll
          for (int u = 0; u < kGroup; ++u) ixn[u] = ix[q0 + 2 * kGroup + u];
        }
      }
      const bool more = q0 + kGroup < qend;
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        FastEvent nxt;
        if (u + 1 < kGroup || more) nxt = kStaged ? stage[q0 - c0 + u + 1] : fv[q0 + u + 1];
        const double w64 = (double)w[u];
        a = fma(w64, cur.es, a);
        bb = fma(w64, cur.e, bb);
        const bool pass = kAlwaysExact ? live : clamp_filter(a, bb, cur, c, inv2c);
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(pass) != 0, 0)) {
          if (pass) {
            const SolveStateZ r = solve_crossings_deferred_stash(
                a, bb, n, cur.zq_lo, cur.zl_hi, EventEnds{ev + q0 + u}, c, inv2c, dtt, tau, K,
                row, out_a, out_x, stash);
            bb = r.b;
            n = r.n;
            if (kAlwaysExact)
              live = r.live != 0;
            else if (!r.live)
              a = -kInfD;
          }
        }
        if (u + 1 < kGroup || more) cur = nxt;
      }
    }
    if (kStaged) {
      // this chunk's stage reads come before the next chunk's write
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
  // spike times: the first kLaneStash of a row from the LDS stash into the row heads (their t
  // slots are written below with the row), later ones in place from their parked words
  if (active) {
#pragma unroll
    for (int k = 0; k < kLaneStash; ++k) {
      if (k < n) {
        const double z = s_z[k * nt + threadIdx.x];
        s_head[threadIdx.x * kLaneStash + k] = (float)(tau * log(z));
        out_z[row + k] = (float)z;
      }
    }
    if (n > kLaneStash) finish_spike_times(row, n, tau, out_t, out_z, kLaneStash);
    out_n[(size_t)b * J + j] = n;
  }
  s_n[threadIdx.x] = active ? n : 0;
  __syncthreads();
  // the block's rows of t, written row by row with consecutive lanes on consecutive slots
  // (whole lines): the row heads, then +inf; slots of later spikes are already written.
  // Streamed past the caches (no later kernel reads them before they have left L2, and they
  // would evict the other workgroups' weights).
  const int jcount = min(nt, J - j0);
  float *tb = out_t + ((size_t)b * J + j0) * K;
  const auto put = [&](int r, int k) {
    const int nr = s_n[r];
    if (k < nr && k < kLaneStash)
      __builtin_nontemporal_store(s_head[r * kLaneStash + k], tb + (size_t)r * K + k);
    else if (k >= nr)
      __builtin_nontemporal_store(kInfF, tb + (size_t)r * K + k);
  };
  if (K <= 32) {
    // two rows per wave and pass (lanes 0-31 one row, lanes 32-63 the next), four passes'
    // LDS reads in one round
    const int lane = threadIdx.x & 63, k = lane & 31;
    const int step = 2 * (nt >> 6);
    for (int r0 = 2 * (threadIdx.x >> 6) + (lane >> 5); r0 < jcount; r0 += 4 * step) {
      int nr[4];
      float hv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = min(r0 + u * step, jcount - 1);
        nr[u] = s_n[r];
        hv[u] = s_head[r * kLaneStash + min(k, kLaneStash - 1)];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = r0 + u * step;
        if (r < jcount && k < K) {
          if (k < nr[u] && k < kLaneStash)
            __builtin_nontemporal_store(hv[u], tb + (size_t)r * K + k);
          else if (k >= nr[u])
            __builtin_nontemporal_store(kInfF, tb + (size_t)r * K + k);
        }
      }
    }
  } else {
    for (int e = threadIdx.x; e < jcount * K; e += nt) {
      const int r = e / K;
      put(r, e - r * K);
    }
  }
  if (chain.lds_keys > 0) {
    // the next layer's event list of this sample (the workgroup holds all its neurons: j0 = 0):
    // (time, slot) keys of every spike at offsets from a block scan of the counts, sorted,
    // then the records, as sort_events_kernel writes them from the same spike trains
    uint64_t *s_keys = reinterpret_cast<uint64_t *>(
        reinterpret_cast<char *>(s_n + nt) + (kStaged ? (size_t)nt * sizeof(FastEvent) : 0));
    int *s_wtot = reinterpret_cast<int *>(s_keys + chain.lds_keys);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = nt >> 6;
    const int cnt = s_n[threadIdx.x];
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(incl, d, 64);
      if (lane >= d) incl += y;
    }
    if (lane == 63) s_wtot[wv] = incl;
    __syncthreads();
    int base = 0, P2 = 0;
    for (int w = 0; w < nw; ++w) {
      const int v = s_wtot[w];
      base += w < wv ? v : 0;
      P2 += v;
    }
    const int excl = base + incl - cnt;
    int N2 = 1;
    while (N2 < P2) N2 <<= 1;
    const bool glob = N2 > chain.lds_keys;
    uint64_t *keys = glob ? chain.gkeys + (size_t)b * chain.gcap : s_keys;
    for (int k = 0; k < cnt; ++k) {
      // a row's later spike times were written in place by this lane
      const float tk = k < kLaneStash ? s_head[threadIdx.x * kLaneStash + k] : out_t[row + k];
      keys[excl + k] = ((uint64_t)float_key(tk) << 32) | (uint32_t)(j * K + k);
    }
    // runs of 128 keys sorted by the waves in registers, then merge-path passes between two
    // LDS halves, when both fit; otherwise a bitonic network in place (in LDS, or in the
    // sample's global key region)
    const int Ns = N2 > 128 ? N2 : 128;
    const bool runs = !glob && 2 * Ns <= chain.lds_keys;
    const int Nfill = runs ? Ns : N2;
    for (int q = P2 + threadIdx.x; q < Nfill; q += nt) keys[q] = ~0ull;
    if (glob) __threadfence();
    __syncthreads();
    if (runs) {
      for (int r0 = wv * 128; r0 < Ns; r0 += nw * 128) {
        uint64_t x0 = keys[r0 + 2 * lane], x1 = keys[r0 + 2 * lane + 1];
        bitonic128_wave(x0, x1, lane);
        keys[r0 + 2 * lane] = x0;
        keys[r0 + 2 * lane + 1] = x1;
      }
      __syncthreads();
      if (Ns > 128) {
        int nthr = 1;
        while (2 * nthr <= nt && 2 * nthr <= Ns / 2) nthr <<= 1;
        keys = merge_runs(keys, keys + Ns, Ns, 128, nthr);
      }
    } else if (P2 > 1) {
      bitonic_sort(keys, N2, glob);
    }
    write_sorted_events(keys, P2, b, J * K, K, chain.tau_s, chain.tau, chain.c, chain.tmax,
                        chain.Pcap, chain.events, chain.fast, chain.idx, chain.counts,
                        chain.pre_es, chain.pre_e);
  }
}

// Dynamic LDS of lif_solve_seq_kernel for nt lanes per workgroup (and lds_keys sort keys of
// the chained event lists).
static size_t seq_solver_lds_bytes(int nt, bool staged, int lds_keys = 0) {
  return (size_t)nt * (kLaneStash * (sizeof(double) + sizeof(float)) + sizeof(int)) +
         (staged ? (size_t)nt * sizeof(FastEvent) : 0) +
         (lds_keys > 0 ? (size_t)lds_keys * sizeof(uint64_t) + sizeof(int) * (nt / 64) : 0);
}

// ------------------------------------------------------------------ solver, first spike only
// Layers with one spike per neuron (K = 1) and few (sample, neuron) pairs per SIMD: before a
// neuron's first spike no reset has happened, so its a and b at any event are the plain prefix
// sums of the inputs, and its walk splits exactly into time segments. A workgroup is 64 post
// neurons of one sample times nseg waves, wave s walking segment s of the event list:
//   1. the segment's sums of w*es and w*e (no filter);
//   2. the segment's start state, the sums of the earlier segments (in segment order);
//   3. the lane solver's walk over the segment from that state (filter, closed form), which
//      stops at the first accepted spike; z, a and x wait in LDS;
//   4. the neuron's spike is that of its earliest segment with one (the later segments' walks
//      assumed no earlier spike and are discarded).
// The running sums are associated by segment instead of event by event (float64 rounding of
// the last bits, as in the wave-per-neuron solver). kSegMax segments at most.
constexpr int kSegMax = 16;

__global__ __launch_bounds__(64 * kSegMax) void lif_solve_seg_kernel(
    const Event *__restrict__ events, const FastEvent *__restrict__ fast,
    const int *__restrict__ evidx, const int *__restrict__ counts, int Pcap,
    const float *__restrict__ WT, int J, int Jp, int wt_bytes, double c, double dtt,
    double tau, int nseg, int *__restrict__ out_n, float *__restrict__ out_t,
    float *__restrict__ out_a, float *__restrict__ out_x, float *__restrict__ out_z) {
  __shared__ double s_sa[kSegMax][64], s_sb[kSegMax][64], s_z[kSegMax][64];
  __shared__ float s_a[kSegMax][64], s_x[kSegMax][64];
  __shared__ int s_hit[kSegMax][64];
  const int b = blockIdx.y;
  const int sg = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 64 + lane;
  const bool active = j < J;
  const int jc = active ? j : 0;
  const int P = counts[b];
  const int Ppad = (P + kGroup - 1) / kGroup * kGroup;
  const int L = (Ppad / kGroup + nseg - 1) / nseg * kGroup;  // events per segment
  const int q_lo = min(sg * L, Ppad), q_hi = min(q_lo + L, Ppad);
  const Event *ev = events + (size_t)b * Pcap;
  const FastEvent *fv = fast + (size_t)b * Pcap;
  const int *ix = evidx + (size_t)b * Pcap;
  const int lane_off = jc * (int)sizeof(float);
  const int row_bytes = Jp * (int)sizeof(float);
  const __amdgpu_buffer_rsrc_t wres =
      __builtin_amdgcn_make_buffer_rsrc((void *)WT, (short)0, wt_bytes, kBufferWord3);
  // 1. the segment's plain sums (the next group's weights in flight, its indices a group
  // earlier still)
  double sa = 0.0, sb = 0.0;
  {
    float pw[kGroup];
    int pix[kGroup];
    if (q_lo < q_hi) {
#pragma unroll
      for (int u = 0; u < kGroup; ++u) pw[u] = load_weight(wres, lane_off, ix[q_lo + u] * row_bytes);
    }
#pragma unroll
    for (int u = 0; u < kGroup; ++u) pix[u] = q_lo + kGroup < q_hi ? ix[q_lo + kGroup + u] : 0;
    for (int q0 = q_lo; q0 < q_hi; q0 += kGroup) {
      float w[kGroup];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) w[u] = pw[u];
      if (q0 + kGroup < q_hi) {
#pragma unroll
        for (int u = 0; u < kGroup; ++u) pw[u] = load_weight(wres, lane_off, pix[u] * row_bytes);
        if (q0 + 2 * kGroup < q_hi) {
#pragma unroll
          for (int u = 0; u < kGroup; ++u) pix[u] = ix[q0 + 2 * kGroup + u];
        }
      }
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        const FastEvent f = fv[q0 + u];
        sa = fma((double)w[u], f.es, sa);
        sb = fma((double)w[u], f.e, sb);
      }
    }
  }
  s_sa[sg][lane] = sa;
  s_sb[sg][lane] = sb;
  __syncthreads();
  // 2. the state before the segment's first event
  double a = 0.0, bb = 0.0;
  for (int s2 = 0; s2 < sg; ++s2) {
    a += s_sa[s2][lane];
    bb += s_sb[s2][lane];
  }
  if (!active) a = -kInfD;  // fails the filter's a > 0 for good
  // 3. the walk over the segment (the next group's weights and indices in flight)
  const double inv2c = 0.5 / c;
  const SegStash stash{&s_z[sg][lane], &s_a[sg][lane], &s_x[sg][lane]};
  int n = 0;
  float wn[kGroup];
  int ixn[kGroup];
  if (q_lo < q_hi) {
#pragma unroll
    for (int u = 0; u < kGroup; ++u) wn[u] = load_weight(wres, lane_off, ix[q_lo + u] * row_bytes);
  }
#pragma unroll
  for (int u = 0; u < kGroup; ++u) ixn[u] = q_lo + kGroup < q_hi ? ix[q_lo + kGroup + u] : 0;
  FastEvent cur{};
  if (q_lo < q_hi) cur = fv[q_lo];
  for (int q0 = q_lo; q0 < q_hi; q0 += kGroup) {
    float w[kGroup];
#pragma unroll
    for (int u = 0; u < kGroup; ++u) w[u] = wn[u];
    if (q0 + kGroup < q_hi) {
#pragma unroll
      for (int u = 0; u < kGroup; ++u) wn[u] = load_weight(wres, lane_off, ixn[u] * row_bytes);
      if (q0 + 2 * kGroup < q_hi) {
#pragma unroll
        for (int u = 0; u < kGroup; ++u) ixn[u] = ix[q0 + 2 * kGroup + u];
      }
    }
    const bool more = q0 + kGroup < q_hi;
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      FastEvent nxt;
      if (u + 1 < kGroup || more) nxt = fv[q0 + u + 1];
      const double w64 = (double)w[u];
      a = fma(w64, cur.es, a);
      bb = fma(w64, cur.e, bb);
      const bool pass = clamp_filter(a, bb, cur, c, inv2c);
      if (__builtin_expect(__builtin_amdgcn_ballot_w64(pass) != 0, 0)) {
        if (pass) {
          // the potential already at the threshold where the interval starts (g <= 0 a little
          // before it): the smaller root lies before the interval, which the closed form
          // would reject, so it is not formed. This is what keeps a segment that starts after
          // the neuron's first spike (its state has no reset) from taking the closed form at
          // every event while the potential stays above the threshold.
          const double gq = fma(fma(c, cur.zq_lo, -bb), cur.zq_lo, a);
          if (!(gq <= 0.0)) {
            const SolveStateZ r = solve_crossings_deferred_stash(
                a, bb, n, cur.zq_lo, cur.zl_hi, EventEnds{ev + q0 + u}, c, inv2c, dtt, tau, 1,
                0, nullptr, nullptr, stash);
            bb = r.b;
            n = r.n;
            if (!r.live) a = -kInfD;
          }
        }
      }
      if (u + 1 < kGroup || more) cur = nxt;
    }
  }
  s_hit[sg][lane] = n;
  __syncthreads();
  // 4. the earliest segment's spike
  if (sg == 0 && active) {
    int win = -1;
    for (int s2 = nseg - 1; s2 >= 0; --s2)
      if (s_hit[s2][lane]) win = s2;
    const size_t row = (size_t)b * J + j;
    if (win >= 0) {
      const double z = s_z[win][lane];
      out_t[row] = (float)(tau * log(z));
      out_z[row] = (float)z;
      out_a[row] = s_a[win][lane];
      out_x[row] = s_x[win][lane];
      out_n[row] = 1;
    } else {
      out_t[row] = kInfF;
      out_n[row] = 0;
    }
  }
}

// Sort keys of a chained sample held in LDS at most (beyond: the global key region).
constexpr int kChainLdsKeys = 8192;

// ------------------------------------------------------------------ solver, wave per neuron
// One wave per (sample, post neuron); lane l holds event q0 + l of the current 64-event chunk.
// a_q and b_q come from float64 inclusive wave scans of w*exp terms plus the carry of the
// previous chunks; every lane evaluates the crossing filter for its own event, and the
// candidates are resolved in time order (lowest lane first) with the reference's closed-form
// loop on wave-uniform values. A reset lowers b for all later events of the chunk, after
// which the filter is re-evaluated for them.
constexpr int kWaveThreads = 256;

// A float64 moved between lanes with one DPP control, half by half; lanes without a source
// (and rows outside row_mask) receive 0.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp_move_f64(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)bits, kCtrl, kRowMask, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), kCtrl, kRowMask, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Inclusive float64 prefix sum over the 64 lanes with DPP moves (no LDS): steps 1, 2, 4, 8
// within each row of 16 lanes (row_shr), then the last lane of row 0 (row 2) is added to row 1
// (row 3) (row_bcast:15) and the last lane of row 1 to rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ double wave_scan_incl(double v) {
  v += dpp_move_f64<0x111, 0xf>(v);  // row_shr:1
  v += dpp_move_f64<0x112, 0xf>(v);  // row_shr:2
  v += dpp_move_f64<0x114, 0xf>(v);  // row_shr:4
  v += dpp_move_f64<0x118, 0xf>(v);  // row_shr:8
  v += dpp_move_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

// The value of a wave-uniform lane: two lane reads into SGPRs.
__device__ __forceinline__ double lane_value(double v, int l) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)bits, l);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__global__ __launch_bounds__(kWaveThreads) void lif_solve_wave_kernel(
    const Event *__restrict__ events, const int *__restrict__ evidx,
    const int *__restrict__ counts, int Pcap,
    const float *__restrict__ W, int I, int J, int BJ, int K, double c, double dtt, double tau,
    int always_exact, int *__restrict__ out_n, float *__restrict__ out_t,
    float *__restrict__ out_a, float *__restrict__ out_x, float *__restrict__ out_z) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (kWaveThreads / 64) + (threadIdx.x >> 6);
  if (wave >= BJ) return;  // uniform per wave; the kernel has no block-level barrier
  const int b = wave / J;
  const int j = wave - b * J;
  const int P = counts[b];
  const Event *ev = events + (size_t)b * Pcap;
  const int *ix = evidx + (size_t)b * Pcap;
  const float *wrow = W + (size_t)j * I;
  const size_t row = ((size_t)b * J + j) * K;
  const double c4 = 4.0 * c;
  double carry_a = 0.0, carry_b = 0.0;
  int n = 0;
  bool live = true;
  // loads in flight ahead of the chunk being solved: the event record and the weight of the
  // next chunk, the input index of the one after (its weight gather needs it)
  Event en{};
  float wn = 0.f;
  int inn = 0;
  if (lane < P) {
    en = ev[lane];
    wn = wrow[ix[lane]];
  }
  if (64 + lane < P) inn = ix[64 + lane];
  for (int q0 = 0; q0 < P && live; q0 += 64) {
    const int q = q0 + lane;
    const bool valid = q < P;
    const Event e = en;
    const float wq = wn;
    if (q + 64 < P) {
      en = ev[q + 64];
      wn = wrow[inn];
    }
    if (q + 128 < P) inn = ix[q + 128];
    double pa = 0.0, pb = 0.0;
    if (valid) {
      const double w = (double)wq;
      pa = w * e.es;
      pb = w * e.e;
    }
    double A = wave_scan_incl(pa) + carry_a;
    double Bv = wave_scan_incl(pb) + carry_b;
    bool cand = valid & (always_exact | crossing_filter(A, Bv, e, c4));
    unsigned long long mask = __ballot(cand);
    while (mask) {
      const int f = __builtin_amdgcn_readfirstlane(__ffsll(mask) - 1);
      const double af = lane_value(A, f);
      const double bin = lane_value(Bv, f);
      const double tq = (double)lane_value(e.t, f);
      const double tl = (double)lane_value(e.tl, f);
      double bf = bin;
      const int n0 = n;
      live = solve_crossings(af, bf, n, tq, tl, c, dtt, tau, K, row, out_t, out_a, out_x, out_z,
                             lane == 0);
      if (!live) break;
      if (n != n0) {
        const double delta = bin - bf;
        if (lane > f) Bv -= delta;
        if (lane == f) Bv = bf;
        cand = valid & (lane > f) & (always_exact | crossing_filter(A, Bv, e, c4));
        mask = __ballot(cand);
      } else {
        mask &= ~(1ull << f);
      }
    }
    const int last_lane = __builtin_amdgcn_readfirstlane(min(63, P - 1 - q0));
    carry_a = lane_value(A, last_lane);
    carry_b = lane_value(Bv, last_lane);
  }
  if (lane == 0) out_n[(size_t)b * J + j] = n;
  for (int k = n + lane; k < K; k += 64) out_t[row + k] = kInfF;
}

// ------------------------------------------------------------------ host side
// Solver forms (per call): 0 auto, 1 lane per neuron, 2 wave per neuron, 3 / 4 lane per
// neuron with staged / scalar event records.

// Lane-per-neuron keeps the event walk sequential in each lane and needs many (sample,
// neuron) pairs to fill 256 CUs; below this many pairs the wave-per-neuron form is used.
constexpr long long kSeqMinLanes = 32768;

static int lds_capacity_entries() {
  static int cached[64];
  static std::once_flag once[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 4096;
  std::call_once(once[dev], [dev]() {
    int bytes = 0;
    if (hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
        hipSuccess)
      bytes = 65536;
    bytes -= 1024;  // static LDS of the kernel + slack
    int cap = 1;
    while ((cap << 1) * 8 <= bytes) cap <<= 1;
    cached[dev] = cap;
    if (cap * 8 > 65536)
      (void)hipFuncSetAttribute((const void *)sort_events_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, cap * 8);
  });
  return cached[dev];
}

// Compute units of the current device (256 on MI355X), read once per device.
static int cu_count() {
  static int cached[64];
  static std::once_flag once[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  std::call_once(once[dev], [dev]() {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n <= 0)
      n = 256;
    cached[dev] = n;
  });
  return cached[dev];
}

// The lane solver's dynamic LDS reaches about 151 KB at 1024 lanes in the staged form with
// chained event lists: the limit above the default 64 KB is raised once per device for all
// four instantiations.
static int allow_seq_solver_lds() {
  static int status[64];
  static std::once_flag once[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
  std::call_once(once[dev], [dev]() {
    const int bytes = (int)seq_solver_lds_bytes(kSolveMaxThreads, true, kChainLdsKeys);
    const void *fns[4] = {(const void *)lif_solve_seq_kernel<true, true>,
                          (const void *)lif_solve_seq_kernel<true, false>,
                          (const void *)lif_solve_seq_kernel<false, true>,
                          (const void *)lif_solve_seq_kernel<false, false>};
    int rc = 0;
    for (const void *f : fns)
      if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
        rc = -1;
    status[dev] = rc;
  });
  return status[dev];
}

static int event_capacity(int S) {
  return std::max(kEventPad, (S + kEventPad - 1) / kEventPad * kEventPad);
}

// Global sort keys per sample: needed when a list may not fit the LDS twice (the second
// buffer of the merge passes, or the whole list beyond the LDS).
static int global_key_capacity(int N, int cap) { return (2 * N > cap) ? N : 0; }

size_t event_lists_bytes(int B, int S) {
  const int cap = lds_capacity_entries();
  int N = 1;
  while (N < S) N <<= 1;
  const int gcap = global_key_capacity(N, cap);
  const int Pcap = event_capacity(S);
  size_t off = 0;
  off = align_up(off + sizeof(Event) * (size_t)B * Pcap, 256);
  off = align_up(off + sizeof(FastEvent) * (size_t)B * Pcap, 256);
  off = align_up(off + sizeof(int) * (size_t)B * Pcap, 256);
  off = align_up(off + sizeof(int) * (size_t)B, 256);
  off = align_up(off + sizeof(uint64_t) * (size_t)B * gcap, 256);
  return off;
}

// The regions of a batch's event lists inside `ws` (event_lists_bytes(B, S) bytes): records,
// fast records, indices, counts and, for lists longer than the LDS sort holds, global sort
// keys of gcap per sample.
static EventLists event_lists_at(void *ws, int B, int S, uint64_t **gkeys, int *gcap_out) {
  const int cap = lds_capacity_entries();
  int N = 1;
  while (N < S) N <<= 1;
  const int gcap = global_key_capacity(N, cap);
  const int Pcap = event_capacity(S);
  char *base = static_cast<char *>(ws);
  EventLists el;
  size_t off = 0;
  el.events = reinterpret_cast<Event *>(base + off);
  off = align_up(off + sizeof(Event) * (size_t)B * Pcap, 256);
  el.fast = reinterpret_cast<FastEvent *>(base + off);
  off = align_up(off + sizeof(FastEvent) * (size_t)B * Pcap, 256);
  el.idx = reinterpret_cast<int *>(base + off);
  off = align_up(off + sizeof(int) * (size_t)B * Pcap, 256);
  el.counts = reinterpret_cast<int *>(base + off);
  off = align_up(off + sizeof(int) * (size_t)B, 256);
  el.Pcap = Pcap;
  if (gkeys) *gkeys = reinterpret_cast<uint64_t *>(base + off);
  if (gcap_out) *gcap_out = gcap;
  return el;
}

int build_event_lists(hipStream_t st, int B, int S, int K_pre, const float *pre_times,
                      float tau_s, float tau, float c, float tmax, float *pre_es, float *pre_e,
                      void *ws, EventLists *out, bool timed) {
  const int cap = lds_capacity_entries();
  int N = 1;
  while (N < S) N <<= 1;
  uint64_t *gkeys = nullptr;
  int gcap = 0;
  *out = event_lists_at(ws, B, S, &gkeys, &gcap);
  const int Pcap = out->Pcap;
  if (S > 0) {
    // room for the list twice (the merge passes) up to the LDS limit; 2048 at least for the
    // short-list merge buffers
    const int lds_entries = std::max(2048, std::min(2 * N, cap));
    BATS_TIMED_IF(timed, "sort_events_kernel", st,
               hipLaunchKernelGGL(sort_events_kernel, dim3(B), dim3(kSortThreads),
                                  sizeof(uint64_t) * lds_entries, st, pre_times, S, K_pre, tau_s,
                                  tau, c, tmax, Pcap, out->events, out->fast, out->idx, out->counts, pre_es,
                                  pre_e, gkeys, gcap, lds_entries));
    return check_launch("sort_events_kernel");
  }
  hipLaunchKernelGGL(fill_kernel<int>, dim3((B + 255) / 256), dim3(256), 0, st, out->counts,
                     (int64_t)B, 0);
  return check_launch("zero_counts");
}

// W^T row length: J rounded up to a multiple of 4 (rows start 16-byte aligned)
static int wt_stride(int J) { return (J + 3) & ~3; }

static size_t fwd_bytes(int B, int I, int K_pre, int J) {
  return align_up(sizeof(float) * (size_t)I * wt_stride(J), 256) +
         event_lists_bytes(B, I * K_pre);
}

}  // namespace bats

using namespace bats;

extern "C" size_t bats_lif_forward_workspace_size(int B, int I, int K_pre, int J) {
  if (B < 0 || I < 0 || K_pre < 0 || J < 0) return 0;
  return fwd_bytes(B, I, K_pre, J);
}

// The forward of one dense layer. solver_mode / solver_blocks choose the solver form and the
// lane solver's workgroups per sample (0: automatic); `timed` allows library timing events.
// The solver launch over a batch's event lists (el): one wave per (sample, post neuron) in
// wave_mode (W read row by row), else one lane per pair over W^T (WT, rows of Jp).
// With `chain`, the lane solver also builds the next layer's event lists when one workgroup
// holds a whole sample (chain->lds_keys is then set; otherwise it stays 0 and the caller
// builds them from the spike trains).
static int solve_event_lists(hipStream_t st, int B, int I, int J, int K_post,
                             const EventLists &el, bool any_events, const float *weights,
                             const float *WT, int Jp, bool wave_mode, int solver_mode,
                             int solver_blocks, float c, float delta_theta_tau, float tau,
                             int *out_n, float *out_t, float *out_a, float *out_x, float *out_z,
                             bool timed, ChainOut *chain = nullptr) {
  const int always_exact = (c > 0.f) ? 0 : 1;
  const long long BJ = (long long)B * J;
  if (wave_mode && any_events) {
    BATS_REQUIRE(BJ < 2147483647LL, "too many (sample, neuron) pairs");
    const unsigned blocks = (unsigned)((BJ + (kWaveThreads / 64) - 1) / (kWaveThreads / 64));
    BATS_TIMED_IF(timed, "lif_solve_wave_kernel", st,
               hipLaunchKernelGGL(lif_solve_wave_kernel, dim3(blocks), dim3(kWaveThreads), 0, st,
                                  el.events, el.idx, el.counts, el.Pcap, weights, I, J, (int)BJ,
                                  K_post,
                                  (double)c, (double)delta_theta_tau, (double)tau, always_exact,
                                  out_n, out_t, out_a, out_x, out_z));
    return check_launch("lif_solve_wave_kernel");
  }
  const int cus = cu_count();
  // one spike per neuron and fewer than two lane-solver waves per SIMD: the walks split into
  // time segments (lif_solve_seg_kernel), about four waves per SIMD in all
  const long long lane_waves = (long long)B * ((J + 63) / 64);
  if (K_post == 1 && !always_exact && any_events &&
      (solver_mode == 5 || (solver_mode == 0 && lane_waves < 2LL * 4 * cus))) {
    int nseg = (int)std::max<long long>(
        2, std::min<long long>(kSegMax, (16LL * cus + lane_waves - 1) / std::max(1LL, lane_waves)));
    if (solver_blocks > 0) nseg = std::min(solver_blocks, kSegMax);  // segments, for tuning
    const size_t wt_bytes = sizeof(float) * (size_t)I * Jp;
    BATS_REQUIRE(wt_bytes < 0x7fffffffULL, "W^T above 2 GiB");
    BATS_TIMED_IF(timed, "lif_solve_seg_kernel", st,
               hipLaunchKernelGGL(lif_solve_seg_kernel, dim3((J + 63) / 64, B), dim3(64 * nseg), 0,
                                  st, el.events, el.fast, el.idx, el.counts, el.Pcap, WT, J, Jp,
                                  (int)wt_bytes, (double)c, (double)delta_theta_tau, (double)tau,
                                  nseg, out_n, out_t, out_a, out_x, out_z));
    return check_launch("lif_solve_seg_kernel");
  }
  // blocks per sample, and lanes per block rounded up to whole waves; a batch with fewer
  // samples than the device has CUs spreads each sample's neurons over more workgroups (of
  // at least one wave), so that the launch reaches every CU
  int nblk = (J + kSolveMaxThreads - 1) / kSolveMaxThreads;
  if ((long long)B * nblk < cus) nblk = std::max(nblk, std::min((cus + B - 1) / B, (J + 63) / 64));
  if (solver_blocks > 0) nblk = std::max(1, std::min(solver_blocks, (J + 63) / 64));
  const int per = (J + nblk - 1) / nblk;
  const int threads = (per + 63) / 64 * 64;
  dim3 sg((J + threads - 1) / threads, B);
  // staged event records when the launch has fewer than 3 waves per SIMD (4 SIMDs per CU):
  // with more, the other waves hide the per-event scalar loads and staging costs a little
  const long long waves = (long long)B * sg.x * (threads / 64);
  const bool staged = solver_mode == 3 ||
                      (solver_mode != 4 && waves < 3LL * 4 * (long long)cus);
  const size_t wt_bytes = sizeof(float) * (size_t)I * Jp;
  BATS_REQUIRE(wt_bytes < 0x7fffffffULL, "W^T above 2 GiB");
  const auto seq = always_exact ? (staged ? lif_solve_seq_kernel<true, true>
                                          : lif_solve_seq_kernel<true, false>)
                                : (staged ? lif_solve_seq_kernel<false, true>
                                          : lif_solve_seq_kernel<false, false>);
  ChainOut co{};
  if (chain && sg.x == 1) {
    // every sample's keys must fit the LDS room or its global key region
    long long n2 = 1;
    while (n2 < (long long)J * K_post) n2 <<= 1;
    if (n2 <= kChainLdsKeys || chain->gcap >= n2) {
      // room for two halves (the merge passes) up to the LDS limit
      chain->lds_keys = (int)std::min<long long>(2 * n2, kChainLdsKeys);
      co = *chain;
    }
  }
  const size_t lds = seq_solver_lds_bytes(threads, staged, co.lds_keys);
  BATS_REQUIRE(allow_seq_solver_lds() == 0, "could not raise the solver's LDS limit");
  BATS_TIMED_IF(timed, "lif_solve_seq_kernel", st,
             hipLaunchKernelGGL(seq, sg, dim3(threads), lds, st, el.events, el.fast, el.idx,
                                el.counts, el.Pcap, WT, J, Jp, (int)wt_bytes, K_post, (double)c,
                                (double)delta_theta_tau, (double)tau, out_n, out_t, out_a, out_x,
                                out_z, co));
  return check_launch("lif_solve_seq_kernel");
}

static int lif_forward_impl(void *stream, int B, int I, int K_pre, int J, int K_post,
                            const float *pre_times, const float *weights, float tau_s,
                            float tau, float c, float delta_theta_tau, float max_simulation,
                            int *out_n, float *out_t, float *out_a, float *out_x, float *out_z,
                            float *pre_exp_tau_s, float *pre_exp_tau, int solver_mode,
                            int solver_blocks, bool timed, void *workspace,
                            size_t workspace_bytes, ChainOut *chain = nullptr) {
  BATS_REQUIRE(B >= 0 && I >= 0 && K_pre >= 0 && J >= 0, "negative size");
  BATS_REQUIRE(K_post >= 1, "max_n_spike (K_post) must be >= 1");
  BATS_REQUIRE((double)B * J * K_post < 2147483647.0 * 8, "output too large");
  BATS_REQUIRE(B <= 65535, "batch above 65535");
  BATS_REQUIRE(solver_mode >= 0 && solver_mode <= 5,
               "solver_mode must be 0 (auto), 1 (lane per neuron), 2 (wave per neuron), 3 / 4 "
               "(lane per neuron, staged / scalar event records) or 5 (time segments, one "
               "spike per neuron)");
  BATS_REQUIRE(solver_blocks >= 0, "solver_blocks must be >= 0 (0: automatic)");
  if (B == 0 || J == 0) return 0;
  BATS_REQUIRE(out_n && out_t && out_a && out_x && out_z, "null output pointer");
  BATS_REQUIRE(tau_s > 0.f && tau > 0.f, "tau_s and tau must be positive");
  const int S = I * K_pre;
  hipStream_t st = as_stream(stream);
  const size_t need = fwd_bytes(B, I, K_pre, J);
  BATS_REQUIRE(workspace != nullptr, "null workspace");
  BATS_REQUIRE(workspace_bytes >= need, "workspace too small");
  char *ws = static_cast<char *>(workspace);
  float *WT = reinterpret_cast<float *>(ws);
  const int Jp = wt_stride(J);
  void *ev_ws = ws + align_up(sizeof(float) * (size_t)I * Jp, 256);
  const int always_exact = (c > 0.f) ? 0 : 1;
  const long long BJ = (long long)B * J;
  const bool wave_mode = solver_mode == 2 || (solver_mode == 0 && BJ < kSeqMinLanes);
  if (S > 0) {
    BATS_REQUIRE(pre_times && weights && pre_exp_tau_s && pre_exp_tau, "null input pointer");
    if (!wave_mode) {  // the wave-per-neuron solver reads W's rows directly
      dim3 tg((I + 31) / 32, (Jp + 31) / 32);
      hipLaunchKernelGGL(transpose_kernel, tg, dim3(256), 0, st, weights, WT, J, I, Jp);
      if (int r = check_launch("transpose_kernel")) return r;
    }
  }
  EventLists el;
  if (int r = build_event_lists(st, B, S, K_pre, pre_times, tau_s, tau, c, max_simulation,
                                pre_exp_tau_s, pre_exp_tau, ev_ws, &el, timed))
    return r;
  return solve_event_lists(st, B, I, J, K_post, el, S > 0, weights, WT, Jp, wave_mode,
                           solver_mode, solver_blocks, c, delta_theta_tau, tau, out_n, out_t,
                           out_a, out_x, out_z, timed, chain);
}

// ------------------------------------------------------------------ chained layers
// The next dense layer's event lists from this layer's spike trains out_t [B, J, K_post]: the
// chain's regions (next_events, laid out by event_lists_at for S = J * K_post slots) and
// constants. `built` says whether the lane solver already wrote them; otherwise they are built
// here with the per-sample sort, from the spike trains.
static int prepare_chain(int B, int J, int K_post, float next_tau_s, float next_tau,
                         float next_c, float next_max_simulation, float *next_pre_exp_tau_s,
                         float *next_pre_exp_tau, void *next_events, size_t next_events_bytes,
                         ChainOut *ch) {
  BATS_REQUIRE(next_tau_s > 0.f && next_tau > 0.f, "next layer: tau_s and tau must be positive");
  const int S2 = J * K_post;
  BATS_REQUIRE(next_events != nullptr && next_pre_exp_tau_s && next_pre_exp_tau,
               "next layer: null event-list or exp pointer");
  BATS_REQUIRE(next_events_bytes >= event_lists_bytes(B, S2), "next layer: event lists too small");
  *ch = ChainOut{};
  int gcap = 0;
  const EventLists el = event_lists_at(next_events, B, S2, &ch->gkeys, &gcap);
  ch->events = el.events;
  ch->fast = el.fast;
  ch->idx = el.idx;
  ch->counts = el.counts;
  ch->pre_es = next_pre_exp_tau_s;
  ch->pre_e = next_pre_exp_tau;
  ch->Pcap = el.Pcap;
  ch->gcap = gcap;
  ch->lds_keys = 0;
  ch->tau_s = next_tau_s;
  ch->tau = next_tau;
  ch->c = next_c;
  ch->tmax = next_max_simulation;
  return 0;
}

static int finish_chain(hipStream_t st, int B, int J, int K_post, const float *out_t,
                        const ChainOut &ch, void *next_events) {
  if (ch.lds_keys > 0) return 0;  // built by the lane solver
  EventLists el;
  return build_event_lists(st, B, J * K_post, K_post, out_t, ch.tau_s, ch.tau, ch.c, ch.tmax,
                           ch.pre_es, ch.pre_e, next_events, &el);
}

extern "C" int bats_lif_forward(void *stream, int B, int I, int K_pre, int J, int K_post,
                                const float *pre_times, const float *weights, float tau_s,
                                float tau, float c, float delta_theta_tau,
                                float max_simulation, int *out_n, float *out_t, float *out_a,
                                float *out_x, float *out_z, float *pre_exp_tau_s,
                                float *pre_exp_tau, void *workspace, size_t workspace_bytes) {
  return lif_forward_impl(stream, B, I, K_pre, J, K_post, pre_times, weights, tau_s, tau, c,
                          delta_theta_tau, max_simulation, out_n, out_t, out_a, out_x, out_z,
                          pre_exp_tau_s, pre_exp_tau, 0, 0, true, workspace, workspace_bytes);
}

extern "C" int bats_lif_forward_solver(void *stream, int B, int I, int K_pre, int J, int K_post,
                                       const float *pre_times, const float *weights,
                                       float tau_s, float tau, float c, float delta_theta_tau,
                                       float max_simulation, int *out_n, float *out_t,
                                       float *out_a, float *out_x, float *out_z,
                                       float *pre_exp_tau_s, float *pre_exp_tau,
                                       int solver_mode, int solver_blocks, void *workspace,
                                       size_t workspace_bytes) {
  return lif_forward_impl(stream, B, I, K_pre, J, K_post, pre_times, weights, tau_s, tau, c,
                          delta_theta_tau, max_simulation, out_n, out_t, out_a, out_x, out_z,
                          pre_exp_tau_s, pre_exp_tau, solver_mode, solver_blocks, true,
                          workspace, workspace_bytes);
}

extern "C" size_t bats_event_lists_size(int B, int S) {
  if (B < 0 || S < 0) return 0;
  return event_lists_bytes(B, S);
}

extern "C" int bats_lif_forward_chain(
    void *stream, int B, int I, int K_pre, int J, int K_post, const float *pre_times,
    const float *weights, float tau_s, float tau, float c, float delta_theta_tau,
    float max_simulation, int *out_n, float *out_t, float *out_a, float *out_x, float *out_z,
    float *pre_exp_tau_s, float *pre_exp_tau, float next_tau_s, float next_tau, float next_c,
    float next_max_simulation, float *next_pre_exp_tau_s, float *next_pre_exp_tau,
    void *next_events, size_t next_events_bytes, void *workspace, size_t workspace_bytes) {
  BATS_REQUIRE(B >= 0 && J >= 0 && K_post >= 1, "bad sizes");
  if (B == 0) return 0;
  ChainOut ch;
  if (int r = prepare_chain(B, J, K_post, next_tau_s, next_tau, next_c, next_max_simulation,
                            next_pre_exp_tau_s, next_pre_exp_tau, next_events, next_events_bytes,
                            &ch))
    return r;
  if (int r = lif_forward_impl(stream, B, I, K_pre, J, K_post, pre_times, weights, tau_s, tau, c,
                               delta_theta_tau, max_simulation, out_n, out_t, out_a, out_x,
                               out_z, pre_exp_tau_s, pre_exp_tau, 0, 0, true, workspace,
                               workspace_bytes, &ch))
    return r;
  return finish_chain(as_stream(stream), B, J, K_post, out_t, ch, next_events);
}

extern "C" size_t bats_lif_forward_events_workspace_size(int I, int J) {
  if (I < 0 || J < 0) return 0;
  return std::max<size_t>(256, align_up(sizeof(float) * (size_t)I * wt_stride(J), 256));
}

extern "C" int bats_lif_forward_events(
    void *stream, int B, int I, int K_pre, int J, int K_post, const void *events,
    size_t events_bytes, const float *weights, float tau_s, float tau, float c,
    float delta_theta_tau, int *out_n, float *out_t, float *out_a, float *out_x, float *out_z,
    float next_tau_s, float next_tau, float next_c, float next_max_simulation,
    float *next_pre_exp_tau_s, float *next_pre_exp_tau, void *next_events,
    size_t next_events_bytes, void *workspace, size_t workspace_bytes) {
  BATS_REQUIRE(B >= 0 && I >= 0 && K_pre >= 0 && J >= 0, "negative size");
  BATS_REQUIRE(K_post >= 1, "max_n_spike (K_post) must be >= 1");
  BATS_REQUIRE((double)B * J * K_post < 2147483647.0 * 8, "output too large");
  BATS_REQUIRE(B <= 65535, "batch above 65535");
  if (B == 0 || J == 0) return 0;
  BATS_REQUIRE(out_n && out_t && out_a && out_x && out_z, "null output pointer");
  BATS_REQUIRE(tau_s > 0.f && tau > 0.f, "tau_s and tau must be positive");
  const int S = I * K_pre;
  BATS_REQUIRE(events != nullptr && events_bytes >= event_lists_bytes(B, S),
               "event lists missing or too small");
  BATS_REQUIRE(workspace != nullptr &&
                   workspace_bytes >= bats_lif_forward_events_workspace_size(I, J),
               "workspace too small");
  BATS_REQUIRE(S == 0 || weights, "null weights");
  ChainOut ch;
  ChainOut *chp = nullptr;
  if (next_events) {
    if (int r = prepare_chain(B, J, K_post, next_tau_s, next_tau, next_c, next_max_simulation,
                              next_pre_exp_tau_s, next_pre_exp_tau, next_events,
                              next_events_bytes, &ch))
      return r;
    chp = &ch;
  }
  hipStream_t st = as_stream(stream);
  float *WT = static_cast<float *>(workspace);
  const int Jp = wt_stride(J);
  const bool wave_mode = (long long)B * J < kSeqMinLanes;
  if (S > 0 && !wave_mode) {
    dim3 tg((I + 31) / 32, (Jp + 31) / 32);
    hipLaunchKernelGGL(transpose_kernel, tg, dim3(256), 0, st, weights, WT, J, I, Jp);
    if (int r = check_launch("transpose_kernel")) return r;
  }
  const EventLists el = event_lists_at(const_cast<void *>(events), B, S, nullptr, nullptr);
  if (int r = solve_event_lists(st, B, I, J, K_post, el, S > 0, weights, WT, Jp, wave_mode, 0, 0,
                                c, delta_theta_tau, tau, out_n, out_t, out_a, out_x, out_z, true,
                                chp))
    return r;
  return chp ? finish_chain(st, B, J, K_post, out_t, ch, next_events) : 0;
}

// ------------------------------------------------------------------ forward from packed lists
// The first layer of an inference server: its inputs arrive as packed, time-ordered event
// lists, so the Event lists are written from them directly (packed_events_kernel) instead of
// going through per-neuron buffers and the per-sample sort. Inference only: no per-slot exps
// are written for a backward.
static size_t packed_lists_bytes(int B, int P_max) {
  const int Pcap = event_capacity(P_max);
  size_t off = 0;
  off = align_up(off + sizeof(Event) * (size_t)B * Pcap, 256);
  off = align_up(off + sizeof(FastEvent) * (size_t)B * Pcap, 256);
  off = align_up(off + sizeof(int) * (size_t)B * Pcap, 256);
  off = align_up(off + sizeof(int) * (size_t)B, 256);
  return off;
}

extern "C" size_t bats_lif_forward_packed_workspace_size(int B, int I, int P_max, int J) {
  if (B < 0 || I < 0 || P_max < 0 || J < 0) return 0;
  return align_up(sizeof(float) * (size_t)I * wt_stride(J), 256) + packed_lists_bytes(B, P_max);
}

static int lif_forward_packed_impl(void *stream, int B, int I, int J, int K_post, int P_max,
                                   const int *row_offsets, const int *indices,
                                   const float *times, const float *weights, float tau_s,
                                   float tau, float c, float delta_theta_tau,
                                   float max_simulation, int *out_n, float *out_t,
                                   float *out_a, float *out_x, float *out_z, void *workspace,
                                   size_t workspace_bytes, ChainOut *chain) {
  BATS_REQUIRE(B >= 0 && I >= 0 && J >= 0 && P_max >= 0, "negative size");
  BATS_REQUIRE(K_post >= 1, "max_n_spike (K_post) must be >= 1");
  BATS_REQUIRE((double)B * J * K_post < 2147483647.0 * 8, "output too large");
  BATS_REQUIRE(B <= 65535, "batch above 65535");
  BATS_REQUIRE((int64_t)B * event_capacity(P_max) < ((int64_t)1 << 31), "too many events");
  if (B == 0 || J == 0) return 0;
  BATS_REQUIRE(out_n && out_t && out_a && out_x && out_z, "null output pointer");
  BATS_REQUIRE(row_offsets && weights, "null input pointer");
  BATS_REQUIRE(P_max == 0 || (indices && times), "null input pointer");
  BATS_REQUIRE(tau_s > 0.f && tau > 0.f, "tau_s and tau must be positive");
  hipStream_t st = as_stream(stream);
  BATS_REQUIRE(workspace != nullptr, "null workspace");
  BATS_REQUIRE(workspace_bytes >= bats_lif_forward_packed_workspace_size(B, I, P_max, J),
               "workspace too small");
  char *ws = static_cast<char *>(workspace);
  float *WT = reinterpret_cast<float *>(ws);
  const int Jp = wt_stride(J);
  const bool wave_mode = (long long)B * J < kSeqMinLanes;
  const bool any = P_max > 0 && I > 0;
  if (any && !wave_mode) {
    dim3 tg((I + 31) / 32, (Jp + 31) / 32);
    hipLaunchKernelGGL(transpose_kernel, tg, dim3(256), 0, st, weights, WT, J, I, Jp);
    if (int r = check_launch("transpose_kernel")) return r;
  }
  EventLists el;
  const int Pcap = event_capacity(P_max);
  char *base = ws + align_up(sizeof(float) * (size_t)I * Jp, 256);
  size_t off = 0;
  el.events = reinterpret_cast<Event *>(base + off);
  off = align_up(off + sizeof(Event) * (size_t)B * Pcap, 256);
  el.fast = reinterpret_cast<FastEvent *>(base + off);
  off = align_up(off + sizeof(FastEvent) * (size_t)B * Pcap, 256);
  el.idx = reinterpret_cast<int *>(base + off);
  off = align_up(off + sizeof(int) * (size_t)B * Pcap, 256);
  el.counts = reinterpret_cast<int *>(base + off);
  el.Pcap = Pcap;
  BATS_TIMED("packed_events_kernel", st,
             hipLaunchKernelGGL(packed_events_kernel, dim3(B), dim3(256), 0, st, row_offsets,
                                indices, times, any ? I : 0, tau_s, tau, c, max_simulation, P_max,
                                Pcap, el.events, el.fast, el.idx, el.counts));
  if (int r = check_launch("packed_events_kernel")) return r;
  return solve_event_lists(st, B, I, J, K_post, el, any, weights, WT, Jp, wave_mode, 0, 0, c,
                           delta_theta_tau, tau, out_n, out_t, out_a, out_x, out_z, true, chain);
}

extern "C" int bats_lif_forward_packed(void *stream, int B, int I, int J, int K_post, int P_max,
                                       const int *row_offsets, const int *indices,
                                       const float *times, const float *weights, float tau_s,
                                       float tau, float c, float delta_theta_tau,
                                       float max_simulation, int *out_n, float *out_t,
                                       float *out_a, float *out_x, float *out_z,
                                       void *workspace, size_t workspace_bytes) {
  return lif_forward_packed_impl(stream, B, I, J, K_post, P_max, row_offsets, indices, times,
                                 weights, tau_s, tau, c, delta_theta_tau, max_simulation, out_n,
                                 out_t, out_a, out_x, out_z, workspace, workspace_bytes,
                                 nullptr);
}

extern "C" int bats_lif_forward_packed_chain(
    void *stream, int B, int I, int J, int K_post, int P_max, const int *row_offsets,
    const int *indices, const float *times, const float *weights, float tau_s, float tau,
    float c, float delta_theta_tau, float max_simulation, int *out_n, float *out_t,
    float *out_a, float *out_x, float *out_z, float next_tau_s, float next_tau, float next_c,
    float next_max_simulation, float *next_pre_exp_tau_s, float *next_pre_exp_tau,
    void *next_events, size_t next_events_bytes, void *workspace, size_t workspace_bytes) {
  BATS_REQUIRE(B >= 0 && J >= 0 && K_post >= 1, "bad sizes");
  if (B == 0) return 0;
  ChainOut ch;
  if (int r = prepare_chain(B, J, K_post, next_tau_s, next_tau, next_c, next_max_simulation,
                            next_pre_exp_tau_s, next_pre_exp_tau, next_events, next_events_bytes,
                            &ch))
    return r;
  if (int r = lif_forward_packed_impl(stream, B, I, J, K_post, P_max, row_offsets, indices,
                                      times, weights, tau_s, tau, c, delta_theta_tau,
                                      max_simulation, out_n, out_t, out_a, out_x, out_z,
                                      workspace, workspace_bytes, &ch))
    return r;
  return finish_chain(as_stream(stream), B, J, K_post, out_t, ch, next_events);
}

// ------------------------------------------------------------------ forward without allocation
// The forward's workspace (W^T and the per-sample event records) in a file-scope device array
// of a fixed size, for callers that cannot allocate. Every GPU the library is loaded on holds
// its own copy until the process ends; calls on one GPU share it, so calls that overlap on one
// GPU (two streams, two host threads) race. Nothing in it needs to start at zero: the sort
// writes every record the solver reads, on every call.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices];
int g_scratch_status[kMaxDevices];
std::once_flag g_scratch_once[kMaxDevices];
}  // namespace

// Device address of `scratch` on the current device, looked up once per device with that
// device current (the host-side name of the array is not a device address).
static int scratch_address(void **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail("bats_lif_forward_noalloc: no usable current device");
  std::call_once(g_scratch_once[dev], [dev]() {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = p;
    g_scratch_status[dev] = (e == hipSuccess && p != nullptr) ? 0 : (e == hipSuccess ? -1 : (int)e);
  });
  if (g_scratch_status[dev] != 0)
    return fail("bats_lif_forward_noalloc: hipGetSymbolAddress of the scratch array failed (" +
                std::to_string(g_scratch_status[dev]) + ")");
  *out = g_scratch_ptr[dev];
  return 0;
}

extern "C" size_t bats_lif_forward_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

extern "C" int bats_lif_forward_noalloc(void *stream, int B, int I, int K_pre, int J, int K_post,
                                        const float *pre_times, const float *weights,
                                        float tau_s, float tau, float c, float delta_theta_tau,
                                        float max_simulation, int *out_n, float *out_t,
                                        float *out_a, float *out_x, float *out_z,
                                        float *pre_exp_tau_s, float *pre_exp_tau) {
  BATS_REQUIRE(B >= 0 && I >= 0 && K_pre >= 0 && J >= 0, "negative size");
  const size_t need = fwd_bytes(B, I, K_pre, J);
  BATS_REQUIRE(need <= kNoAllocScratchBytes,
               "these sizes need " + std::to_string(need) + " bytes of workspace, above the " +
                   std::to_string(kNoAllocScratchBytes) +
                   "-byte scratch array of the no-allocation entry point (use bats_lif_forward "
                   "with a caller-allocated workspace)");
  void *ws = nullptr;
  if (int r = scratch_address(&ws)) return r;
  return lif_forward_impl(stream, B, I, K_pre, J, K_post, pre_times, weights, tau_s, tau, c,
                          delta_theta_tau, max_simulation, out_n, out_t, out_a, out_x, out_z,
                          pre_exp_tau_s, pre_exp_tau, 0, 0, false, ws, kNoAllocScratchBytes);
}
