// Binned backward of dense layers with multi-spike pre layers (gfx950).
//
// With V1_j(t) = sum_{k: t_jk > t} d1_jk, V2_j likewise (d1 = e f1, d2 = e f2 from
// row_prep_kernel), the reference's weight gradient and pre-spike errors regroup as
//   dW[j,i]        = sum_b sum_m  es_m V1_j(t_m) - e_m V2_j(t_m)            (m: spikes of i)
//   pre_err[b,i,m] = sum_j W[j,i] (V1_j(t_m) es_m / tau_s - V2_j(t_m) e_m / tau)
// (compute_weights_gradient.cu:33-45, propagate_errors_to_pre_spikes.cu:41-53). Each sample's
// post-spike time range [lo, hi] is cut into kBins equal bins; a time t < lo is bin -1 and a
// pre time t > hi bin kBins. The bin of a time is monotone in the time, so a pre spike in an
// earlier bin than a post spike precedes it and one in a later bin follows it; only pairs in
// one bin need the time comparison. With
//   S1[b,r,j] = sum of d1 over row j's spikes in bin r,   S2 likewise,
//   P1[b,r,i] = sum of es over neuron i's spikes in bins < r,   P2 likewise with e,
// the weight gradient is the GEMM   sum_{(b,r)} S1[b,r,j] P1[b,r,i] - S2[b,r,j] P2[b,r,i]
// plus the pairs inside one bin, and the pre-spike errors use Q = [S1; -S2] W:
//   pre_err = es/tau_s sum_{r > bin(m)} Q1[b,r,i] + e/tau sum_{r > bin(m)} Q2[b,r,i]
// plus the pairs inside the pre spike's bin. The GEMMs run on the matrix cores in float32
// (gemm_f32.hpp): products and sums are rounded to float32, which keeps the results well inside
// the 1e-4 relative bar; the in-bin pairs are summed as in the spike-lane kernels (float32
// terms within a tile, float64 across).
//
// Kernels: bin_range_kernel (per-sample [lo, hi]), post_bins_kernel (S and the tile records
// ordered by (bin, row) with per-bin offsets), pre_bins_kernel (P and the pre spikes listed
// bin by bin), the weight-gradient GEMM and wgrad_bin_kernel (in-bin pairs); then, for the
// pre-spike errors, the transposed weights, the GEMM for Q and pre_err_bin_kernel.
#include "backward_bins.hpp"
#include "gemm_f32.hpp"

#include <algorithm>

namespace bats {

constexpr int kBT = kWT;  // post rows per tile record (lanes of one 16-lane group)

__device__ __forceinline__ int post_bin(float t, float lo, float scale) {
  const int r = (int)floorf((t - lo) * scale);
  return min(max(r, 0), kBins - 1);
}
__device__ __forceinline__ int pre_bin(float t, float lo, float hi, float scale) {
  if (t < lo) return -1;
  if (t > hi) return kBins;
  return post_bin(t, lo, scale);
}

// ------------------------------------------------------------------ per-sample time range
// One workgroup per sample: the smallest first spike and the largest last spike of its rows.
// range[b] = {lo, hi, scale = kBins / (hi - lo) (0 when hi == lo or no spike)}.
__global__ __launch_bounds__(256) void bin_range_kernel(int J, int K, const int *__restrict__ n_eff,
                                                       const float *__restrict__ t,
                                                       float4 *__restrict__ range) {
  __shared__ float s_lo[4], s_hi[4];
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  float lo = kInfF, hi = -kInfF;
  for (int j = tid; j < J; j += 256) {
    const size_t r = (size_t)b * J + j;
    const int ne = n_eff[r];
    if (ne > 0) {
      lo = fminf(lo, t[r * K]);
      hi = fmaxf(hi, t[r * K + ne - 1]);
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    lo = fminf(lo, __shfl_xor(lo, d, 64));
    hi = fmaxf(hi, __shfl_xor(hi, d, 64));
  }
  if (lane == 0) {
    s_lo[wv] = lo;
    s_hi[wv] = hi;
  }
  __syncthreads();
  if (tid == 0) {
    lo = fminf(fminf(s_lo[0], s_lo[1]), fminf(s_lo[2], s_lo[3]));
    hi = fmaxf(fmaxf(s_hi[0], s_hi[1]), fmaxf(s_hi[2], s_hi[3]));
    if (!(lo <= hi)) lo = hi = 0.f;  // no post spike in the sample
    const float w = hi - lo;
    float scale = w > 0.f ? (float)kBins / w : 0.f;
    if (!(scale < kInfF)) scale = 0.f;  // a range too narrow to divide: one bin (still exact)
    range[b] = make_float4(lo, hi, scale, 0.f);
  }
}

// ------------------------------------------------------------------ post side
// Lanes are (sample, tile of kBT rows, row): the row's spikes give its per-bin sums S1, -S2
// (float64 sums of the float32 terms, one float32 rounding), written as rows b*2*kBins + r and
// b*2*kBins + kBins + r, column j, of A; and its entries {t, d1, d2, (row << 9) | bin} in the
// tile record (after its count slot), ordered by (bin, row) with the offsets of the bins in
// boff[tile][0..kBins] (boff[tile][kBins] = the tile's spike count). A row's spikes are in
// time order, so their bins do not decrease.
__global__ __launch_bounds__(256) void post_bins_kernel(
    int B, int J, int K, int gj, const int *__restrict__ n_eff, const float *__restrict__ t,
    const float2 *__restrict__ D, const float4 *__restrict__ range, int lda,
    float *__restrict__ A, float4 *__restrict__ recb, int *__restrict__ boff) {
  const int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tile = lane / kBT;
  const int u = (int)(lane % kBT);
  const bool valid = tile < (int64_t)B * gj;
  const int b = valid ? (int)(tile / gj) : 0;
  const int j = valid ? (int)(tile % gj) * kBT + u : 0;
  const bool active = valid && j < J;
  const size_t r = active ? (size_t)b * J + j : 0;
  const int ne = active ? n_eff[r] : 0;
  const float4 rg = valid ? range[b] : make_float4(0.f, 0.f, 0.f, 0.f);
  int cnt[kBins];
  double s1[kBins], s2[kBins];
#pragma unroll
  for (int q = 0; q < kBins; ++q) {
    cnt[q] = 0;
    s1[q] = 0.0;
    s2[q] = 0.0;
  }
  for (int k = 0; k < ne; ++k) {
    const int bin = post_bin(t[r * K + k], rg.x, rg.z);
    const float2 d = D[r * K + k];
#pragma unroll
    for (int q = 0; q < kBins; ++q) {
      if (q == bin) {
        cnt[q] += 1;
        s1[q] += (double)d.x;
        s2[q] += (double)d.y;
      }
    }
  }
  if (active) {
#pragma unroll
    for (int q = 0; q < kBins; ++q) {
      A[((size_t)b * 2 * kBins + q) * lda + j] = (float)s1[q];
      A[((size_t)b * 2 * kBins + kBins + q) * lda + j] = (float)(-s2[q]);
    }
  }
  // (bin, row) order: the row's entries of bin q start at base[q] + (spikes of bin q in the
  // tile's earlier rows)
  int excl[kBins], base[kBins];
  int run_total = 0;
#pragma unroll
  for (int q = 0; q < kBins; ++q) {
    int incl = cnt[q];
#pragma unroll
    for (int d = 1; d < kBT; d <<= 1) {
      const int y = __shfl_up(incl, d, kBT);
      if (u >= d) incl += y;
    }
    excl[q] = incl - cnt[q];
    base[q] = run_total;
    run_total += __shfl(incl, kBT - 1, kBT);
  }
  if (!valid) return;
  float4 *rec = recb + (size_t)tile * tile_rec_slots(K);
  if (u == 0) {
    int *bo = boff + (size_t)tile * (kBins + 1);
#pragma unroll
    for (int q = 0; q < kBins; ++q) bo[q] = base[q];
    bo[kBins] = run_total;
    rec[0] = make_float4(__int_as_float(run_total), 0.f, 0.f, 0.f);
  }
  float4 *rb = rec + kRecHead;
  int seen[kBins];
#pragma unroll
  for (int q = 0; q < kBins; ++q) seen[q] = 0;
  for (int k = 0; k < ne; ++k) {
    const float tk = t[r * K + k];
    const int bin = post_bin(tk, rg.x, rg.z);
    const float2 d = D[r * K + k];
    int pos = 0;
#pragma unroll
    for (int q = 0; q < kBins; ++q) {
      if (q == bin) {
        pos = base[q] + excl[q] + seen[q];
        seen[q] += 1;
      }
    }
    rb[pos] = make_float4(tk, d.x, d.y, __int_as_float((u << 9) | bin));
  }
}

// ------------------------------------------------------------------ pre side
// One workgroup per sample; a thread owns npt consecutive pre neurons. Per neuron:
// the exclusive per-bin prefix sums of es and e (bins -1 .. r-1), written as rows
// b*2*kBins + r and b*2*kBins + kBins + r, column i, of Bm; and the neuron's spikes in the
// sample's bin-major list: bins -1 .. kBins in turn, within a bin by (thread, neuron, spike).
// spk entries are {t, es, e, i | (bin + 1) << 24}, slot the index i*K_pre + m of the spike,
// poff[b][0 .. kBins+2] the start of each bin (bin -1 first) and the total.
constexpr int kPreMaxThreads = 1024;
constexpr int kPreBinsN = kBins + 2;  // bins -1 .. kBins

// Leading finite slots of a spike row, eight loads in flight at a time.
__device__ __forceinline__ int leading_finite_row(const float *__restrict__ row, int K) {
  int n = 0;
  for (int m0 = 0; m0 < K; m0 += 8) {
    bool run = true;
    int add = 0;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const float v = m0 + m < K ? row[m0 + m] : kInfF;
      run = run & (bool)isfinite(v);
      add += run ? 1 : 0;
    }
    n += add;
    if (add < 8) break;
  }
  return n;
}

// A neuron's spike counts per bin (-1 .. kBins), and with `sums` its per-bin es / e sums.
template <bool kSums>
__device__ __forceinline__ void neuron_bins(const float *__restrict__ t, const float *__restrict__ es,
                                            const float *__restrict__ e, int n, float4 rg,
                                            int (&cnt)[kPreBinsN], double (&hs)[kBins + 1],
                                            double (&he)[kBins + 1]) {
#pragma unroll
  for (int q = 0; q < kPreBinsN; ++q) cnt[q] = 0;
#pragma unroll
  for (int q = 0; q <= kBins; ++q) {
    hs[q] = 0.0;
    he[q] = 0.0;
  }
  for (int m = 0; m < n; ++m) {
    const int bin = pre_bin(t[m], rg.x, rg.y, rg.z);
    double vs = 0.0, ve = 0.0;
    if (kSums) {
      vs = (double)es[m];
      ve = (double)e[m];
    }
#pragma unroll
    for (int q = 0; q < kPreBinsN; ++q) {
      if (q == bin + 1) {
        cnt[q] += 1;
        if (kSums && q <= kBins) {
          hs[q] += vs;
          he[q] += ve;
        }
      }
    }
  }
}

// One workgroup (up to 1024 threads, one neuron per thread per round) per sample. With one
// round (I <= threads) a single pass does it all: each thread reads its neuron's spikes once
// and one block scan of the per-bin counts gives the spikes' positions and the bins' starts.
// Otherwise pass 1 counts every bin's spikes and pass 2 does the rest. Per neuron: the
// exclusive per-bin prefix sums of es and e (bins -1 .. r-1) as rows b*2*kBins + r and
// b*2*kBins + kBins + r, column i, of Bm, and the neuron's spikes in the sample's bin-major
// list: bins -1 .. kBins in turn, within a bin by neuron and spike. spk entries are
// {t, es, e, i | (bin + 1) << 24}, slot the index i*K_pre + m of the spike, poff[b][0 ..
// kBins+2] the start of each bin and the total.
__global__ __launch_bounds__(kPreMaxThreads) void pre_bins_kernel(
    int I, int K_pre, const float *__restrict__ pre_t, const float *__restrict__ pre_es,
    const float *__restrict__ pre_e, const float4 *__restrict__ range, int ldb,
    float *__restrict__ Bm, int cap, float4 *__restrict__ spk, int *__restrict__ slot,
    int *__restrict__ poff) {
  __shared__ int s_wave[kPreMaxThreads / 64][kPreBinsN];
  __shared__ int s_run[kPreBinsN];  // next free position of each bin
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nth = blockDim.x, nw = nth >> 6;
  const float4 rg = range[b];
  int cnt[kPreBinsN];
  double hs[kBins + 1], he[kBins + 1];
  float4 *o = spk + (size_t)b * cap;
  int *os = slot + (size_t)b * cap;
  // the neuron's spikes in the list from their positions pos[bin + 1] on
  auto list_spikes = [&](int i, size_t row, int n, int (&pos)[kPreBinsN]) {
    for (int m = 0; m < n; ++m) {
      const float tm = pre_t[row + m];
      const int bin = pre_bin(tm, rg.x, rg.y, rg.z);
      int at = 0;
#pragma unroll
      for (int q = 0; q < kPreBinsN; ++q) {
        if (q == bin + 1) {
          at = pos[q];
          pos[q] += 1;
        }
      }
      o[at] = make_float4(tm, pre_es[row + m], pre_e[row + m],
                          __int_as_float(i | ((bin + 1) << 24)));
      os[at] = i * K_pre + m;
    }
  };
  auto write_prefix = [&](int i) {
    double ps = 0.0, pe = 0.0;
#pragma unroll
    for (int q = 0; q < kBins; ++q) {  // P[r] = bins -1 .. r-1
      ps += hs[q];
      pe += he[q];
      Bm[((size_t)b * 2 * kBins + q) * ldb + i] = (float)ps;
      Bm[((size_t)b * 2 * kBins + kBins + q) * ldb + i] = (float)pe;
    }
  };
  if (I <= nth) {
    // one neuron per thread: its counts and sums in one read, one block scan gives both its
    // positions and the bins' totals (the same list as the two passes below)
    const int i = tid;
    const bool act = i < I;
    const size_t row = ((size_t)b * I + (act ? i : 0)) * K_pre;
    const int n = act ? leading_finite_row(pre_t + row, K_pre) : 0;
    neuron_bins<true>(pre_t + row, pre_es + row, pre_e + row, n, rg, cnt, hs, he);
    if (act) write_prefix(i);
    int pos[kPreBinsN];
#pragma unroll
    for (int q = 0; q < kPreBinsN; ++q) {
      int incl = cnt[q];
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
      }
      pos[q] = incl - cnt[q];
      if (lane == 63) s_wave[wv][q] = incl;
    }
    __syncthreads();
    int start = 0;
    int *po = poff + (size_t)b * (kPreBinsN + 1);
#pragma unroll
    for (int q = 0; q < kPreBinsN; ++q) {
      int woff = 0, tq = 0;
      for (int w = 0; w < nw; ++w) {
        const int c = s_wave[w][q];
        woff += w < wv ? c : 0;
        tq += c;
      }
      pos[q] += woff + start;
      if (tid == 0) po[q] = start;
      start += tq;
    }
    if (tid == 0) po[kPreBinsN] = start;
    if (act) list_spikes(i, row, n, pos);
    return;
  }
  // pass 1: per-bin totals
  int tot[kPreBinsN];
#pragma unroll
  for (int q = 0; q < kPreBinsN; ++q) tot[q] = 0;
  for (int i = tid; i < I; i += nth) {
    const size_t row = ((size_t)b * I + i) * K_pre;
    const int n = leading_finite_row(pre_t + row, K_pre);
    neuron_bins<false>(pre_t + row, nullptr, nullptr, n, rg, cnt, hs, he);
#pragma unroll
    for (int q = 0; q < kPreBinsN; ++q) tot[q] += cnt[q];
  }
#pragma unroll
  for (int q = 0; q < kPreBinsN; ++q) {
    int v = tot[q];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (lane == 0) s_wave[wv][q] = v;
  }
  __syncthreads();
  if (tid == 0) {
    int a = 0;
    int *po = poff + (size_t)b * (kPreBinsN + 1);
    for (int q = 0; q < kPreBinsN; ++q) {
      int sq = 0;
      for (int w = 0; w < nw; ++w) sq += s_wave[w][q];
      po[q] = a;
      s_run[q] = a;
      a += sq;
    }
    po[kPreBinsN] = a;
  }
  __syncthreads();
  // pass 2: prefix sums and the bin-major list, one block scan per round of nth neurons
  for (int ib = 0; ib < I; ib += nth) {
    const int i = ib + tid;
    const bool act = i < I;
    const size_t row = ((size_t)b * I + (act ? i : 0)) * K_pre;
    const int n = act ? leading_finite_row(pre_t + row, K_pre) : 0;
    neuron_bins<true>(pre_t + row, pre_es + row, pre_e + row, n, rg, cnt, hs, he);
    if (act) write_prefix(i);
    int pos[kPreBinsN];
#pragma unroll
    for (int q = 0; q < kPreBinsN; ++q) {
      int incl = cnt[q];
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
      }
      pos[q] = incl - cnt[q];
      if (lane == 63) s_wave[wv][q] = incl;
    }
    __syncthreads();
    int round_tot[kPreBinsN];
#pragma unroll
    for (int q = 0; q < kPreBinsN; ++q) {
      int woff = 0, t_ = 0;
      for (int w = 0; w < nw; ++w) {
        const int c = s_wave[w][q];
        woff += w < wv ? c : 0;
        t_ += c;
      }
      pos[q] += woff + s_run[q];
      round_tot[q] = t_;
    }
    __syncthreads();  // every thread has read s_wave and s_run
    if (tid < kPreBinsN) s_run[tid] += round_tot[tid];
    if (act) list_spikes(i, row, n, pos);
  }
}

// ------------------------------------------------------------------ in-bin weight gradient
// The spike-lane walk restricted to pairs in one bin: one workgroup per (tile of kBT rows,
// batch slice of up to kWalkMaxSamples samples) keeps acc[row][pre neuron] in float64 LDS.
// The slice's samples are taken in rounds of as many consecutive samples as the LDS beside acc
// holds the real entries of (packed, each record at its start in the round): their tile
// records are copied to LDS together, and the waves share out the 64-spike chunks of all the
// round's samples' pre spikes of bins 0 .. kBins-1 (bin-major, so a chunk spans few bins),
// each walking only the record entries of its chunk's bins and adding es d1 - e d2 where the
// bins match and t_pre < t_post. A lane
// sums one row's terms in a register (entries of one bin are row-ordered, so a row change is
// uniform) and adds them to acc with one no-return LDS atomic when the row ends. A wave
// requests its next chunk's spikes before walking the current one; there are two barriers
// per round, not per sample.
constexpr int kWalkMaxSamples = 64;  // samples per in-bin walk slice (one lane each in a scan)
constexpr int kWalkThreads = 1024;
constexpr int kWalkCopy = 4;  // record entries per thread in flight while a round is copied
constexpr int kWalkLds = 160 * 1024 - 1024;
// per sample: p0, p1, chunks, bin offsets 0 .. kBins, the record's start in the round's LDS
constexpr int kSInfo = 3 + kBins + 2;
constexpr int kSBase = 3 + kBins + 1;

__host__ __device__ inline size_t bin_walk_fixed_bytes(int I) {
  const size_t b = (size_t)kBT * I * sizeof(double) +
                   (size_t)(kWalkMaxSamples * kSInfo + kWalkMaxSamples + 1 + 1) * sizeof(int);
  return (b + 15) / 16 * 16;
}

// Record entries the LDS holds beside acc: a round takes consecutive samples while their
// records (only their real entries, packed) fit.
__host__ __device__ inline int bin_walk_entries(int I) {
  const long avail = (long)kWalkLds - (long)bin_walk_fixed_bytes(I);
  return avail > 0 ? (int)(avail / (long)sizeof(float4)) : 0;
}

// Whole records (kBT * K entries) the LDS holds: at least one is needed.
__host__ __device__ inline int bin_walk_group(int I, int K) {
  return min(bin_walk_entries(I) / (kBT * K), kWalkMaxSamples);
}

__host__ __device__ inline size_t bin_wgrad_lds_bytes(int I, int K) {
  (void)K;
  return bin_walk_fixed_bytes(I) + (size_t)bin_walk_entries(I) * sizeof(float4);
}

__global__ __launch_bounds__(kWalkThreads) void wgrad_bin_kernel(
    int B, int I, int J, int K, int gj, int b_per_slice, int capE, int cap,
    const float4 *__restrict__ spk, const int *__restrict__ poff, const float4 *__restrict__ recb,
    const int *__restrict__ boff, float *__restrict__ partial) {
  extern __shared__ __align__(16) char smem[];
  double *acc = reinterpret_cast<double *>(smem);  // [kBT][I]
  int *info = reinterpret_cast<int *>(acc + (size_t)kBT * I);  // [kWalkMaxSamples][kSInfo]
  int *rstart = info + kWalkMaxSamples * kSInfo;  // first sample of each round, then nS
  int *nrounds = rstart + kWalkMaxSamples + 1;
  float4 *rb = reinterpret_cast<float4 *>(smem + bin_walk_fixed_bytes(I));  // [capE]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int nth = kWalkThreads, nw = kWalkThreads / 64;
  const int tj = blockIdx.x, sl = blockIdx.y;
  const int j0 = tj * kBT, jn = min(kBT, J - j0);
  const int b0 = sl * b_per_slice, nS = max(0, min(B, b0 + b_per_slice) - b0);
  const int R = tile_rec_slots(K);
  for (int e = tid; e < kBT * I; e += nth) acc[e] = 0.0;
  if (tid < nS) {
    const int *po = poff + (size_t)(b0 + tid) * (kPreBinsN + 1);
    const int p0 = po[1], p1 = po[kBins + 1];  // the pre spikes of bins 0 .. kBins-1
    info[tid * kSInfo] = p0;
    info[tid * kSInfo + 1] = p1;
    info[tid * kSInfo + 2] = (p1 - p0 + 63) >> 6;
  }
  for (int e = tid; e < nS * (kBins + 1); e += nth) {
    const int s = e / (kBins + 1), q = e - s * (kBins + 1);
    info[s * kSInfo + 3 + q] = boff[((size_t)(b0 + s) * gj + tj) * (kBins + 1) + q];
  }
  __syncthreads();
  // rounds: consecutive samples whose records fit capE entries together (a record has at most
  // kBT * K <= capE entries), each record's start in the round
  if (tid == 0) {
    int nr = 0, used = 0;
    rstart[0] = 0;
    for (int s = 0; s < nS; ++s) {
      const int c = info[s * kSInfo + 3 + kBins];
      if (used + c > capE) {
        rstart[++nr] = s;
        used = 0;
      }
      info[s * kSInfo + kSBase] = used;
      used += c;
    }
    rstart[++nr] = nS;
    *nrounds = nS > 0 ? nr : 0;
  }
  __syncthreads();
  const int nr = *nrounds;
  for (int r = 0; r < nr; ++r) {
    const int g0 = rstart[r], gn = rstart[r + 1] - g0;
    const int *last = info + (g0 + gn - 1) * kSInfo;
    const int E = last[kSBase] + last[3 + kBins];
    // the round's records, packed, kWalkCopy entries per thread in flight together; an
    // entry's sample is the last of the round whose start is at or before it
    for (int e0 = 0; e0 < E; e0 += kWalkCopy * nth) {
      float4 v[kWalkCopy];
      int at[kWalkCopy];
#pragma unroll
      for (int h = 0; h < kWalkCopy; ++h) {
        const int e = e0 + tid + h * nth;
        at[h] = -1;
        if (e < E) {
          int lo = 0, hi = gn - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (info[(g0 + mid) * kSInfo + kSBase] <= e)
              lo = mid;
            else
              hi = mid - 1;
          }
          const int k = e - info[(g0 + lo) * kSInfo + kSBase];
          v[h] = recb[((size_t)(b0 + g0 + lo) * gj + tj) * R + kRecHead + k];
          at[h] = e;
        }
      }
#pragma unroll
      for (int h = 0; h < kWalkCopy; ++h)
        if (at[h] >= 0) rb[at[h]] = v[h];
    }
    __syncthreads();
    // chunk prefix of the round's samples: lane s holds the inclusive count up to sample s
    int incl = lane < gn ? info[(g0 + lane) * kSInfo + 2] : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(incl, d, 64);
      if (lane >= d) incl += y;
    }
    const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
    // sample, chunk in sample and spikes of chunk c
    struct Chunk {
      int s, p1;
      float4 S;
    };
    auto load_chunk = [&](int c) {
      Chunk ch;
      const int s = __builtin_amdgcn_readfirstlane(__popcll(__ballot(incl <= c)));
      const int before = s > 0 ? __builtin_amdgcn_readfirstlane(__shfl(incl, s - 1, 64)) : 0;
      const int *in = info + (g0 + s) * kSInfo;
      const int sidx = in[0] + (c - before) * 64 + lane;
      ch.s = s;
      ch.p1 = in[1];
      ch.S = sidx < ch.p1 ? spk[(size_t)(b0 + g0 + s) * cap + sidx]
                          : make_float4(kInfF, 0.f, 0.f, 0.f);
      return ch;
    };
    Chunk cur;
    if (wv < total) cur = load_chunk(wv);
    for (int c = wv; c < total; c += nw) {
      Chunk nxt;
      if (c + nw < total) nxt = load_chunk(c + nw);
      const float4 S = cur.S;
      const bool valid = S.x < kInfF;
      const int w = __float_as_int(S.w);
      const int nb = valid ? (w >> 24) - 1 : -1;  // 0 .. kBins-1 for valid lanes
      double *al = acc + (valid ? (w & 0xffffff) : 0);
      const unsigned long long vm = __ballot(valid);
      const int blo = __builtin_amdgcn_readfirstlane(nb);
      const int bhi = __builtin_amdgcn_readlane(nb, 63 - __clzll((long long)vm));
      const int *sb = info + (g0 + cur.s) * kSInfo + 3;
      const float4 *rr = rb + info[(g0 + cur.s) * kSInfo + kSBase];
      const int e0 = sb[blo], e1 = sb[bhi + 1];
      double racc = 0.0;
      int crow = e0 < e1 ? __builtin_amdgcn_readfirstlane(__float_as_int(rr[e0].w) >> 9) : 0;
      auto flush = [&]() {
        if (racc != 0.0)
          __hip_atomic_fetch_add(al + (size_t)crow * I, racc, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
        racc = 0.0;
      };
      auto term = [&](const float4 E) {
        const int ew = __float_as_int(E.w);
        const int row = __builtin_amdgcn_readfirstlane(ew >> 9);
        if (row != crow) {
          flush();
          crow = row;
        }
        const float q = fmaf(S.y, E.y, -(S.z * E.z));
        if ((S.x < E.x) & ((ew & 511) == nb)) racc += (double)q;
      };
      int k = e0;
      for (; k + 4 <= e1; k += 4) {
        const float4 E0 = rr[k], E1 = rr[k + 1], E2 = rr[k + 2], E3 = rr[k + 3];
        term(E0);
        term(E1);
        term(E2);
        term(E3);
      }
      for (; k < e1; ++k) term(rr[k]);
      flush();
      cur = nxt;
    }
    __syncthreads();  // the round's records are consumed
  }
  for (int e = tid; e < jn * I; e += nth) {
    const int u = e / I, i = e - u * I;
    partial[((size_t)sl * J + j0 + u) * I + i] = (float)acc[e];
  }
}

// ------------------------------------------------------------------ pre-spike errors
// W^T with rows padded to ldt (a multiple of 16) by zeros: WT[i * ldt + j] = W[j * I + i].
__global__ __launch_bounds__(256) void transpose_pad_kernel(int J, int I, int ldt,
                                                            const float *__restrict__ W,
                                                            float *__restrict__ WT) {
  __shared__ float tile[32][33];
  const int j0 = blockIdx.y * 32, i0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int j = j0 + r, i = i0 + tx;
    tile[r][tx] = (j < J && i < I) ? W[(size_t)j * I + i] : 0.f;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + r, j = j0 + tx;
    if (i < I && j < ldt) WT[(size_t)i * ldt + j] = tile[tx][r];
  }
}

// Lane per pre spike of the sample's bin-major list, no LDS. The bins after the spike's come
// from Q (rows b*2*kBins + r and b*2*kBins + kBins + r, column i, summed over the GEMM's K
// slices); the pairs inside the spike's bin from the tile records: a wave covers the bins
// [blo, bhi] of its 64 spikes (a narrow range in the bin-major list) and per tile reads the
// record entries of those bins,
// their offsets and the entries themselves through scalar loads (the same for every lane), and
// its lane's 16 weights from W^T (four 16-byte loads, one tile ahead). An entry adds
// w[row] (es' d1 - e' d2) where its bin is the lane's and t_pre < t_post. The grid is
// (chunk slots x G tile groups, B); slot s takes chunks s, s + slots, ... of its sample's
// list and walks the tiles of its group only. A chunk's spikes belong to neurons spread over
// the layer, so each lane reads its own line of W^T per tile: with G = 8 and the group taken
// from blockIdx.x % 8 (workgroups are dealt to the 8 XCDs round robin), every XCD reads one
// eighth of W^T and of the records, which stays in its L2. With G > 1 each group writes its
// partial sums (float64; group 0 adds the bins after the spike's) and pre_err_groups_kernel
// adds them in group order. Slots without chunks (a sample with few pre spikes) exit at once;
// the slot index turns with the sample so that they spread over the XCD's CUs (with 8 slots
// and no turn they fell on the same CUs for every sample: 125 against 102 us per launch on
// the deep network's 800 -> 800 layers; 7 slots, odd, spread them by themselves: 100 us).
constexpr int kPbWaves = 4;
constexpr int kPbSlots = 7;

__global__ __launch_bounds__(64 * kPbWaves) void pre_err_bin_kernel(
    int I, int K_pre, int K, int gj, int cap, const float4 *__restrict__ spk,
    const int *__restrict__ slot, const int *__restrict__ poff, const float *__restrict__ WT,
    int ldt, const float4 *__restrict__ recb, const int *__restrict__ boff,
    const float *__restrict__ Q, int ldq, int q_splits, int64_t q_stride, double inv_tau_s,
    double inv_tau, float *__restrict__ pre_errors, int G, double *__restrict__ gpart) {
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.y;
  // tile group g of G (blocks of one group run on one XCD when G divides 8)
  // the slot turns with the sample, so that the slots left without chunks (short lists) do
  // not fall on the same CUs of the XCD for every sample
  const int g = blockIdx.x % G, nslot = gridDim.x / G;
  const int sl = (blockIdx.x / G + blockIdx.y) % nslot;
  const int tj0 = g * gj / G, tj1 = (g + 1) * gj / G;
  const int P = poff[(size_t)b * (kPreBinsN + 1) + kPreBinsN];
  const int R = tile_rec_slots(K);
  uniform_f32 *srec = (uniform_f32 *)(recb + (size_t)b * gj * R);
  uniform_i32 *sbo = (uniform_i32 *)(boff + (size_t)b * gj * (kBins + 1));
  for (int c = sl * kPbWaves + wv; c * 64 < P; c += nslot * kPbWaves) {
    const int sidx = c * 64 + lane;
    const bool valid = sidx < P;
    const float4 S = valid ? spk[(size_t)b * cap + sidx] : make_float4(kInfF, 0.f, 0.f, 0.f);
    const int sw = __float_as_int(S.w);
    const int i = valid ? (sw & 0xffffff) : 0;
    const int nb = valid ? (sw >> 24) - 1 : kBins;  // -1 .. kBins
    // the bins after the spike's (by tile group 0)
    double q1 = 0.0, q2 = 0.0;
    for (int sp = 0; sp < (g == 0 ? q_splits : 0); ++sp) {
      const float *q = Q + sp * q_stride + (size_t)b * 2 * kBins * ldq + i;
#pragma unroll
      for (int r = 0; r < kBins; ++r) {
        const float a1 = q[(size_t)r * ldq], a2 = q[(size_t)(kBins + r) * ldq];
        if (r > nb) {
          q1 += (double)a1;
          q2 += (double)a2;
        }
      }
    }
    double pe = (double)S.y * inv_tau_s * q1 + (double)S.z * inv_tau * q2;
    // the pairs inside the spike's bin
    const bool walks = valid & (nb >= 0) & (nb < kBins);
    const unsigned long long wm = __ballot(walks);
    if (wm != 0) {
      const int blo = __builtin_amdgcn_readlane(nb, __ffsll((long long)wm) - 1);
      const int bhi = __builtin_amdgcn_readlane(nb, 63 - __clzll((long long)wm));
      const float esd = (float)((double)S.y * inv_tau_s), ed = (float)((double)S.z * inv_tau);
      const float4 *wt = reinterpret_cast<const float4 *>(WT + (size_t)i * ldt);
      float w[kBT], wn[kBT];
#pragma unroll
      for (int h = 0; h < kBT / 4; ++h) {
        const float4 v = wt[tj0 * (kBT / 4) + h];
        w[4 * h] = v.x;
        w[4 * h + 1] = v.y;
        w[4 * h + 2] = v.z;
        w[4 * h + 3] = v.w;
      }
      for (int tj = tj0; tj < tj1; ++tj) {
        const int tn = min(tj + 1, tj1 - 1);
#pragma unroll
        for (int h = 0; h < kBT / 4; ++h) {
          const float4 v = wt[tn * (kBT / 4) + h];
          wn[4 * h] = v.x;
          wn[4 * h + 1] = v.y;
          wn[4 * h + 2] = v.z;
          wn[4 * h + 3] = v.w;
        }
        const int e0 = sbo[tj * (kBins + 1) + blo], e1 = sbo[tj * (kBins + 1) + bhi + 1];
        uniform_f32 *tr = srec + (size_t)tj * R * 4;
        float pt = 0.f;
        auto term = [&](const float4 E) {
          const int ew = __float_as_int(E.w);
          const float wr = w[ew >> 9];
          const float qv = fmaf(esd, E.y, -(ed * E.z));
          const bool m = (S.x < E.x) & ((ew & 511) == nb);
          pt = fmaf(m ? wr : 0.f, qv, pt);
        };
        int k = e0;
        for (; k + 4 <= e1; k += 4) {
          const float4 E0 = uniform_load4(tr, kRecHead + k);
          const float4 E1 = uniform_load4(tr, kRecHead + k + 1);
          const float4 E2 = uniform_load4(tr, kRecHead + k + 2);
          const float4 E3 = uniform_load4(tr, kRecHead + k + 3);
          term(E0);
          term(E1);
          term(E2);
          term(E3);
        }
        for (; k < e1; ++k) term(uniform_load4(tr, kRecHead + k));
        pe += (double)pt;
#pragma unroll
        for (int u = 0; u < kBT; ++u) w[u] = wn[u];
      }
    }
    if (valid) {
      if (G == 1)
        pre_errors[(size_t)b * I * K_pre + slot[(size_t)b * cap + sidx]] = (float)pe;
      else
        gpart[((size_t)g * gridDim.y + b) * cap + sidx] = pe;
    }
  }
}

// pre_err = the tile groups' partial sums, added in group order.
__global__ __launch_bounds__(256) void pre_err_groups_kernel(
    int I, int K_pre, int cap, int G, const int *__restrict__ slot, const int *__restrict__ poff,
    const double *__restrict__ gpart, float *__restrict__ pre_errors) {
  const int b = blockIdx.y, B = gridDim.y;
  const int P = poff[(size_t)b * (kPreBinsN + 1) + kPreBinsN];
  for (int sidx = blockIdx.x * 256 + threadIdx.x; sidx < P; sidx += gridDim.x * 256) {
    double pe = 0.0;
    for (int g = 0; g < G; ++g) pe += gpart[((size_t)g * B + b) * cap + sidx];
    pre_errors[(size_t)b * I * K_pre + slot[(size_t)b * cap + sidx]] = (float)pe;
  }
}

// ------------------------------------------------------------------ host side
// Tile groups of the pre-spike error walk: one per XCD, so that each XCD's L2 holds its
// group's columns of W^T and records (the deep network's 800 -> 800 layers: 227 -> 144 us per
// launch plus 12 us for the group sums; 4 groups 204 us, 16 groups 155 us, 32 groups 175 us).
constexpr int kPeGroups = 8;
static int pre_err_groups(int gj) { return std::max(1, std::min(kPeGroups, gj)); }

BinLayout bin_layout(int B, int I, int K_pre, int J, int K) {
  BinLayout L{};
  L.lda = (J + 3) & ~3;
  L.ldb = (I + 3) & ~3;
  L.ldt = (J + kBT - 1) / kBT * kBT;
  L.gj = (J + kBT - 1) / kBT;
  const size_t rows = (size_t)2 * kBins * B;
  const size_t cap = (size_t)I * K_pre;
  // GEMMs: K split so that the grid is at most one round of 2 workgroups per CU (the LDS of
  // two fits a CU), each split at least 64 deep. The weight gradient has K = 2 kBins B, Q has
  // K = J; Q's slices are summed per pre spike by pre_err_bin_kernel, so it takes at most
  // two (deep network's 800 -> 800 layers: the Q GEMM takes 47 us with two or four slices, 82
  // with one; the walk 143 us reading four slices, 125 us reading two).
  const int tiles = ((J + kGemmBM - 1) / kGemmBM) * ((I + kGemmBN - 1) / kGemmBN);
  L.gemm_splits = std::max(1, std::min((int)(rows / 64), kGemmSlots / tiles));
  const int qtiles = (int)((rows + kGemmBM - 1) / kGemmBM) * ((I + kGemmBN - 1) / kGemmBN);
  L.q_splits = std::max(1, std::min(std::min(J / 64, 2), kGemmSlots / qtiles));
  // in-bin walk: slices so that the grid fills the 256 CUs about once (one workgroup each, as
  // the LDS accumulators take most of a CU's LDS), with at most kWalkMaxSamples per slice
  int sl = std::max(1, std::min(std::max(B, 1), 256 / L.gj));
  sl = std::max(sl, (B + kWalkMaxSamples - 1) / kWalkMaxSamples);
  L.walk_b_per_slice = B > 0 ? (B + sl - 1) / sl : 1;
  L.walk_slices = B > 0 ? (B + L.walk_b_per_slice - 1) / L.walk_b_per_slice : 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = align_up(off + bytes, 256);
    return at;
  };
  L.range = take(sizeof(float4) * (size_t)B);
  L.A = take(sizeof(float) * rows * L.lda);
  L.Bm = take(sizeof(float) * rows * L.ldb);
  L.Q = take(sizeof(float) * rows * L.ldb * L.q_splits);
  L.wt = take(sizeof(float) * (size_t)I * L.ldt);
  L.recb = take(sizeof(float4) * (size_t)B * L.gj * tile_rec_slots(K));
  L.boff = take(sizeof(int) * (size_t)B * L.gj * (kBins + 1));
  L.spk = take(sizeof(float4) * (size_t)B * cap);
  L.slot = take(sizeof(int) * (size_t)B * cap);
  L.poff = take(sizeof(int) * (size_t)B * (kPreBinsN + 1));
  L.D = take(sizeof(float2) * (size_t)B * J * K);
  L.n_eff = take(sizeof(int) * (size_t)B * J);
  L.partial = take(sizeof(float) * (size_t)(L.gemm_splits + L.walk_slices) * J * I);
  L.pe_groups = pre_err_groups(L.gj);
  L.gpart = take(L.pe_groups > 1 ? sizeof(double) * (size_t)L.pe_groups * B * cap : 0);
  L.total = off;
  return L;
}

bool bin_path_applies(int I, int K_pre, int J, int K, int grad_mode) {
  return K_pre > 1 && J > 16 && grad_mode != 1 && I < (1 << 24) && kBT * K <= 1024 &&
         (int64_t)I * J < ((int64_t)1 << 29) && bin_walk_group(I, K) >= 1;
}

static int allow_lds_bins(const void *kernel, size_t bytes) {
  if (bytes <= 64 * 1024) return 0;
  if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) !=
      hipSuccess)
    return fail("binned backward: LDS request above the device limit");
  return 0;
}

int bins_backward(hipStream_t st, const BinLayout &L, int B, int I, int K_pre, int J, int K,
                  const float *t, const float *pre_times, const float *pre_exp_tau_s,
                  const float *pre_exp_tau, const float *weights, double inv_tau_s,
                  double inv_tau, float *pre_errors, char *ws) {
  float4 *range = reinterpret_cast<float4 *>(ws + L.range);
  float *A = reinterpret_cast<float *>(ws + L.A);
  float *Bm = reinterpret_cast<float *>(ws + L.Bm);
  float4 *recb = reinterpret_cast<float4 *>(ws + L.recb);
  int *boff = reinterpret_cast<int *>(ws + L.boff);
  float4 *spk = reinterpret_cast<float4 *>(ws + L.spk);
  int *slot = reinterpret_cast<int *>(ws + L.slot);
  int *poff = reinterpret_cast<int *>(ws + L.poff);
  const float2 *D = reinterpret_cast<const float2 *>(ws + L.D);
  const int *n_eff = reinterpret_cast<const int *>(ws + L.n_eff);
  float *partial = reinterpret_cast<float *>(ws + L.partial);
  const int cap = I * K_pre;
  const int rows = 2 * kBins * B;

  BATS_TIMED("bin_range_kernel", st,
             hipLaunchKernelGGL(bin_range_kernel, dim3(B), dim3(256), 0, st, J, K, n_eff, t,
                                range));
  if (int r = check_launch("bin_range_kernel")) return r;
  const int64_t lanes = (int64_t)B * L.gj * kBT;
  BATS_TIMED("post_bins_kernel", st,
             hipLaunchKernelGGL(post_bins_kernel, dim3((unsigned)((lanes + 255) / 256)),
                                dim3(256), 0, st, B, J, K, L.gj, n_eff, t, D, range, L.lda, A,
                                recb, boff));
  if (int r = check_launch("post_bins_kernel")) return r;
  const int pth = std::min(kPreMaxThreads, (I + 63) / 64 * 64);
  BATS_TIMED("pre_bins_kernel", st,
             hipLaunchKernelGGL(pre_bins_kernel, dim3(B), dim3(pth), 0, st, I, K_pre,
                                pre_times, pre_exp_tau_s, pre_exp_tau, range, L.ldb, Bm, cap,
                                spk, slot, poff));
  if (int r = check_launch("pre_bins_kernel")) return r;
  // weight gradient: GEMM slices, then the in-bin slices after them
  int grc = 0;
  BATS_TIMED("gemm_wgrad", st,
             grc = launch_gemm_f32<true>(st, J, I, rows, A, L.lda, Bm, L.ldb, partial, I,
                                         (int64_t)J * I, L.gemm_splits, 1.f));
  if (grc) return grc;
  const size_t lds = bin_wgrad_lds_bytes(I, K);
  if (int r = allow_lds_bins((const void *)wgrad_bin_kernel, lds)) return r;
  BATS_TIMED("wgrad_bin_kernel", st,
             hipLaunchKernelGGL(wgrad_bin_kernel, dim3(L.gj, L.walk_slices), dim3(kWalkThreads),
                                lds, st, B, I, J, K, L.gj, L.walk_b_per_slice,
                                bin_walk_entries(I), cap, spk, poff, recb, boff,
                                partial + (size_t)L.gemm_splits * J * I));
  if (int r = check_launch("wgrad_bin_kernel")) return r;
  if (!pre_errors) return 0;
  // pre-spike errors: W^T, Q = [S1; -S2] W, then the per-spike suffix sums and in-bin pairs
  float *WT = reinterpret_cast<float *>(ws + L.wt);
  float *Q = reinterpret_cast<float *>(ws + L.Q);
  BATS_TIMED("transpose_pad_kernel", st,
             hipLaunchKernelGGL(transpose_pad_kernel, dim3((I + 31) / 32, (L.ldt + 31) / 32),
                                dim3(256), 0, st, J, I, L.ldt, weights, WT));
  if (int r = check_launch("transpose_pad_kernel")) return r;
  BATS_TIMED("gemm_pre_err", st,
             grc = launch_gemm_f32<false>(st, rows, I, J, A, L.lda, weights, I, Q, L.ldb,
                                          (int64_t)rows * L.ldb, L.q_splits, 1.f));
  if (grc) return grc;
  const int slots = std::max(1, std::min(kPbSlots, (cap + 64 * kPbWaves - 1) / (64 * kPbWaves)));
  const int G = L.pe_groups;
  double *gpart = reinterpret_cast<double *>(ws + L.gpart);
  BATS_TIMED("pre_err_bin_kernel", st,
             hipLaunchKernelGGL(pre_err_bin_kernel, dim3(slots * G, B), dim3(64 * kPbWaves), 0,
                                st, I, K_pre, K, L.gj, cap, spk, slot, poff, WT, L.ldt, recb,
                                boff, Q, L.ldb, L.q_splits, (int64_t)rows * L.ldb, inv_tau_s,
                                inv_tau, pre_errors, G, gpart));
  if (int r = check_launch("pre_err_bin_kernel")) return r;
  if (G == 1) return 0;
  BATS_TIMED("pre_err_groups_kernel", st,
             hipLaunchKernelGGL(pre_err_groups_kernel, dim3(8, B), dim3(256), 0, st, I, K_pre,
                                cap, G, slot, poff, gpart, pre_errors));
  return check_launch("pre_err_groups_kernel");
}

}  // namespace bats
