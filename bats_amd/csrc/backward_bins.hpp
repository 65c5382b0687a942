// Binned backward of dense layers whose pre layer fires several times per neuron: the weight
// gradient and the pre-spike errors as float32 GEMMs on the matrix cores over per-sample time
// bins, plus an exact walk of the spike pairs that share a bin (lif_backward_bins.hip).
#pragma once

#include "common.hpp"

namespace bats {

// Time bins per sample. The GEMMs grow with it (4 J I B kBins multiply-adds for each of the
// weight gradient and the pre-spike errors) and the exact walk's pairs shrink with it (about
// 1/kBins of the spike pairs), but the walk's cost does not follow its pairs: a pre spike still
// reads the weights of every tile with a post spike in its bin. On the deep network's two
// 800 -> 800 layers (batch 256) the whole step measured 2.55, 2.39, 2.33, 2.32, 2.33, 2.38,
// 2.55 and 2.66 ms at 2, 3, 4, 5, 6, 8, 12 and 16 bins.
constexpr int kBins = 4;

// Tile records of the dense backward: per (sample, tile of kWT post rows) tile_rec_slots(K)
// float4, slot 0 holding the tile's spike count (int bits in x), the spikes from slot kRecHead.
constexpr int kWT = 16;      // post rows per tile
constexpr int kRecHead = 1;  // first spike slot of a tile record
__host__ __device__ inline int tile_rec_slots(int K) { return kRecHead + kWT * K; }

struct BinLayout {
  size_t range, A, Bm, Q, wt, recb, boff, spk, slot, poff, D, n_eff, partial, gpart, total;
  int lda, ldb, ldt, gj, gemm_splits, q_splits, walk_slices, walk_b_per_slice, pe_groups;
};

// Workspace layout of the binned path for a layer (B samples, I pre neurons with up to K_pre
// spikes, J post neurons with up to K spikes).
BinLayout bin_layout(int B, int I, int K_pre, int J, int K);

// Whether the binned path serves this call: pre neurons with several spikes, more than the
// narrow-layer limit of post neurons, a batch-reduced gradient, and the walk's LDS fits.
bool bin_path_applies(int I, int K_pre, int J, int K, int grad_mode);

// After row_prep_kernel has run the reset chain (errors in place) and written the per-spike
// terms D = {d1, d2} [B*J*K] and the leading-finite counts n_eff [B*J] into the workspace
// regions of L: the weight gradient as L.gemm_splits + L.walk_slices partial slices [J][I]
// at ws + L.partial (to be summed in order and scaled by the caller), and with pre_errors
// (zero-filled by the caller) the pre-spike errors.
int bins_backward(hipStream_t st, const BinLayout &L, int B, int I, int K_pre, int J, int K,
                  const float *t, const float *pre_times, const float *pre_exp_tau_s,
                  const float *pre_exp_tau, const float *weights, double inv_tau_s,
                  double inv_tau, float *pre_errors, char *ws);

}  // namespace bats
