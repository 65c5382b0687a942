/*
 * bats_amd.h -- C ABI of the MI355X (gfx950) BATS hot path.
 *
 * One shared library, libbats_amd.so, built with hipcc for gfx950. Every entry point
 * takes device pointers, sizes and scalars plus a HIP stream (passed as an opaque
 * pointer; NULL is the default stream) and returns 0 on success or a nonzero status,
 * with the message available from bats_last_error(). Outputs and workspaces are
 * allocated by the caller; the library allocates no device memory, never synchronises
 * the stream and issues no memset, so every call can be captured in a HIP graph. Every
 * choice that changes which kernels run is a per-call argument; the only process-wide
 * switch is the opt-in kernel timing at the end of this file, which changes no result.
 *
 * Layouts (row-major, float32 unless stated):
 *   spike times  [B, N, K]  ascending per neuron, +inf padded ("no spike")
 *   spike counts [B, N]     int32
 *   dense weights [J, I]    post-major, as the reference's LIFLayer weights
 *
 * Each declaration names the reference interface it replaces (paths relative to the
 * reference repository root, Florian-BACHO/bats).
 */
#ifndef BATS_AMD_H
#define BATS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failing call on this host thread ("" if none). */
const char *bats_last_error(void);

/* Library version string. */
const char *bats_version(void);

/* ------------------------------------------------------------------ dense LIF forward */

/* Workspace bytes bats_lif_forward needs for these sizes. */
size_t bats_lif_forward_workspace_size(int B, int I, int K_pre, int J);

/*
 * Dense LIF layer forward: pre-spike exps, per-sample time sort, weight gather and the
 * exact multi-spike spike-time solver, fused.
 * Replaces, in one call, bats/Layers/LIFLayer.py:63-91 and the operators it launches:
 *   compute_pre_exps        bats/CudaKernels/Wrappers/Inference/compute_pre_exps.py:10
 *   get_sorted_spikes_indices  bats/CudaKernels/Wrappers/Inference/get_sorted_spikes_indices.py:7
 *   get_spike_weights       bats/CudaKernels/Wrappers/Inference/get_spike_weights.py:9
 *   compute_spike_times     bats/CudaKernels/Wrappers/Inference/compute_spike_times.py:11
 * Inputs : pre_times [B,I,K_pre]; weights [J,I]; tau_s, tau = 2 tau_s, c = theta/tau,
 *          delta_theta_tau = delta_theta/tau (float32 as the layer stores them);
 *          max_simulation (+inf allowed).
 * Outputs: out_n [B,J] int32; out_t [B,J,K_post] (+inf padded); out_a, out_x, out_z
 *          [B,J,K_post] (valid for slots < out_n, other slots untouched);
 *          pre_exp_tau_s, pre_exp_tau [B,I,K_pre] (valid for finite pre slots only).
 */
int bats_lif_forward(void *stream, int B, int I, int K_pre, int J, int K_post,
                     const float *pre_times, const float *weights,
                     float tau_s, float tau, float c, float delta_theta_tau,
                     float max_simulation,
                     int *out_n, float *out_t, float *out_a, float *out_x, float *out_z,
                     float *pre_exp_tau_s, float *pre_exp_tau,
                     void *workspace, size_t workspace_bytes);

/*
 * bats_lif_forward with the solver form chosen by the caller for this call only (the library
 * keeps no tuning state): solver_mode 0 = automatic (what bats_lif_forward does: one lane per
 * (sample, post neuron) when there are at least 32768 such pairs, else one wave per pair),
 * 1 = always one lane per pair, 2 = always one wave per pair, 3 / 4 = one lane per pair with
 * the event records staged in LDS / read with scalar loads (1 chooses by the launch's waves
 * per SIMD), 5 = for K_post = 1 (one spike per neuron; otherwise as 1) the event walk split
 * into time segments walked at once (0 takes it when the lane-per-pair form would have fewer
 * than two waves per SIMD; solver_blocks > 0 then sets the number of segments, at most 16);
 * solver_blocks = workgroups per sample of the lane-per-pair solver (0: one per
 * 1024 neurons, more when the batch has fewer samples than the device has CUs). Every form
 * gives the same results; the arguments exist for testing and tuning. Same reference
 * interfaces as bats_lif_forward (bats/Layers/LIFLayer.py:63-91).
 */
int bats_lif_forward_solver(void *stream, int B, int I, int K_pre, int J, int K_post,
                            const float *pre_times, const float *weights,
                            float tau_s, float tau, float c, float delta_theta_tau,
                            float max_simulation,
                            int *out_n, float *out_t, float *out_a, float *out_x, float *out_z,
                            float *pre_exp_tau_s, float *pre_exp_tau,
                            int solver_mode, int solver_blocks,
                            void *workspace, size_t workspace_bytes);

/*
 * Two dense layers in a row (the second's inputs are the first's spike trains): the second
 * layer's pre-spike exps and per-sample time-ordered event lists are made by the first
 * layer's call, from its own spikes, and the second layer's call solves on them. This replaces
 * the second layer's part of bats/Layers/LIFLayer.py:63-91 that only depends on its inputs
 * (compute_pre_exps, bats/CudaKernels/Wrappers/Inference/compute_pre_exps.py:10, and
 * get_sorted_spikes_indices, bats/CudaKernels/Wrappers/Inference/get_sorted_spikes_indices.py:7)
 * by the first layer's solver: when one workgroup holds all of a sample's neurons, the solver
 * sorts the sample's spikes in LDS after its walk (no per-sample sort of a [B, J, K_post]
 * buffer). Results are those of two bats_lif_forward calls, bit for bit.
 *
 * bats_event_lists_size(B, S): bytes of the event lists of B samples with S input slots each
 * (for the second layer, S = J * K_post of the first).
 *
 * bats_lif_forward_chain: bats_lif_forward of the first layer (same arguments and outputs up
 * to pre_exp_tau), plus the second layer's constants next_tau_s, next_tau, next_c and its
 * next_max_simulation, its per-slot exps next_pre_exp_tau_s / next_pre_exp_tau [B, J, K_post]
 * (valid at the first layer's spikes, as bats_lif_forward writes pre exps) and its event lists
 * (next_events, bats_event_lists_size(B, J * K_post) bytes, device memory the caller keeps
 * until the second layer's call has run).
 *
 * bats_lif_forward_events: the second layer's forward on those event lists (events /
 * events_bytes; I, K_pre = the first layer's J, K_post), with a workspace of
 * bats_lif_forward_events_workspace_size(I, J) bytes; outputs out_n .. out_z as
 * bats_lif_forward (its pre exps are the ones the chained call wrote). With next_events not
 * NULL it also makes the event lists of the layer after it, as bats_lif_forward_chain does
 * (next_events NULL: the next_* arguments are ignored).
 */
size_t bats_event_lists_size(int B, int S);
int bats_lif_forward_chain(void *stream, int B, int I, int K_pre, int J, int K_post,
                           const float *pre_times, const float *weights,
                           float tau_s, float tau, float c, float delta_theta_tau,
                           float max_simulation,
                           int *out_n, float *out_t, float *out_a, float *out_x, float *out_z,
                           float *pre_exp_tau_s, float *pre_exp_tau,
                           float next_tau_s, float next_tau, float next_c,
                           float next_max_simulation,
                           float *next_pre_exp_tau_s, float *next_pre_exp_tau,
                           void *next_events, size_t next_events_bytes,
                           void *workspace, size_t workspace_bytes);
size_t bats_lif_forward_events_workspace_size(int I, int J);
int bats_lif_forward_events(void *stream, int B, int I, int K_pre, int J, int K_post,
                            const void *events, size_t events_bytes, const float *weights,
                            float tau_s, float tau, float c, float delta_theta_tau,
                            int *out_n, float *out_t, float *out_a, float *out_x, float *out_z,
                            float next_tau_s, float next_tau, float next_c,
                            float next_max_simulation,
                            float *next_pre_exp_tau_s, float *next_pre_exp_tau,
                            void *next_events, size_t next_events_bytes,
                            void *workspace, size_t workspace_bytes);

/*
 * bats_lif_forward for a first layer fed by packed event lists, for inference: sample b's
 * input events are indices/times[row_offsets[b] .. row_offsets[b+1]) (row_offsets [B+1] int32
 * on the device), in ascending time, at most P_max per sample (later ones are dropped); an
 * index outside [0, I) or a time that is not finite ends a sample's list. The events become the
 * solver's time-ordered event lists directly, without the per-neuron buffers and the per-sample
 * sort of bats_lif_forward (events of equal time keep their list order; bats_lif_forward
 * orders them by input neuron, which can change float64 sums in their last bits). No per-slot
 * exps are written: there is no backward from this entry point. Outputs as bats_lif_forward.
 * Same reference interfaces (bats/Layers/LIFLayer.py:63-91, and the event-list conversion of
 * bats/CudaKernels/Wrappers/Inference/convert_to_spike_per_neuron.py:20-33 that it makes
 * unnecessary).
 */
size_t bats_lif_forward_packed_workspace_size(int B, int I, int P_max, int J);
int bats_lif_forward_packed(void *stream, int B, int I, int J, int K_post, int P_max,
                            const int *row_offsets, const int *indices, const float *times,
                            const float *weights, float tau_s, float tau, float c,
                            float delta_theta_tau, float max_simulation,
                            int *out_n, float *out_t, float *out_a, float *out_x, float *out_z,
                            void *workspace, size_t workspace_bytes);

/*
 * bats_lif_forward_packed chained to the next dense layer, as bats_lif_forward_chain is to
 * bats_lif_forward: the next layer's pre exps and event lists are made in the same call (for
 * bats_lif_forward_events). Same reference interfaces as bats_lif_forward_chain.
 */
int bats_lif_forward_packed_chain(void *stream, int B, int I, int J, int K_post, int P_max,
                                  const int *row_offsets, const int *indices, const float *times,
                                  const float *weights, float tau_s, float tau, float c,
                                  float delta_theta_tau, float max_simulation,
                                  int *out_n, float *out_t, float *out_a, float *out_x,
                                  float *out_z,
                                  float next_tau_s, float next_tau, float next_c,
                                  float next_max_simulation,
                                  float *next_pre_exp_tau_s, float *next_pre_exp_tau,
                                  void *next_events, size_t next_events_bytes,
                                  void *workspace, size_t workspace_bytes);

/*
 * bats_lif_forward without a workspace argument and without any allocation: the workspace is
 * a file-scope device array of bats_lif_forward_noalloc_max_bytes() bytes (64 MiB) in the
 * library, one per GPU, held until the process ends. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Sizes whose workspace
 * (bats_lif_forward_workspace_size) exceeds the array are refused with an error before any
 * launch. Calls on one GPU share the array, so calls that overlap on one GPU (several streams
 * or host threads) race; calls in stream order do not. Library timing does not apply to it.
 * Same arguments, outputs and results as bats_lif_forward otherwise.
 */
size_t bats_lif_forward_noalloc_max_bytes(void);
int bats_lif_forward_noalloc(void *stream, int B, int I, int K_pre, int J, int K_post,
                             const float *pre_times, const float *weights,
                             float tau_s, float tau, float c, float delta_theta_tau,
                             float max_simulation,
                             int *out_n, float *out_t, float *out_a, float *out_x, float *out_z,
                             float *pre_exp_tau_s, float *pre_exp_tau);

/* ------------------------------------------------------------------ dense LIF backward */

size_t bats_lif_backward_workspace_size(int B, int I, int K_pre, int J, int K_post);

/*
 * Dense LIF layer BATS backward.
 * Replaces bats/Layers/LIFLayer.py:93-110 and its operators:
 *   propagate_recurrent_errors bats/CudaKernels/Wrappers/Backpropagation/propagate_recurrent_errors.py:9
 *   compute_factors            bats/CudaKernels/Wrappers/Backpropagation/compute_factors.py:11
 *   compute_weights_gradient   bats/CudaKernels/Wrappers/Backpropagation/compute_weights_gradient.py:10
 *   propagate_errors_to_pre_spikes bats/CudaKernels/Wrappers/Backpropagation/propagate_errors_to_pre_spikes.py:11
 * and the caller's batch mean (experiments/mnist/train_spike_count.py:167).
 * errors [B,J,K_post] is updated in place with the reset-chain errors, as the reference's
 * propagate_recurrent_errors does (only slots holding a spike change).
 * grad_mode 0: grad is [J,I] = batch mean of the per-sample gradients.
 * grad_mode 1: grad is [B,J,I], the per-sample gradients of the reference.
 * grad_mode 2: grad is [J,I] = batch sum (the gradient of a summed loss, for autograd).
 * pre_errors [B,I,K_pre] may be NULL (previous layer not trainable).
 */
int bats_lif_backward(void *stream, int B, int I, int K_pre, int J, int K_post,
                      const int *n, const float *t, const float *a, const float *x,
                      const float *z, float *errors,
                      const float *pre_times, const float *pre_exp_tau_s,
                      const float *pre_exp_tau, const float *weights,
                      float tau_s, float tau, float c, float delta_theta_tau,
                      float *grad, int grad_mode, float *pre_errors,
                      void *workspace, size_t workspace_bytes);

/* ------------------------------------------------------------------ conv LIF and pooling */

/*
 * Shapes: pre_shape = {X, Y, C_in} of the input layer, filters_shape = {fx, fy, C_out}
 * (the reference's ConvLIFLayer(filters_shape=[fx, fy, C_out])); weights [C_out, fx, fy, C_in];
 * neuron index (x*Y + y)*C + c; valid, stride 1, output {X-fx+1, Y-fy+1, C_out}.
 */
size_t bats_conv_lif_forward_workspace_size(int B, const int *pre_shape, int K_pre,
                                            const int *filters_shape);

/*
 * Conv LIF forward. Replaces bats/Layers/ConvLIFLayer.py:77-107 with compute_pre_exps,
 * get_sorted_spikes_indices and compute_spike_times_conv
 * (bats/CudaKernels/Wrappers/Inference/compute_spike_times_conv.py:11-35). Outputs as
 * bats_lif_forward, with J = (X-fx+1)*(Y-fy+1)*C_out.
 */
int bats_conv_lif_forward(void *stream, int B, const int *pre_shape, int K_pre,
                          const int *filters_shape, int K_post, const float *pre_times,
                          const float *weights, float tau_s, float tau, float c,
                          float delta_theta_tau, float max_simulation, int *out_n,
                          float *out_t, float *out_a, float *out_x, float *out_z,
                          float *pre_exp_tau_s, float *pre_exp_tau, void *workspace,
                          size_t workspace_bytes);

size_t bats_conv_lif_backward_workspace_size(int B, const int *pre_shape, int K_pre,
                                             const int *filters_shape, int K_post);

/*
 * Conv LIF backward. Replaces bats/Layers/ConvLIFLayer.py:109-134 with
 * propagate_recurrent_errors, compute_factors, compute_weights_gradient_conv
 * (bats/CudaKernels/Wrappers/Backpropagation/compute_weights_gradient_conv.py:11-32) and
 * propagate_errors_to_pre_spikes_conv (.../propagate_errors_to_pre_spikes_conv.py:11-32).
 * grad: grad_mode 0 [C_out,fx,fy,C_in] batch mean, 1 [B,C_out,fx,fy,C_in] per sample,
 * 2 batch sum. errors updated in place; pre_errors may be NULL.
 */
int bats_conv_lif_backward(void *stream, int B, const int *pre_shape, int K_pre,
                           const int *filters_shape, int K_post, const int *n, const float *t,
                           const float *a, const float *x, const float *z, float *errors,
                           const float *pre_times, const float *pre_exp_tau_s,
                           const float *pre_exp_tau, const float *weights, float tau_s,
                           float tau, float c, float delta_theta_tau, float *grad,
                           int grad_mode, float *pre_errors, void *workspace,
                           size_t workspace_bytes);

/*
 * 2x2 spike pooling. Replaces bats/Layers/PoolingLayer.py:49-53 with aggregate_spikes_conv
 * (bats/CudaKernels/Wrappers/Inference/aggregate_spikes_conv.py:10-28): out_n [B,N_post],
 * out_t [B,N_post,4K] ascending (+inf last), out_src [B,N_post,4K] int32 flat index of each
 * slot's source in t [B,N_pre,K]. N_post = (X/2)*(Y/2)*C.
 */
int bats_pool2x2_forward(void *stream, int B, const int *pre_shape, int K, const int *n,
                         const float *t, int *out_n, float *out_t, int *out_src);

/* Pooling backward (bats/Layers/PoolingLayer.py:55-66): pre_errors[src] = errors. */
int bats_pool2x2_backward(void *stream, int B, const int *pre_shape, int K, const float *errors,
                          const int *src, float *pre_errors);

/* ------------------------------------------------------------------ losses */

/*
 * Spike-count class loss. Replaces bats/Losses/SpikeCountClassLoss.py:15-33 with
 * bats/Losses/SpikeCountLoss.py:17-38: loss[b] = 0.5*sum_j (target - n)^2,
 * errors[b,j,k] = target - n, pred[b] = argmax_j n. Any output pointer may be NULL.
 */
int bats_count_loss(void *stream, int B, int O, int K, const int *n, const int64_t *labels,
                    float target_true, float target_false,
                    float *loss, float *errors, int64_t *pred);

/*
 * Generic spike-count loss with explicit targets [B,O] (bats/Losses/SpikeCountLoss.py:20-38).
 */
int bats_count_loss_targets(void *stream, int B, int O, int K, const int *n,
                            const float *targets, float *loss, float *errors, int64_t *pred);

/*
 * Softmax cross-entropy on the first spike time. Replaces
 * bats/Losses/TTFSSoftmaxCrossEntropy.py:20-56. t is [B,O,K]; errors [B,O,K] (slot 0 set,
 * others 0); pred has B*K entries (argmin over neurons, as the reference flattens it).
 */
int bats_ttfs_loss(void *stream, int B, int O, int K, const float *t, const int64_t *labels,
                   float tau, float *loss, float *errors, int64_t *pred);

/* ------------------------------------------------------------------ optimizers */

/*
 * Adam deltas. Replaces bats/Optimizers/AdamOptimizer.py:35-55 for one tensor of n
 * elements: m, v updated in place, delta = -lr * m_hat / (sqrt(v_hat) + eps) with
 * bias1 = 1 - beta_1^t and bias2 = 1 - beta_2^t. If weights is not NULL the delta is also
 * added to it (the fused form of bats/Network.py:50-54).
 */
int bats_adam_step(void *stream, int64_t n, const float *grad, float *m, float *v,
                   float *delta, float *weights, float lr, float beta_1,
                   float one_minus_beta_1, float beta_2, float one_minus_beta_2,
                   float epsilon, double bias1, double bias2);

/* NAdam deltas. Replaces bats/Optimizers/NAdamOptimizer.py:36-58. */
int bats_nadam_step(void *stream, int64_t n, const float *grad, float *m, float *v,
                    float *delta, float *weights, float lr, float beta_1,
                    float one_minus_beta_1, float beta_2, float one_minus_beta_2,
                    float epsilon, double bias1, double bias2);

/*
 * Adam (nadam = 0) or NAdam (nadam = 1) with the step counter t in device memory (*step,
 * incremented by the caller before the call): the bias corrections are computed on the
 * device, so a captured graph of the training step stays correct on every replay.
 * Same arithmetic as bats_adam_step / bats_nadam_step otherwise.
 */
int bats_adam_step_dev(void *stream, int64_t n, const float *grad, float *m, float *v,
                       float *delta, float *weights, float lr, float beta_1,
                       float one_minus_beta_1, float beta_2, float one_minus_beta_2,
                       float epsilon, const int64_t *step, int nadam);

/* Gradient-descent deltas. Replaces bats/Optimizers/GradientDescentOptimizer.py:8-9. */
int bats_sgd_step(void *stream, int64_t n, const float *grad, float *delta,
                  float *weights, float lr);

/* weights += delta. Replaces bats/Layers/LIFLayer.py:112-113 (add_deltas). */
int bats_add_deltas(void *stream, int64_t n, float *weights, const float *delta);

/*
 * One optimizer step of every layer in one launch: Adam (nadam = 0) or NAdam (nadam = 1) over
 * count <= 16 tensors, tensor k of sizes[k] elements at grads[k], m[k], v[k], writing
 * deltas[k] and/or adding to weights[k] (either array, or an entry of it, may be NULL where the
 * other is given). The pointer arrays are host memory, read before the call returns; step as
 * in bats_adam_step_dev (NULL: bias1/bias2 are used). Replaces the per-layer loop of
 * bats/Optimizers/AdamOptimizer.py:35-55 (and NAdamOptimizer.py:36-58) with one kernel.
 */
int bats_adam_step_multi(void *stream, int count, const int64_t *sizes,
                         const float *const *grads, float *const *m, float *const *v,
                         float *const *deltas, float *const *weights, float lr, float beta_1,
                         float one_minus_beta_1, float beta_2, float one_minus_beta_2,
                         float epsilon, double bias1, double bias2, const int64_t *step,
                         int nadam);

/*
 * weights[k] += deltas[k] for count <= 16 tensors in one launch (host pointer arrays). Replaces
 * the per-layer loop of bats/Network.py:50-54 (apply_deltas).
 */
int bats_add_deltas_multi(void *stream, int count, const int64_t *sizes, float *const *weights,
                          const float *const *deltas);

/* ------------------------------------------------------------------ input encoding */

/*
 * Latency code of pixel intensities. Replaces the host-side Dataset.__to_spikes of
 * experiments/mnist/Dataset.py:66-71: t = time_window * (1 - px / max_value) (evaluated in
 * float64, stored as float32), +inf and a count of 0 where px = 0.
 * pixels [n] uint8 (pixel_type 0) or float32 (pixel_type 1); times [n] (the [B,N,1] spike
 * buffer of an input layer), counts [n] int32.
 */
int bats_latency_encode(void *stream, int64_t n, const void *pixels, int pixel_type,
                        double time_window, double max_value, float *times, int *counts);

/* ------------------------------------------------------------------ event lists */

/*
 * Time-ordered event list -> per-neuron buffers. Replaces
 * bats/CudaKernels/Wrappers/Inference/convert_to_spike_per_neuron.py:20-33.
 * indices [B,P] int32 (a negative index ends the list), times [B,P].
 * out_times [B,N,K] is fully written (+inf padded); out_counts [B,N] counts every
 * event of the neuron, events beyond K are dropped.
 */
int bats_events_to_per_neuron(void *stream, int B, int P, int N, int K,
                              const int *indices, const float *times,
                              float *out_times, int *out_counts);

/*
 * The same conversion from a packed list: sample b's events are
 * indices/times[row_offsets[b] .. row_offsets[b+1]) (row_offsets [B+1] int32, ascending),
 * in time order; a negative index also ends a sample's list. Replaces the same reference
 * wrapper (bats/CudaKernels/Wrappers/Inference/convert_to_spike_per_neuron.py:20-33) for
 * callers that hold requests of uneven lengths back to back (no [B,P] padding to build).
 * out_row_max [B] (may be NULL) receives each sample's largest per-neuron count, dropped
 * events included, so a caller can tell whether K held every event. 1 <= N <= 4096.
 */
int bats_events_csr_to_per_neuron(void *stream, int B, int N, int K, const int *row_offsets,
                                  const int *indices, const float *times, float *out_times,
                                  int *out_counts, int *out_row_max);

/*
 * Host-side helper of an inference server (no GPU work, no counterpart in the reference, which
 * has no server): gathers n event lists from a host arena into one packed list.
 * Request r's events are src_indices/src_times[src_offsets[r] .. + lengths[r]); they are
 * copied back to back into dst_indices/dst_times, and row_offsets [n+1] receives their
 * offsets (the layout bats_events_csr_to_per_neuron reads). The copies run on a small pool of
 * host threads kept for the life of the process. Fails if the total exceeds dst_capacity
 * events. Calls from several host threads are serialised.
 */
int bats_pack_event_lists(int n, const int64_t *src_offsets, const int32_t *lengths,
                          const int32_t *src_indices, const float *src_times,
                          int32_t *row_offsets, int32_t *dst_indices, float *dst_times,
                          int64_t dst_capacity);

/* ------------------------------------------------------------------ observability */

/*
 * Per-kernel HIP-event timing of the library's main launches (sort_events_kernel,
 * lif_solve_seq_kernel, lif_solve_wave_kernel, wgrad_kernel). No counterpart in the
 * reference (it has no profiler). Off by default; a process-wide observability switch (like
 * a profiler's): enabling it adds two event records per timed launch of every caller, changes
 * no result, and must not be used while a stream is being captured.
 */
void bats_timing_enable(int on);
void bats_timing_reset(void);
/* Waits for the recorded launches and returns the total milliseconds and launch count. */
int bats_timing_query(const char *kernel_name, double *total_ms, long long *count);

#ifdef __cplusplus
}
#endif

#endif /* BATS_AMD_H */
