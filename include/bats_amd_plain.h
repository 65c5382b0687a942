/*
 * bats_amd_plain.h -- the plain FP32 straight-translation form (measured baseline only).
 *
 * These entry points are not part of the optimised path. Each launches one kernel with the
 * reference's launch geometry (one thread per sample, block = batch <= 1024) and
 * materialisations, computing in float32, so that the optimised path can be timed against
 * the reference's structure on the same GPU. Outputs that the reference zero- or
 * inf-initialises are initialised by the caller (bats_amd/plain.py).
 */
#ifndef BATS_AMD_PLAIN_H
#define BATS_AMD_PLAIN_H

#ifdef __cplusplus
extern "C" {
#endif

/* bats/CudaKernels/Wrappers/Inference/compute_pre_exps.py:10-18 */
int bats_plain_pre_exps(void *stream, int B, int N, int K, const float *t, float *es, float *e,
                        float tau_s, float tau);
/* bats/CudaKernels/Wrappers/Inference/get_spike_weights.py:9-18; out [B,J,P] */
int bats_plain_spike_weights(void *stream, int B, int P, int J, int I, const float *W,
                             const int *idx, float *out);
/* bats/CudaKernels/Wrappers/Inference/compute_spike_times.py:11-31 */
int bats_plain_solve(void *stream, int B, int P, int J, int K, const float *st, const float *ses,
                     const float *se, const float *sw, float c, float dtt, float tau, float tmax,
                     int *n, float *a, float *x, float *t, float *z);
/* bats/CudaKernels/Wrappers/Backpropagation/propagate_recurrent_errors.py:9-15 */
int bats_plain_recurrent(void *stream, int B, int J, int K, const float *x, const float *z,
                         float *errors, float dtt);
/* bats/CudaKernels/Wrappers/Backpropagation/compute_factors.py:11-22 */
int bats_plain_factors(void *stream, int B, int J, int K, const float *t, const float *a, float c,
                       const float *x, const float *z, float tau, float *f1, float *f2);
/* bats/CudaKernels/Wrappers/Backpropagation/compute_weights_gradient.py:10-24; grad [B,J,I] */
int bats_plain_wgrad(void *stream, int B, int J, int I, int K, int K_pre, const float *f1,
                     const float *f2, const float *post_t, const float *pre_t,
                     const float *pre_es, const float *pre_e, const float *errors, float *grad);
/* bats/CudaKernels/Wrappers/Backpropagation/propagate_errors_to_pre_spikes.py:11-27; out [B,I,K_pre,J] */
int bats_plain_pre_errors(void *stream, int B, int I, int K_pre, int J, int K, const float *f1,
                          const float *f2, const float *post_t, const float *pre_t,
                          const float *pre_es, const float *pre_e, const float *W,
                          const float *errors, float *out, float tau_s, float tau);

#ifdef __cplusplus
}
#endif

#endif /* BATS_AMD_PLAIN_H */
