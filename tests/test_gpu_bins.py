"""GPU parity of the binned backward of multi-spike dense layers against the FP64 oracle.

The binned form serves batch-reduced weight gradients of layers whose pre layer fires several
times per neuron: per-sample time bins, a float32 GEMM on the matrix cores for the pairs in
different bins, an exact walk of the pairs inside one bin, and the pre-spike errors from a
second GEMM plus their in-bin pairs. These cases check it at its edges: ragged sizes (post rows
not a multiple of the 16-row tile, pre neurons not a multiple of 4), samples whose post layer
is silent or fires once (an empty or zero-width time range), pre spikes before the first and
after the last post spike, times tied across pre and post, several GEMM and walk slices.

Tolerance (as the other backward tests): |gpu - ref| <= 1e-4 |ref| + 1e-6 max|ref| elementwise,
for the reset-chain errors, the batch-mean (or batch-sum) gradient and the pre-spike errors.
"""
import numpy as np
import pytest
import torch

from oracle import lif as olif
from tests.helpers import assert_close_rel, multi_spike_inputs

pytestmark = pytest.mark.gpu

F32 = np.float32


def binned_backward(ref, errors, batch_sum=False):
    from bats_amd import ops
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    err = T(errors.astype(F32))
    ops.timing(True, reset=True)
    try:
        grad, pre = ops.lif_backward(T(ref.n), T(ref.t), T(ref.a), T(ref.x), T(ref.z), err,
                                     T(ref.pre_times), T(ref.pre_exp_s), T(ref.pre_exp),
                                     T(ref.W), ref.tau_s, ref.tau, ref.c, ref.dtt,
                                     per_sample=False, pre_trainable=True, batch_sum=batch_sum)
        torch.cuda.synchronize()
        _, walks = ops.timing_query("wgrad_bin_kernel")
        _, pre_walks = ops.timing_query("pre_err_bin_kernel")
    finally:
        ops.timing(False, reset=True)
    assert walks == 1 and pre_walks == 1, "the binned path did not serve this layer"
    return err.cpu().numpy(), grad.cpu().numpy(), pre.cpu().numpy()


def check_binned(ref, errors, label, batch_sum=False):
    print(f"[binned backward] {label}: B={ref.t.shape[0]} I={ref.pre_times.shape[1]} "
          f"K_pre={ref.pre_times.shape[2]} J={ref.t.shape[1]} K={ref.t.shape[2]}", flush=True)
    e_ref, g_ref, p_ref = olif.lif_backward(ref, errors, pre_trainable=True)
    e_g, g_g, p_g = binned_backward(ref, errors, batch_sum)
    g_want = np.asarray(g_ref, np.float64).sum(axis=0) if batch_sum else olif.batch_mean(g_ref)
    assert_close_rel(e_g, e_ref, 1e-4, 1e-6, "reset-chain errors")
    assert_close_rel(g_g, g_want, 1e-4, 1e-6, "batch gradient")
    assert_close_rel(p_g, p_ref, 1e-4, 1e-6, "pre-spike errors")


def test_binned_ragged_shapes_and_silent_samples():
    rng = np.random.default_rng(21)
    B, I, Kp, J, K = 9, 101, 6, 37, 5
    pre, _ = multi_spike_inputs(rng, B, I, Kp, t_max=1.0)
    pre[2] = np.inf  # a sample with no input: its post layer is silent (empty time range)
    W = (2.5 * rng.normal(0, 1.0, size=(J, I))).astype(F32)
    ref = olif.lif_forward(pre, W, 0.130, 1.0, 1.0, K, tmax=1.0)
    assert ref.n[2].sum() == 0 and ref.n.sum() > 50
    check_binned(ref, rng.normal(size=ref.t.shape).astype(F32), "ragged, silent sample")


def test_binned_single_post_spike_and_outer_bins():
    """One sample whose post layer fires exactly once (lo == hi: one bin), pre spikes long
    after the last post spike (bin after the range, no contribution) and before the first
    (bin before the range, every post spike follows them)."""
    rng = np.random.default_rng(5)
    B, I, Kp, J, K = 4, 64, 4, 20, 3
    pre, _ = multi_spike_inputs(rng, B, I, Kp, t_max=0.6)
    W = (2.0 * rng.normal(0, 1.0, size=(J, I))).astype(F32)
    # sample 1: a single strong input early -> one post neuron fires once
    pre[1] = np.inf
    pre[1, 0, 0] = 0.05
    W[3, 0] = 20.0
    ref0 = olif.lif_forward(pre, W, 0.130, 1.0, 1.0, K, tmax=1.0)
    assert ref0.n[1].sum() == 1
    # late pre spikes (after every post spike of their sample) on the spare slots of sample 0
    n0 = np.isfinite(pre[0]).sum(axis=1)
    spare = np.nonzero(n0 < Kp)[0]
    pre[0, spare, n0[spare]] = 0.99
    ref = olif.lif_forward(pre, W, 0.130, 1.0, 1.0, K, tmax=1.0)
    assert ref.n[1].sum() == 1
    check_binned(ref, rng.normal(size=ref.t.shape).astype(F32), "one post spike, outer bins")


def test_binned_tied_times_many_slices_and_batch_sum():
    """Pre and post times on a coarse grid (ties across bins' edges and within bins, counted
    only strictly before), a batch large enough for several GEMM and walk slices, and the
    batch-sum reduction."""
    rng = np.random.default_rng(9)
    B, I, Kp, J, K = 70, 96, 5, 48, 6
    pre, _ = multi_spike_inputs(rng, B, I, Kp, t_max=1.0, p_silent=0.1)
    pre = np.where(np.isfinite(pre), np.round(pre * 32) / 32, np.inf).astype(F32)
    pre = np.sort(pre, axis=2)
    W = (1.5 * rng.normal(0, 1.0, size=(J, I))).astype(F32)
    ref = olif.lif_forward(pre, W, 0.130, 1.0, 1.0, K, tmax=1.0)
    ref.t = np.where(np.isfinite(ref.t), np.round(np.asarray(ref.t, np.float64) * 32) / 32,
                     np.inf).astype(F32)
    post = ref.t[np.isfinite(ref.t)]
    assert np.isin(post, pre[np.isfinite(pre)]).mean() > 0.5
    errors = rng.normal(size=ref.t.shape).astype(F32)
    check_binned(ref, errors, "tied times, several slices")
    check_binned(ref, errors, "tied times, batch sum", batch_sum=True)


def test_binned_pre_layer_wider_than_one_workgroup_round():
    """More pre neurons than the pre-spike listing's workgroup has threads (1024): the
    listing counts the bins in a first pass and lists the spikes round by round."""
    rng = np.random.default_rng(13)
    B, I, Kp, J, K = 6, 1100, 3, 24, 4
    pre, _ = multi_spike_inputs(rng, B, I, Kp, t_max=1.0)
    W = (0.8 * rng.normal(0, 1.0, size=(J, I))).astype(F32)
    ref = olif.lif_forward(pre, W, 0.130, 1.0, 1.0, K, tmax=1.0)
    assert ref.n.sum() > 30
    check_binned(ref, rng.normal(size=ref.t.shape).astype(F32), "pre layer of 1100 neurons")
