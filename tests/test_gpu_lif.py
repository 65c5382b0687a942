"""GPU parity of the dense LIF forward and backward against the FP64 oracle.

Tolerances (stated here, asserted below):
  * spike counts: identical (bit-exact);
  * spike times, a, x, z of every emitted spike: relative error <= 1e-5;
  * pre-spike exps: identical float32 values;
  * reset-chain errors, weight gradients, pre-spike errors:
    |gpu - ref| <= 1e-4 |ref| + 1e-6 max|ref| (elementwise).
"""
import numpy as np
import pytest
import torch

from oracle import lif as olif
from tests.helpers import (assert_close_rel, assert_forward_match, latency_inputs,
                           multi_spike_inputs)

pytestmark = pytest.mark.gpu

F32 = np.float32


def gpu_forward(pre_t, W, tau_s, theta, dth, K, tmax, solver_mode=0):
    from bats_amd import ops
    from bats_amd._constants import lif_constants
    ts, tau, c, dtt = lif_constants(tau_s, theta, dth)
    pt = torch.as_tensor(pre_t, device="cuda")
    Wd = torch.as_tensor(W, device="cuda")
    out = ops.lif_forward(pt, Wd, ts, tau, c, dtt, tmax, K, solver_mode=solver_mode)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def check_forward(pre_t, W, tau_s, theta, dth, K, tmax, label, solver_mode=0):
    print(f"[forward] {label}: pre {pre_t.shape} W {W.shape} K={K}", flush=True)
    ref = olif.lif_forward(pre_t, W, tau_s, theta, dth, K, tmax=tmax)
    n, t, a, x, z, es, e = gpu_forward(pre_t, W, tau_s, theta, dth, K, tmax, solver_mode)
    assert_forward_match(n, t, a, x, z, ref)
    fin = np.isfinite(pre_t)
    assert np.array_equal(es[fin], ref.pre_exp_s[fin])
    assert np.array_equal(e[fin], ref.pre_exp[fin])
    return ref, (n, t, a, x, z, es, e)


def test_forward_reference_seed42_kernel_case():
    # inputs of tests/CudaKernels/Inference/compute_spike_times_tests.py:44-55
    from oracle.events import events_to_per_neuron
    np.random.seed(42)
    idx = np.random.randint(10, size=(1, 100))
    tt = np.sort(np.random.uniform(0.0, 0.5, size=(1, 100)), axis=1)
    tt[0, -28:] = np.inf
    W = np.random.normal(0, 1.0, size=(5, 10)).astype(F32)
    ii = np.where(np.isfinite(tt), idx, -1).astype(np.int32)
    per, _ = events_to_per_neuron(ii, tt.astype(F32), 10)
    ref, _ = check_forward(per, W, 0.1, 0.1, 0.2, 128, 0.5, "seed42 kernel")
    assert ref.n.sum() > 10


def test_forward_reference_seed42_layer_case():
    # inputs of tests/Layers/LIFLayer_tests.py:44-58 (I=1000, J=50, up to 4 spikes per input)
    np.random.seed(42)
    n_in = np.random.randint(0, 5, size=(1, 1000))
    spk = np.full((1, 1000, 5), np.inf, dtype=F32)
    for s, ns in zip(spk, n_in):
        for row, k in zip(s, ns):
            row[:k] = np.sort(np.random.uniform(0.0, 0.5, size=(k,)))
    W = np.random.normal(0, 1.0, size=(50, 1000)).astype(F32)
    ref, _ = check_forward(spk, W, 0.1, 0.1, 0.2, 128, 0.5, "seed42 layer")
    assert ref.n.sum() > 100


@pytest.mark.parametrize("seed", [0, 1])
def test_forward_multi_spike_hidden_c2_layer1(seed):
    # 784 -> 800, K = 30, B = 256, W ~ N(0, 1) (the active variant), theta = 1, t_max = 1
    rng = np.random.default_rng(seed)
    t, _ = latency_inputs(rng, 256, 784)
    W = rng.normal(0, 1.0, size=(800, 784)).astype(F32)
    ref, _ = check_forward(t, W, 0.130, 1.0, 1.0, 30, 1.0, f"C2 layer1 seed{seed}")
    assert ref.n.sum() > 256 * 100


def test_forward_output_layer_from_hidden_spikes():
    rng = np.random.default_rng(3)
    t, _ = latency_inputs(rng, 64, 784)
    W1 = (2.0 * rng.normal(0, 1.0, size=(800, 784))).astype(F32)
    h = olif.lif_forward(t, W1, 0.130, 1.0, 1.0, 30, tmax=1.0)
    W2 = rng.normal(0, 1.0, size=(10, 800)).astype(F32)
    ref, _ = check_forward(h.t, W2, 0.130, 1.3, 1.3, 30, 1.0, "output layer")
    assert ref.n.sum() > 0


def test_forward_ttfs_single_spike():
    rng = np.random.default_rng(4)
    t, _ = latency_inputs(rng, 50, 784)
    W = rng.normal(0, 1.0, size=(800, 784)).astype(F32)
    check_forward(t, W, 0.130, 1.0, 1.0, 1, 1.0, "TTFS K=1")


def test_forward_buffer_full_and_tmax():
    rng = np.random.default_rng(5)
    t, _ = multi_spike_inputs(rng, 8, 100, 5, t_max=1.0)
    W = np.abs(rng.normal(0, 3.0, size=(64, 100))).astype(F32)
    ref, _ = check_forward(t, W, 0.05, 0.2, 0.2, 3, 0.6, "buffer full + t_max")
    assert (ref.n == 3).any()


def test_forward_empty_batch():
    t = np.full((4, 30, 2), np.inf, dtype=F32)
    W = np.ones((7, 30), dtype=F32)
    ref, (n, tt, *_rest) = check_forward(t, W, 0.1, 0.1, 0.1, 4, 1.0, "empty")
    assert n.sum() == 0 and np.all(np.isinf(tt))


@pytest.mark.parametrize("n_events", [1024, 1025, 2800])
def test_forward_event_list_sort_paths(n_events):
    # up to 1024 events per sample the sort runs in registers, above it in LDS; the first two
    # cases sit on either side of that boundary with every slot finite
    rng = np.random.default_rng(7)
    if n_events <= 1025:
        t = rng.uniform(0.0, 1.0, size=(3, n_events, 1)).astype(F32)
    else:
        t, _ = multi_spike_inputs(rng, 3, 1000, 8, t_max=1.0)
        counts = np.isfinite(t).sum(axis=(1, 2))
        assert (counts > 1025).all() and (counts < 16384).all()
    W = (rng.normal(0.2, 1.0, size=(40, t.shape[1])) * (30.0 / np.sqrt(t.shape[1]))).astype(F32)
    ref, _ = check_forward(t, W, 0.130, 1.0, 1.0, 12, 1.0, f"sort path {n_events}")
    assert ref.n.sum() > 0


@pytest.mark.parametrize("n_in,k_in,lo,hi", [(800, 10, 4097, 8192), (1000, 12, 8193, 16384)])
def test_forward_long_event_lists_merge_sort(n_in, k_in, lo, hi):
    # lists of 4097..8192 events (merged between two LDS buffers) and of 8193..16384 events in
    # a layer of more slots (the merge passes alternate between the LDS and global memory)
    rng = np.random.default_rng(8)
    n = rng.integers(k_in - 3, k_in + 1, size=(2, n_in))
    t = np.sort(rng.uniform(0.0, 1.0, size=(2, n_in, k_in)), axis=2).astype(F32)
    t[np.arange(k_in)[None, None, :] >= n[:, :, None]] = np.inf
    counts = np.isfinite(t).sum(axis=(1, 2))
    assert ((counts >= lo) & (counts <= hi)).all(), counts
    W = rng.normal(0, 0.3, size=(10, n_in)).astype(F32)
    check_forward(t, W, 0.130, 1.3, 1.3, 30, 1.0, f"merge sort, {counts.tolist()} events")


def test_forward_large_event_list_global_sort():
    # 800 x 30 slots per sample with ~20k spikes: exceeds the LDS sort capacity
    rng = np.random.default_rng(6)
    n = rng.integers(22, 31, size=(2, 800))
    t = np.sort(rng.uniform(0.0, 1.0, size=(2, 800, 30)), axis=2).astype(F32)
    t[np.arange(30)[None, None, :] >= n[:, :, None]] = np.inf
    assert (np.isfinite(t).sum(axis=(1, 2)) > 16384).all()
    W = rng.normal(0, 0.3, size=(10, 800)).astype(F32)
    check_forward(t, W, 0.130, 1.3, 1.3, 30, 1.0, "global sort")


# ------------------------------------------------------------------ backward
def gpu_backward(ref, errors, per_sample, pre_trainable):
    from bats_amd import ops
    dev = "cuda"
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    err = T(errors.astype(F32))
    grad, pre = ops.lif_backward(T(ref.n), T(ref.t), T(ref.a), T(ref.x), T(ref.z), err,
                                 T(ref.pre_times), T(ref.pre_exp_s), T(ref.pre_exp), T(ref.W),
                                 ref.tau_s, ref.tau, ref.c, ref.dtt, per_sample=per_sample,
                                 pre_trainable=pre_trainable)
    torch.cuda.synchronize()
    return err.cpu().numpy(), grad.cpu().numpy(), None if pre is None else pre.cpu().numpy()


def check_backward(ref, errors, label):
    print(f"[backward] {label}", flush=True)
    e_ref, g_ref, p_ref = olif.lif_backward(ref, errors, pre_trainable=True)
    e_g, g_g, p_g = gpu_backward(ref, errors, per_sample=True, pre_trainable=True)
    assert_close_rel(e_g, e_ref, 1e-4, 1e-6, "reset-chain errors")
    assert_close_rel(g_g, g_ref, 1e-4, 1e-6, "per-sample gradient")
    assert_close_rel(p_g, p_ref, 1e-4, 1e-6, "pre-spike errors")
    _, gm, none = gpu_backward(ref, errors, per_sample=False, pre_trainable=False)
    assert none is None
    assert_close_rel(gm, olif.batch_mean(g_ref), 1e-4, 1e-6, "batch-mean gradient")


def test_backward_seed42_layer_case():
    np.random.seed(42)
    n_in = np.random.randint(0, 5, size=(1, 1000))
    spk = np.full((1, 1000, 5), np.inf, dtype=F32)
    for s, ns in zip(spk, n_in):
        for row, k in zip(s, ns):
            row[:k] = np.sort(np.random.uniform(0.0, 0.5, size=(k,)))
    W = np.random.normal(0, 1.0, size=(50, 1000)).astype(F32)
    ref = olif.lif_forward(spk, W, 0.1, 0.1, 0.2, 128, tmax=0.5)
    rng = np.random.default_rng(0)
    check_backward(ref, rng.normal(size=ref.t.shape).astype(F32), "seed42 layer")


def test_backward_hidden_and_output_layers():
    rng = np.random.default_rng(7)
    t, _ = latency_inputs(rng, 32, 784)
    W1 = (1.5 * rng.normal(0, 1.0, size=(800, 784))).astype(F32)
    h = olif.lif_forward(t, W1, 0.130, 1.0, 1.0, 30, tmax=1.0)
    W2 = rng.normal(0, 1.0, size=(10, 800)).astype(F32)
    o = olif.lif_forward(h.t, W2, 0.130, 1.3, 1.3, 30, tmax=1.0)
    check_backward(o, rng.normal(size=o.t.shape).astype(F32), "output layer")
    check_backward(h, rng.normal(size=h.t.shape).astype(F32), "hidden layer")


def test_backward_hidden_to_hidden_c3_shape():
    # 784 -> 800 -> 800 with 30-spike buffers: the second hidden layer's pre-spike errors
    # (the previous layer is trainable), its pre neurons firing several times
    rng = np.random.default_rng(3)
    t, _ = latency_inputs(rng, 8, 784)
    W1 = (2.0 * rng.normal(0, 1.0, size=(800, 784))).astype(F32)
    h1 = olif.lif_forward(t, W1, 0.130, 1.0, 1.0, 30, tmax=1.0)
    assert h1.n.max() >= 3
    W2 = (1.5 * rng.normal(0, 1.0, size=(800, 800))).astype(F32)
    h2 = olif.lif_forward(h1.t, W2, 0.130, 1.0, 1.0, 30, tmax=1.0)
    check_backward(h2, rng.normal(size=h2.t.shape).astype(F32), "hidden to hidden")


def _on_grid(t, step):
    """Spike times moved onto a grid (ascending order per row kept; +inf stays +inf)."""
    return np.where(np.isfinite(t), np.round(np.asarray(t, np.float64) / step) * step,
                    np.inf).astype(F32)


def test_tied_spike_times():
    """Many equal times. Latency inputs on a 1/32 grid tie across inputs (the event order
    must then follow the slot order, as the oracle's stable sort does); the hidden layer's
    spikes moved onto a 1/64 grid feed a second layer (multi-spike ties, the merge sort path);
    and the second layer's own spike times moved onto the same grid before its backward, so
    that pre and post spikes coincide: a pre spike counts only strictly before a post spike."""
    rng = np.random.default_rng(12)
    t, _ = latency_inputs(rng, 8, 784)
    t = _on_grid(t, 1.0 / 32)
    W1 = (2.0 * rng.normal(0, 1.0, size=(800, 784))).astype(F32)
    h, _ = check_forward(t, W1, 0.130, 1.0, 1.0, 30, 1.0, "tied latency inputs")
    assert h.n.sum() > 1000
    ht = _on_grid(h.t, 1.0 / 64)
    W2 = (1.5 * rng.normal(0, 1.0, size=(800, 800))).astype(F32)
    h2, _ = check_forward(ht, W2, 0.130, 1.0, 1.0, 30, 1.0, "tied multi-spike inputs")
    h2.t = _on_grid(h2.t, 1.0 / 64)
    pre = np.isfinite(h2.pre_times)
    post = h2.t[np.isfinite(h2.t)]
    assert np.isin(post, h2.pre_times[pre]).mean() > 0.5  # most post spikes tie with a pre spike
    check_backward(h2, rng.normal(size=h2.t.shape).astype(F32), "tied pre and post spikes")


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_forward_both_solver_forms(mode):
    # the lane-per-neuron (automatic, staged and scalar event records) and wave-per-neuron
    # solvers on the same hidden and output shapes, an odd number of post neurons, and
    # theta = 0 (no threshold quadratic: every interval takes the closed form); the form is
    # chosen per call
    rng = np.random.default_rng(11)
    t, _ = latency_inputs(rng, 16, 784)
    W1 = (2.0 * rng.normal(0, 1.0, size=(800, 784))).astype(F32)
    h, _ = check_forward(t, W1, 0.130, 1.0, 1.0, 30, 1.0, f"hidden mode {mode}", mode)
    W2 = rng.normal(0, 1.0, size=(10, 800)).astype(F32)
    check_forward(h.t, W2, 0.130, 1.3, 1.3, 30, 1.0, f"output mode {mode}", mode)
    tb, _ = multi_spike_inputs(rng, 8, 100, 5, t_max=1.0)
    W3 = np.abs(rng.normal(0, 3.0, size=(64, 100))).astype(F32)
    check_forward(tb, W3, 0.05, 0.2, 0.2, 3, 0.6, f"buffer full mode {mode}", mode)
    W4 = rng.normal(0.2, 1.0, size=(96, 100)).astype(F32)
    check_forward(tb, W4, 0.05, 0.0, 0.3, 4, 0.6, f"theta 0 mode {mode}", mode)
    W5 = rng.normal(0.3, 1.0, size=(97, 100)).astype(F32)
    check_forward(tb, W5, 0.05, 0.2, 0.2, 5, 0.6, f"odd J mode {mode}", mode)
    check_forward(tb, W5, 0.05, 0.0, 0.2, 5, 0.6, f"odd J theta 0 mode {mode}", mode)


@pytest.mark.parametrize("mode", [0, 5])
def test_forward_first_spike_time_segments(mode):
    # one spike per neuron with few (sample, neuron) pairs: the walk split into time segments
    # (mode 5 forces it; mode 0 takes it below two lane-solver waves per SIMD): single and
    # multi-spike inputs, spikes early and late in the list, silent neurons, an odd number of
    # post neurons, one sample, and a list shorter than one group per segment
    rng = np.random.default_rng(12)
    t, _ = latency_inputs(rng, 50, 784)
    W = rng.normal(0, 1.0, size=(800, 784)).astype(F32)
    ref, _ = check_forward(t, W, 0.130, 1.0, 1.0, 1, 1.0, f"TTFS hidden mode {mode}", mode)
    assert 0 < ref.n.sum() < ref.n.size
    W2 = rng.normal(0, 1.0, size=(10, 800)).astype(F32)
    check_forward(ref.t, W2, 0.130, 1.3, 1.3, 1, 1.0, f"TTFS output mode {mode}", mode)
    tb, _ = multi_spike_inputs(rng, 8, 300, 6, t_max=1.0)
    for scale, label in ((0.5, "late spikes"), (3.0, "early spikes")):
        Wm = rng.normal(0.1, scale, size=(97, 300)).astype(F32)
        r, _ = check_forward(tb, Wm, 0.05, 1.0, 1.0, 1, 1.0, f"{label} mode {mode}", mode)
        assert r.n.sum() > 0
    check_forward(t[:1], W, 0.130, 1.0, 1.0, 1, 0.5, f"one sample, t_max 0.5, mode {mode}", mode)
    short = np.full((4, 784, 1), np.inf, F32)
    short[:, :20, 0] = rng.uniform(0, 1, size=(4, 20)).astype(F32)
    check_forward(short, np.abs(W), 0.130, 1.0, 1.0, 1, 1.0, f"short lists mode {mode}", mode)


# ------------------------------------------------------------------ input encoding
def test_latency_encoder_matches_reference_formula():
    # experiments/mnist/Dataset.py:66-71 on uint8 and float pixels, zeros included
    from bats_amd import ops
    rng = np.random.default_rng(9)
    px = rng.integers(0, 256, size=(5, 28, 28)).astype(np.uint8)
    px[:, :3] = 0
    want = 100e-3 * (1 - (px.reshape(5, 784, 1) / 255))
    want[want == 100e-3] = np.inf
    for arg in (px, torch.as_tensor(px, device="cuda"), px.astype(np.float32)):
        t, n = ops.latency_encode(arg)
        torch.cuda.synchronize()
        assert np.array_equal(t.cpu().numpy(), want.astype(F32))
        assert np.array_equal(n.cpu().numpy(), np.isfinite(want[:, :, 0]).astype(np.int32))
