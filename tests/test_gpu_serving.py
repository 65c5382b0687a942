"""The continuous-batching server: every request's result is bitwise the same alone and in
a mixed batch, matches the FP64 oracle (identical counts, times within 1e-5 relative), and
admission never exceeds the memory budget."""
import numpy as np
import pytest
import torch

from oracle import events as oev
from oracle import lif as olif

pytestmark = pytest.mark.gpu
F32 = np.float32


def _layers(seed=0):
    from bats_amd.serving import LayerSpec
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, 1.0, size=(800, 784)).astype(F32) * 0.5
    W2 = rng.normal(0, 1.0, size=(10, 800)).astype(F32)
    return [LayerSpec(W1, 0.130, 1.0, 1.0, 30), LayerSpec(W2, 0.130, 1.3, 1.3, 30)]


def test_alone_equals_mixed_and_oracle():
    from bats_amd.serving import InferenceServer, make_requests
    layers = _layers()
    reqs = make_requests(40, p_min=50, p_max=600, seed=3)
    mixed = InferenceServer(layers, 784, max_batch=16)
    for r in reqs:
        mixed.submit(r)
    mixed.drain()
    torch.cuda.synchronize()
    res_mixed = {r.rid: (r.counts.copy(), r.spike_times.copy(), r.pred) for r in mixed.completed}
    assert len(res_mixed) == 40
    alone = InferenceServer(layers, 784, max_batch=1, min_batch_bucket=1)
    for r in make_requests(40, p_min=50, p_max=600, seed=3)[:12]:
        alone.submit(r)
        alone.drain()
        n_m, t_m, p_m = res_mixed[r.rid]
        assert np.array_equal(r.counts, n_m)
        assert np.array_equal(r.spike_times, t_m)
        assert r.pred == p_m
    # oracle on a few requests
    for r in reqs[:4]:
        per, _ = oev.events_to_per_neuron(r.indices[None, :], r.times[None, :], 784)
        h = olif.lif_forward(per, layers[0].weights, 0.130, 1.0, 1.0, 30, tmax=1.0)
        o = olif.lif_forward(h.t, layers[1].weights, 0.130, 1.3, 1.3, 30, tmax=1.0)
        n_m, t_m, _ = res_mixed[r.rid]
        assert np.array_equal(n_m, o.n[0])
        m = np.arange(30)[None, :] < o.n[0][:, None]
        assert np.all(np.abs(t_m[m] - o.t[0][m]) <= 1e-5 * np.abs(o.t[0][m]))


def test_memory_budget_admission():
    from bats_amd.serving import InferenceServer, make_requests
    layers = _layers()
    reqs = make_requests(30, p_min=50, p_max=300, seed=4)
    probe = InferenceServer(layers, 784, max_batch=64)
    for r in reqs:
        probe.submit(r)
    budget = 5 * max(probe.request_bytes(r) for r in reqs)
    srv = InferenceServer(layers, 784, max_batch=64, memory_budget=budget)
    for r in make_requests(30, p_min=50, p_max=300, seed=4):
        srv.submit(r)
    peak = 0
    while srv.queue or srv.inflight:
        srv.poll()
        peak = max(peak, srv.memory_in_use)
        assert srv.memory_in_use <= budget
    assert len(srv.completed) == 30
    assert srv.memory_in_use == 0


def test_spikes_bucket_overflow_is_served_again():
    # with a spikes-per-input-neuron bucket of 1, every request with two events on one input
    # neuron overflows on the device and is served again in a bucket that holds it; results
    # equal those of a server whose bucket holds every request at once
    from bats_amd.serving import InferenceServer, make_requests
    layers = _layers()
    wide = InferenceServer(layers, 784, max_batch=32, min_spikes_bucket=16)
    for r in make_requests(48, p_min=50, p_max=900, seed=6):
        wide.submit(r)
    wide.drain()
    ref = {r.rid: (r.counts, r.spike_times, r.pred) for r in wide.completed}
    assert all(r.multiplicity == 0 for r in wide.completed)
    tight = InferenceServer(layers, 784, max_batch=32, min_spikes_bucket=1)
    for r in make_requests(48, p_min=50, p_max=900, seed=6):
        tight.submit(r)
    tight.drain()
    assert len(tight.completed) == 48 and tight.memory_in_use == 0
    redone = 0
    for r in tight.completed:
        true_mult = int(np.bincount(r.indices, minlength=784).max())
        if true_mult > 1:
            assert r.multiplicity == true_mult
            redone += 1
        n, t, p = ref[r.rid]
        assert np.array_equal(r.counts, n) and np.array_equal(r.spike_times, t) and r.pred == p
    assert redone > 10
