"""The continuous-batching server: every request's result is bitwise the same alone and in
a mixed batch, matches the FP64 oracle (identical counts, times within 1e-5 relative), batches
follow the length classes, and the resident graphs stay within the memory budget."""
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
    # oracle on a few requests (their events again: a served request's leave the arena)
    for r in make_requests(40, p_min=50, p_max=600, seed=3)[:4]:
        per, _ = oev.events_to_per_neuron(r.indices[None, :], r.times[None, :], 784)
        h = olif.lif_forward(per, layers[0].weights, 0.130, 1.0, 1.0, 30, tmax=1.0)
        o = olif.lif_forward(h.t, layers[1].weights, 0.130, 1.3, 1.3, 30, tmax=1.0)
        n_m, t_m, _ = res_mixed[r.rid]
        assert np.array_equal(n_m, o.n[0])
        m = np.arange(30)[None, :] < o.n[0][:, None]
        assert np.all(np.abs(t_m[m] - o.t[0][m]) <= 1e-5 * np.abs(o.t[0][m]))


def test_memory_budget_bounds_graphs_and_results_hold():
    """The budget bounds the device memory of the resident batch graphs, measured when they
    are made: with room for about one 32-request graph, a mix of lengths (graphs for 128 to 1024 events)
    and batch sizes makes the server evict least-recently-used graphs and shrink batches, its device
    memory never exceeds what it holds by more than the graph being made, and every result
    equals that of an unbounded server."""
    from bats_amd.serving import InferenceServer, make_requests
    layers = _layers()
    free = InferenceServer(layers, 784, max_batch=64, min_events_bucket=128,
                           length_classes=(256, 512))
    for r in make_requests(120, p_min=50, p_max=1000, seed=4):
        free.submit(r)
    free.drain()
    ref = {r.rid: (r.counts, r.spike_times, r.pred) for r in free.completed}
    del free
    probe = InferenceServer(layers, 784, max_batch=64, min_events_bucket=128)
    budget = int(1.3 * probe.graph_bytes_estimate(32, 1024))
    del probe
    srv = InferenceServer(layers, 784, max_batch=64, memory_budget=budget,
                          min_events_bucket=128, length_classes=(256, 512))
    base = torch.cuda.memory_reserved()
    for r in make_requests(120, p_min=50, p_max=1000, seed=4):
        srv.submit(r)
    while srv._waiting() or srv.inflight:
        srv.poll()
        assert srv.memory_in_use <= budget
        assert srv.memory_in_use == sum(g.nbytes for g in srv.graphs.values())
        assert torch.cuda.memory_reserved() - base <= budget + (64 << 20)
    assert len(srv.completed) == 120
    assert srv.graphs_evicted > 0
    for r in srv.completed:
        n, t, p = ref[r.rid]
        assert np.array_equal(r.counts, n) and np.array_equal(r.spike_times, t) and r.pred == p


def test_length_classes_and_longest_first():
    """Batches take the class of the oldest request and shorter ones only; results are the
    same as in arrival-order batches."""
    from bats_amd.serving import InferenceServer, make_requests
    layers = _layers()
    srv = InferenceServer(layers, 784, max_batch=16, length_classes=(256, 512, 700))
    reqs = make_requests(40, p_min=50, p_max=900, seed=8)
    for k, r in enumerate(reqs):
        r.arrival = float(k)
        srv.submit(r)
    assert len(srv.queues) == 4
    batch = srv._take()
    oldest = srv.events_class(reqs[0])
    assert 0 < oldest < 3  # (the seed's first request is of a middle class)
    assert batch[0] is reqs[0]
    assert all(srv.events_class(r) <= oldest for r in batch)
    assert [r.arrival for r in batch] == sorted(r.arrival for r in batch)
    srv._give_back(batch)
    assert sorted(r.rid for r in srv.queue) == list(range(40))
    srv.drain()
    assert len(srv.completed) == 40
    one = InferenceServer(layers, 784, max_batch=1, min_batch_bucket=1)
    for r in make_requests(40, p_min=50, p_max=900, seed=8)[:8]:
        one.submit(r)
        one.drain()
    got = {r.rid: r for r in srv.completed}
    for r in one.completed:
        assert np.array_equal(got[r.rid].counts, r.counts)
        assert np.array_equal(got[r.rid].spike_times, r.spike_times)


def test_many_events_on_one_input():
    """The first layer reads the packed lists directly, so any number of events per input
    neuron is served in one pass: a request with 300 events on one input neuron and one with
    40 events on each of 20 neurons, in a mix, against the oracle and against themselves alone."""
    from bats_amd.serving import InferenceServer, Request, make_requests
    layers = _layers()
    rng = np.random.default_rng(9)
    special = []
    for k, (neurons, per) in enumerate(((np.array([5]), 300), (np.arange(20) * 37, 40))):
        idx = np.repeat(neurons, per).astype(np.int32)
        tm = rng.uniform(0.0, 1.0, size=len(idx)).astype(F32)
        order = np.argsort(tm, kind="stable")
        special.append((idx[order], tm[order]))
    mixed = InferenceServer(layers, 784, max_batch=32)
    for r in make_requests(20, p_min=50, p_max=900, seed=6):
        mixed.submit(r)
    for k, (idx, tm) in enumerate(special):
        mixed.submit(Request(rid=100 + k, indices=idx.copy(), times=tm.copy(), arrival=0.0))
    mixed.drain()
    got = {r.rid: r for r in mixed.completed}
    assert len(got) == 22
    alone = InferenceServer(layers, 784, max_batch=1, min_batch_bucket=1)
    for k, (idx, tm) in enumerate(special):
        r = Request(rid=100 + k, indices=idx.copy(), times=tm.copy(), arrival=0.0)
        alone.submit(r)
        alone.drain()
        assert np.array_equal(r.counts, got[100 + k].counts)
        assert np.array_equal(r.spike_times, got[100 + k].spike_times)
        per, _ = oev.events_to_per_neuron(idx[None, :], tm[None, :], 784)
        h = olif.lif_forward(per, layers[0].weights, 0.130, 1.0, 1.0, 30, tmax=1.0)
        o = olif.lif_forward(h.t, layers[1].weights, 0.130, 1.3, 1.3, 30, tmax=1.0)
        assert np.array_equal(r.counts, o.n[0])
        m = np.arange(30)[None, :] < o.n[0][:, None]
        assert np.all(np.abs(r.spike_times[m] - o.t[0][m]) <= 1e-5 * np.abs(o.t[0][m]))
