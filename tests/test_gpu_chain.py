"""Chained dense layers (bats_lif_forward_chain / bats_lif_forward_events): the second layer's
pre exps and event lists made by the first layer's call, from its own spikes, against two
plain bats_lif_forward calls on the same inputs and weights.

Both ways sort the same (time, slot) keys, so the second layer sees the same event lists and
every output is bit-identical: counts, spike times (+inf tails included), a, x, z at the
spikes, and the second layer's pre exps at the first layer's spikes. Cases: the batch of the
784-800-10 training step (256 samples: one workgroup per sample, the solver sorts the spikes
in LDS), a small batch (several workgroups per sample: the chained call builds the lists with
the per-sample sort instead), single-spike layers (K = 1), a layer whose samples have more
than 8192 spikes (the sort keys go to the global key region), three layers in a row (the
middle one consumes lists and makes the next), a sample with no input spikes, and the Network
API, whose forward chains dense layers by itself (compared with calling the layers' forward
one by one, which does not chain).
"""
import numpy as np
import pytest
import torch

from bats_amd import Network, ops
from bats_amd._constants import lif_constants
from bats_amd.Layers import InputLayer, LIFLayer
from tests.helpers import latency_inputs, spike_slots

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
T_MAX = 1.0


def _dev(a, dtype=torch.float32):
    return torch.as_tensor(a, dtype=dtype, device=DEV)


def _assert_same_layer(got, ref, what):
    n_g, t_g, a_g, x_g, z_g = (v.cpu().numpy() for v in got[:5])
    n_r, t_r, a_r, x_r, z_r = (v.cpu().numpy() for v in ref[:5])
    assert np.array_equal(n_g, n_r), f"{what}: spike counts differ"
    assert np.array_equal(t_g, t_r), f"{what}: spike times differ"
    m = spike_slots(n_r, t_r.shape[2])
    for name, g, r in (("a", a_g, a_r), ("x", x_g, x_r), ("z", z_g, z_r)):
        assert np.array_equal(g[m], r[m]), f"{what}: {name} differs at the spikes"


def _assert_same_exps(got_es, got_e, ref_es, ref_e, pre_n, what):
    m = spike_slots(pre_n.cpu().numpy(), ref_es.shape[2])
    for name, g, r in (("es", got_es, ref_es), ("e", got_e, ref_e)):
        assert np.array_equal(g.cpu().numpy()[m], r.cpu().numpy()[m]), f"{what}: pre {name} differ"


def _two_layers(B, I, J1, J2, K1, K2, seed, w_scale=1.0, w_abs=False):
    rng = np.random.default_rng(seed)
    t, _ = latency_inputs(rng, B, I)
    w1 = rng.normal(0.0, w_scale, size=(J1, I)).astype(np.float32)
    if w_abs:
        w1 = np.abs(w1)
    w2 = rng.normal(0.0, 1.0, size=(J2, J1)).astype(np.float32)
    return _dev(t), _dev(w1), _dev(w2)


def _check_pair(B, I, J1, J2, K1, K2, seed, **kw):
    t, w1, w2 = _two_layers(B, I, J1, J2, K1, K2, seed, **kw)
    c1 = lif_constants(0.130, 1.0, 1.0)
    c2 = lif_constants(0.130, 1.3, 1.3)
    ref1 = ops.lif_forward(t, w1, *c1, T_MAX, K1)
    ref2 = ops.lif_forward(ref1[1], w2, *c2, T_MAX, K2)
    got1, lists = ops.lif_forward_chain(t, w1, *c1, T_MAX, K1, (c2[0], c2[1], c2[2], T_MAX))
    got2 = ops.lif_forward_events(lists, w2, *c2, K2)
    torch.cuda.synchronize()
    _assert_same_layer(got1, ref1, "first layer")
    _assert_same_exps(got1[5], got1[6], ref1[5], ref1[6], torch.isfinite(t).sum(2), "first layer")
    _assert_same_layer(got2, ref2, "second layer")
    _assert_same_exps(got2[5], got2[6], ref2[5], ref2[6], ref1[0], "second layer")
    return ref1, ref2


def test_chain_training_batch_fused():
    ref1, ref2 = _check_pair(256, 784, 800, 10, 30, 30, seed=0)
    assert ref1[0].sum().item() > 256 * 100 and ref2[0].sum().item() > 0


def test_chain_small_batch_sorted_path():
    _check_pair(16, 784, 800, 10, 30, 30, seed=1)


def test_chain_single_spike_layers():
    _check_pair(256, 784, 800, 10, 1, 1, seed=2)


def test_chain_lists_longer_than_the_lds_sort():
    # strongly excitatory weights: more than 8192 hidden spikes in a sample
    ref1, _ = _check_pair(256, 784, 1000, 10, 30, 30, seed=3, w_scale=2.0, w_abs=True)
    assert ref1[0].sum(1).max().item() > 8192


def test_chain_three_layers_and_a_silent_sample():
    rng = np.random.default_rng(4)
    B = 256
    t, _ = latency_inputs(rng, B, 784)
    t[7] = np.inf  # a sample without input spikes
    t = _dev(t)
    ws = [_dev(rng.normal(0, 1.0, size=s).astype(np.float32))
          for s in ((800, 784), (800, 800), (800, 800), (10, 800))]
    c = lif_constants(0.130, 1.0, 1.0)
    co = lif_constants(0.130, 1.3, 1.3)
    consts = [c, c, c, co]
    refs, x = [], t
    for w, k in zip(ws, consts):
        refs.append(ops.lif_forward(x, w, *k, T_MAX, 30))
        x = refs[-1][1]
    nxt = lambda k: (k[0], k[1], k[2], T_MAX)  # noqa: E731
    got = []
    out, lists = ops.lif_forward_chain(t, ws[0], *consts[0], T_MAX, 30, nxt(consts[1]))
    got.append(out)
    for li in (1, 2):
        out, lists = ops.lif_forward_events(lists, ws[li], *consts[li], 30, chain=nxt(consts[li + 1]))
        got.append(out)
    got.append(ops.lif_forward_events(lists, ws[3], *consts[3], 30))
    torch.cuda.synchronize()
    for li in range(4):
        _assert_same_layer(got[li], refs[li], f"layer {li + 1}")
    for li in range(1, 4):
        _assert_same_exps(got[li][5], got[li][6], refs[li][5], refs[li][6], refs[li - 1][0],
                          f"layer {li + 1}")
    assert refs[0][0][7].sum().item() == 0


def test_network_forward_chains_dense_layers():
    rng = np.random.default_rng(5)
    B = 256
    t, n = latency_inputs(rng, B, 784)
    w1 = rng.normal(0, 1.0, size=(800, 784)).astype(np.float32)
    w2 = rng.normal(0, 1.0, size=(10, 800)).astype(np.float32)

    def build():
        net = Network()
        inp = InputLayer(n_neurons=784, name="input")
        net.add_layer(inp, input=True)
        hid = LIFLayer(inp, tau_s=0.13, theta=1.0, delta_theta=1.0,
                       weight_initializer=lambda a, b: w1, max_n_spike=30, n_neurons=800,
                       name="h")
        net.add_layer(hid)
        out = LIFLayer(hid, tau_s=0.13, theta=1.3, delta_theta=1.3,
                       weight_initializer=lambda a, b: w2, max_n_spike=30, n_neurons=10,
                       name="o")
        net.add_layer(out)
        return net, inp, hid, out

    net, _, hid, out = build()
    net.forward(t, n, max_simulation=T_MAX, training=True)
    ref_net, inp, rhid, rout = build()
    inp.set_spike_trains(t, n)
    rhid.forward(T_MAX, True)  # one layer at a time: no chaining
    rout.forward(T_MAX, True)
    torch.cuda.synchronize()
    for a, b in ((hid, rhid), (out, rout)):
        ga, gb = a.spike_trains, b.spike_trains
        assert torch.equal(ga[1], gb[1]) and torch.equal(ga[0], gb[0])
    # lists made for spike trains that were written to afterwards are not used
    hid.forward(T_MAX, True, next_layer=out)
    hid.spike_trains[0][3].fill_(float("inf"))
    out.forward(T_MAX, True)
    ref = ops.lif_forward(hid.spike_trains[0], out.weights, *out.constants, T_MAX, 30)
    torch.cuda.synchronize()
    assert torch.equal(out.spike_trains[1], ref[0])
    assert torch.equal(out.spike_trains[0], ref[1])


def test_packed_first_layer_chained():
    """The inference server's first layer (packed, time-ordered event lists) chained to the
    output layer against the packed forward followed by a plain forward."""
    rng = np.random.default_rng(6)
    B, I, J1, J2, K = 256, 784, 800, 10, 30
    lists = []
    for _ in range(B):
        P = int(rng.integers(50, 2001))
        idx = rng.integers(0, I, size=P).astype(np.int32)
        tm = np.sort(rng.uniform(0.0, 1.0, size=P)).astype(np.float32)
        lists.append((idx, tm))
    off = np.zeros(B + 1, np.int32)
    off[1:] = np.cumsum([len(i) for i, _ in lists])
    idx = _dev(np.concatenate([i for i, _ in lists]), torch.int32)
    tm = _dev(np.concatenate([t for _, t in lists]))
    off = _dev(off, torch.int32)
    P_max = 2048
    w1 = _dev(rng.normal(0, 1.0, size=(J1, I)).astype(np.float32))
    w2 = _dev(rng.normal(0, 1.0, size=(J2, J1)).astype(np.float32))
    c1 = lif_constants(0.130, 1.0, 1.0)
    c2 = lif_constants(0.130, 1.3, 1.3)
    ref1 = ops.lif_forward_packed(off, idx, tm, P_max, w1, *c1, T_MAX, K)
    ref2 = ops.lif_forward(ref1[1], w2, *c2, T_MAX, K)
    got1, ev = ops.lif_forward_packed(off, idx, tm, P_max, w1, *c1, T_MAX, K,
                                      chain=(c2[0], c2[1], c2[2], T_MAX))
    got2 = ops.lif_forward_events(ev, w2, *c2, K)
    torch.cuda.synchronize()
    _assert_same_layer(got1, ref1, "first layer")
    _assert_same_layer(got2, ref2, "second layer")
    _assert_same_exps(got2[5], got2[6], ref2[5], ref2[6], ref1[0], "second layer")
