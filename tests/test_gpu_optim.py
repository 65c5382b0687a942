"""Optimizer steps on the GPU against the FP64 oracle: Adam and NAdam over several layers in
one launch (with a layer that has no gradient), with the step counter on the host and on the
device, and Network.apply_deltas adding every layer's delta in one launch. Bar: deltas, m and
v within 1e-6 relative of the oracle (float64 arithmetic with float32 storage on both sides),
weights after apply_deltas equal to the float32 sum."""
import numpy as np
import pytest
import torch

from oracle import optim as oopt

pytestmark = pytest.mark.gpu
F32 = np.float32


def _close(g, r, rtol=1e-6):
    g = np.asarray(g, np.float64)
    r = np.asarray(r, np.float64)
    assert np.all(np.abs(g - r) <= rtol * np.abs(r) + 1e-12), float(np.max(np.abs(g - r)))


@pytest.mark.parametrize("nadam", [False, True])
@pytest.mark.parametrize("capturable", [False, True])
def test_multi_layer_steps_match_oracle(nadam, capturable):
    from bats_amd.Optimizers import AdamOptimizer, NAdamOptimizer
    rng = np.random.default_rng(8)
    shapes = [(800, 784), None, (10, 800), (37, 5)]
    cls = NAdamOptimizer if nadam else AdamOptimizer
    opt = cls(learning_rate=0.003, capturable=capturable)
    ref = (oopt.NAdam if nadam else oopt.Adam)(0.003)
    for step in range(4):
        grads = [None if s is None else rng.normal(size=s).astype(F32) for s in shapes]
        dg = [None if g is None else torch.as_tensor(g, device="cuda") for g in grads]
        out = opt.step(dg)
        want = ref.step(grads)
        for k, (o, w) in enumerate(zip(out, want)):
            if w is None:
                assert o is None
                continue
            _close(o.cpu().numpy(), w)
            _close(opt._m[k].cpu().numpy(), ref.m[k])
            _close(opt._v[k].cpu().numpy(), ref.v[k])


def test_apply_deltas_in_one_launch():
    from bats_amd import Network
    from bats_amd.Layers import InputLayer, LIFLayer
    rng = np.random.default_rng(9)
    w1 = rng.normal(size=(50, 30)).astype(F32)
    w2 = rng.normal(size=(4, 50)).astype(F32)
    net = Network()
    inp = InputLayer(n_neurons=30, name="in")
    net.add_layer(inp, input=True)
    l1 = LIFLayer(inp, n_neurons=50, tau_s=0.1, theta=1.0, delta_theta=1.0,
                  weight_initializer=lambda a, b: w1, max_n_spike=3, name="l1")
    net.add_layer(l1)
    l2 = LIFLayer(l1, n_neurons=4, tau_s=0.1, theta=1.0, delta_theta=1.0,
                  weight_initializer=lambda a, b: w2, max_n_spike=3, name="l2")
    net.add_layer(l2)
    d1 = rng.normal(size=(50, 30)).astype(F32)
    d2 = rng.normal(size=(1, 4, 50)).astype(F32)  # a leading unit axis, as a batch mean has
    net.apply_deltas([None, torch.as_tensor(d1, device="cuda"), torch.as_tensor(d2, device="cuda")])
    assert np.array_equal(l1.weights.cpu().numpy(), w1 + d1)
    assert np.array_equal(l2.weights.cpu().numpy(), w2 + d2[0])
