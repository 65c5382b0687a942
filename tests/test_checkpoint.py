"""Checkpoints: the reference's weight files plus the Adam state; resuming is exact."""
import numpy as np
import pytest
import torch

from bats_amd import checkpoint
from bats_amd.Optimizers import AdamOptimizer, NAdamOptimizer


def test_adam_state_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    opt = AdamOptimizer(learning_rate=0.003, beta_1=0.8)
    m = [None, torch.as_tensor(rng.normal(size=(5, 7)).astype(np.float32)), None,
         torch.as_tensor(rng.normal(size=(3, 5)).astype(np.float32))]
    v = [None if x is None else x.abs() for x in m]
    opt.load_state(m, v, 17)
    opt.store(tmp_path)
    other = AdamOptimizer(learning_rate=0.1, beta_1=0.8)
    other.restore(tmp_path, device="cpu")
    m2, v2, t2 = other.state
    assert t2 == 17 and other.learning_rate == np.float32(0.003)
    for a, b in zip(m + v, m2 + v2):
        assert (a is None) == (b is None)
        if a is not None:
            assert torch.equal(a, b)
    with pytest.raises(ValueError):
        NAdamOptimizer(learning_rate=0.1).restore(tmp_path, device="cpu")


def _network(w1, w2):
    from bats_amd import Network
    from bats_amd.Layers import InputLayer, LIFLayer
    net = Network()
    inp = InputLayer(n_neurons=w1.shape[1], name="input")
    net.add_layer(inp, input=True)
    net.add_layer(LIFLayer(inp, tau_s=0.13, theta=1.0, delta_theta=1.0, n_neurons=w1.shape[0],
                           weight_initializer=lambda a, b: w1, max_n_spike=8, name="hidden"))
    net.add_layer(LIFLayer(net.layers[-1], tau_s=0.13, theta=1.3, delta_theta=1.3,
                           n_neurons=w2.shape[0], weight_initializer=lambda a, b: w2,
                           max_n_spike=8, name="output"))
    return net


@pytest.mark.gpu
@pytest.mark.parametrize("capturable", [False, True])
def test_resume_is_exact(tmp_path, capturable):
    """Two steps, a checkpoint, two more steps from the restored state equal four steps in a
    row, bit for bit (weights and the step counter), with the step counter on the host or on
    the device."""
    from bats_amd.Losses import SpikeCountClassLoss
    rng = np.random.default_rng(5)
    B, I = 16, 784
    x = rng.uniform(size=(B, I))
    t = np.where(x > 0.1, 1.0 - x, np.inf).astype(np.float32)[:, :, None]
    n = np.isfinite(t[:, :, 0]).astype(np.int32)
    labels = rng.integers(0, 10, size=B)
    w1 = rng.normal(0, 1.0, size=(64, I)).astype(np.float32)
    w2 = rng.normal(0, 1.0, size=(10, 64)).astype(np.float32)
    loss = SpikeCountClassLoss(target_true=6, target_false=2)

    def steps(net, opt, k):
        for _ in range(k):
            net.forward(t, n, max_simulation=1.0, training=True)
            _, err = loss.compute_loss_and_errors(*net.output_spike_trains, labels)
            grads = net.backward(err)
            net.apply_deltas(opt.step([None if g is None else g[0] for g in grads]))

    net_a, opt_a = _network(w1, w2), AdamOptimizer(learning_rate=0.01, capturable=capturable)
    steps(net_a, opt_a, 4)
    net_b, opt_b = _network(w1, w2), AdamOptimizer(learning_rate=0.01, capturable=capturable)
    steps(net_b, opt_b, 2)
    checkpoint.save(tmp_path, net_b, opt_b)
    assert (tmp_path / "hidden_weights.npy").exists() and (tmp_path / "adam_state.npz").exists()
    net_c = _network(np.zeros_like(w1), np.zeros_like(w2))
    opt_c = AdamOptimizer(learning_rate=0.5, capturable=capturable)
    checkpoint.load(tmp_path, net_c, opt_c)
    steps(net_c, opt_c, 2)
    torch.cuda.synchronize()
    for la, lc in zip(net_a.layers[1:], net_c.layers[1:]):
        assert torch.equal(la.weights, lc.weights), la.name
    assert opt_c.state[2] == 4
    for ma, mc in zip(opt_a.state[0] + opt_a.state[1], opt_c.state[0] + opt_c.state[1]):
        assert (ma is None) == (mc is None)
        if ma is not None:
            assert torch.equal(ma, mc)
