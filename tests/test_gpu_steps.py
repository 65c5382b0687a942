"""Whole training steps on the GPU against the FP64 oracle, at the workloads' own shapes.

* TTFS network (784-800-10, K = 1, batch 50, thresholds 1.0 / 1.3, W ~ N(0,1), softmax
  cross-entropy on first spike times with tau = 0.005, Adam), two consecutive steps. Most
  output rows are silent or have first spikes whose exp(-t/tau) underflows, so the NaN loss ->
  zero error path runs.
* Spike-count network (784-800-10, K = 30/30, batch 256, targets 15 / 3, Adam), one step at the
  headline batch (the batch-slice weight-gradient reduction runs).
* Convolutional network at its own settings (28x28 input, conv 5x5x15 K 30 -> pool -> conv
  5x5x40 K 3 -> pool -> 300 K 10 -> 10 K 30, thresholds 0.04 / 0.8 / 0.6 / 0.3, weight gains
  1 / 8 / 8 / 8), batch 4: layer by layer in both directions, including the pooled 120-slot
  input of the second conv layer.
* The reference's end-to-end layer test (tests/Layers/LIFLayer_tests.py:91-132) on the GPU.

Each step's oracle runs on the weights the GPU holds at the start of that step. The Adam step
is checked against the oracle's Adam applied to the GPU's gradients (deltas, m and v within
1e-6 relative), so every stage is compared on identical inputs. Bars: spike counts identical;
spike times within 1e-5 relative; losses within 1e-5 relative (NaN where the oracle's is NaN);
errors, gradients and pre-spike errors within 1e-4 |ref| + 1e-6 max|ref|; weights after the
update equal to the float32 sum of weights and deltas.
"""
import numpy as np
import pytest
import torch

from oracle import conv as oconv
from oracle import lif as olif
from oracle import losses as oloss
from oracle import optim as oopt
from tests.helpers import assert_close_rel, latency_inputs, spike_slots

pytestmark = pytest.mark.gpu
F32 = np.float32
TAU_S, T_MAX, LR = 0.130, 1.0, 0.003


def _np(x):
    return x.detach().cpu().numpy()


def _check_forward(layer, ref, what):
    t, n = (_np(v) for v in layer.spike_trains)
    assert np.array_equal(n, ref.n), f"{what}: spike counts differ"
    m = spike_slots(ref.n, t.shape[2])
    if m.any():
        rel = np.abs(t[m].astype(np.float64) - ref.t[m]) / np.abs(ref.t[m].astype(np.float64))
        assert rel.max() <= 1e-5, f"{what}: spike time rel err {rel.max():.3e}"
    assert np.all(np.isinf(t[~m])), f"{what}: empty slots must be +inf"


def _dense_net(sizes, Ks, thetas, Ws):
    from bats_amd import Network
    from bats_amd.Layers import InputLayer, LIFLayer
    net = Network()
    inp = InputLayer(n_neurons=sizes[0], name="input")
    net.add_layer(inp, input=True)
    prev = inp
    for k, (K, th, W) in enumerate(zip(Ks, thetas, Ws)):
        lay = LIFLayer(previous_layer=prev, n_neurons=sizes[k + 1], tau_s=TAU_S, theta=th,
                       delta_theta=th, weight_initializer=lambda a, b, W=W: W,
                       max_n_spike=K, name=f"l{k}")
        net.add_layer(lay)
        prev = lay
    return net


def _check_adam(opt, ref_opt, grads, deltas):
    """The GPU Adam against the oracle's Adam fed the same (GPU) gradients."""
    g_np = [None if g is None else _np(g[0]) for g in grads]
    want = ref_opt.step(g_np)
    for k, (d, w) in enumerate(zip(deltas, want)):
        if w is None:
            assert d is None
            continue
        assert_close_rel(_np(d), w, 1e-6, 0.0, f"Adam delta {k}")
        assert_close_rel(_np(opt._m[k]), ref_opt.m[k], 1e-6, 0.0, f"Adam m {k}")
        assert_close_rel(_np(opt._v[k]), ref_opt.v[k], 1e-6, 0.0, f"Adam v {k}")


def _dense_step(net, Ks, thetas, t, n, labels, loss_kind, opt, ref_opt):
    """One GPU training step through the Network API, each stage against the oracle."""
    lays = net.layers[1:]
    Ws = [_np(l.weights) for l in lays]
    # oracle forward on the GPU's current weights
    states, cur = [], t
    for W, K, th in zip(Ws, Ks, thetas):
        s = olif.lif_forward(cur, W, TAU_S, th, th, K, tmax=T_MAX)
        states.append(s)
        cur = s.t
    if loss_kind == "ttfs":
        from bats_amd.Losses import TTFSSoftmaxCrossEntropy
        loss_fct = TTFSSoftmaxCrossEntropy(tau=0.005)
    else:
        from bats_amd.Losses import SpikeCountClassLoss
        loss_fct = SpikeCountClassLoss(target_true=15, target_false=3)
    # GPU step
    net.reset()
    net.forward(t, n, max_simulation=T_MAX, training=True)
    ot, on = net.output_spike_trains
    lab = torch.as_tensor(labels, device="cuda")
    loss, err = loss_fct.compute_loss_and_errors(ot, on, lab)
    pred = loss_fct.predict(ot, on)
    err_before = _np(err).copy()
    grads = net.backward(err)
    deltas = opt.step([None if g is None else g[0] for g in grads])
    net.apply_deltas(deltas)
    torch.cuda.synchronize()
    for k, (lay, s) in enumerate(zip(lays, states)):
        _check_forward(lay, s, f"layer {k}")
    # the loss stage on the GPU's own output trains (the TTFS loss scales time differences by
    # 1/tau = 200, so it is compared on identical inputs)
    ot_np, on_np = _np(ot), _np(on)
    if loss_kind == "ttfs":
        ref_loss, ref_err = oloss.ttfs_loss_and_errors(ot_np, labels, 0.005)
        ref_pred = oloss.ttfs_predict(ot_np)
    else:
        tg = oloss.count_targets(on_np, labels, 15, 3)
        ref_loss, ref_err = oloss.count_loss_and_errors(on_np, tg, Ks[-1])
        ref_pred = oloss.count_predict(on_np)
    gl = _np(loss)
    assert np.array_equal(np.isnan(gl), np.isnan(ref_loss))
    fin = np.isfinite(ref_loss)
    assert_close_rel(gl[fin], ref_loss[fin], 1e-5, 0.0, "loss")
    assert_close_rel(err_before, ref_err, 1e-5, 1e-6, "loss errors")
    assert np.array_equal(_np(pred), ref_pred)
    # oracle backward, its own errors carried down the chain
    ref_grads, e = [None] * len(lays), ref_err
    for k in range(len(lays) - 1, -1, -1):
        e_chain, g, pe = olif.lif_backward(states[k], e, pre_trainable=k > 0)
        ref_grads[k] = olif.batch_mean(g)
        if k == len(lays) - 1:
            assert_close_rel(_np(err), e_chain, 1e-4, 1e-6, "reset-chain errors (in place)")
        e = pe
    assert grads[0] is None
    for k in range(len(lays)):
        assert_close_rel(_np(grads[k + 1][0]), ref_grads[k], 1e-4, 1e-6, f"layer {k} gradient")
    _check_adam(opt, ref_opt, grads, [None] + deltas[1:])
    for k, lay in enumerate(lays):
        d = _np(deltas[k + 1])
        assert np.array_equal(_np(lay.weights), (Ws[k] + d.reshape(Ws[k].shape)).astype(F32))
    return states


def test_ttfs_network_two_steps_batch50():
    from bats_amd.Optimizers import AdamOptimizer
    rng = np.random.default_rng(41)
    B = 50
    t, n = latency_inputs(rng, B, 784)
    Ws = [rng.normal(size=(800, 784)).astype(F32), rng.normal(size=(10, 800)).astype(F32)]
    Ks, thetas = [1, 1], [1.0, 1.3]
    net = _dense_net([784, 800, 10], Ks, thetas, Ws)
    opt, ref_opt = AdamOptimizer(learning_rate=LR), oopt.Adam(LR)
    for step in range(2):
        labels = rng.integers(0, 10, size=B)
        states = _dense_step(net, Ks, thetas, t, n, labels, "ttfs", opt, ref_opt)
        exps = np.exp(-states[-1].t[:, :, 0].astype(np.float64) / 0.005).astype(F32)
        silent = exps.sum(axis=1) == 0
        print(f"[ttfs] step {step}: hidden spikes/sample {states[0].n.sum() / B:.1f}, "
              f"output {states[1].n.sum() / B:.2f}, samples with an all-zero softmax "
              f"{int(silent.sum())}", flush=True)
        assert states[0].n.sum() > 0 and states[1].n.sum() > 0
        assert silent.any()


def test_count_network_step_batch256():
    from bats_amd.Optimizers import AdamOptimizer
    rng = np.random.default_rng(42)
    B = 256
    t, n = latency_inputs(rng, B, 784)
    Ws = [rng.normal(size=(800, 784)).astype(F32), rng.normal(size=(10, 800)).astype(F32)]
    Ks, thetas = [30, 30], [1.0, 1.3]
    net = _dense_net([784, 800, 10], Ks, thetas, Ws)
    labels = rng.integers(0, 10, size=B)
    states = _dense_step(net, Ks, thetas, t, n, labels, "count", AdamOptimizer(learning_rate=LR),
                         oopt.Adam(LR))
    print(f"[count] hidden spikes/sample {states[0].n.sum() / B:.1f}, output "
          f"{states[1].n.sum() / B:.2f}", flush=True)
    assert states[0].n.sum() / B > 500


def test_conv_network_step_batch4():
    from bats_amd import Network
    from bats_amd.Layers import ConvInputLayer, ConvLIFLayer, LIFLayer, PoolingLayer
    from bats_amd.Losses import SpikeCountClassLoss
    from bats_amd.Optimizers import AdamOptimizer
    rng = np.random.default_rng(43)
    B = 4
    t, n_in = latency_inputs(rng, B, 784)
    w1 = rng.normal(0.0, 1.0 / 5.0, size=(15, 5, 5, 1)).astype(F32)
    w2 = (8.0 * rng.normal(0.0, 1.0 / np.sqrt(375), size=(40, 5, 5, 15))).astype(F32)
    w3 = (8.0 * rng.normal(0.0, 1.0 / np.sqrt(640), size=(300, 640))).astype(F32)
    w4 = (8.0 * rng.normal(0.0, 1.0 / np.sqrt(300), size=(10, 300))).astype(F32)
    labels = rng.integers(0, 10, size=B)
    net = Network()
    inp = ConvInputLayer(neurons_shape=np.array([28, 28, 1]), name="input")
    net.add_layer(inp, input=True)
    c1 = ConvLIFLayer(inp, filters_shape=np.array([5, 5, 15]), tau_s=TAU_S, theta=0.04,
                      delta_theta=0.04, weight_initializer=lambda *a: w1, max_n_spike=30,
                      name="conv1")
    p1 = PoolingLayer(c1, name="pool1")
    c2 = ConvLIFLayer(p1, filters_shape=np.array([5, 5, 40]), tau_s=TAU_S, theta=0.8,
                      delta_theta=0.8, weight_initializer=lambda *a: w2, max_n_spike=3,
                      name="conv2")
    p2 = PoolingLayer(c2, name="pool2")
    d3 = LIFLayer(p2, n_neurons=300, tau_s=TAU_S, theta=0.6, delta_theta=0.6,
                  weight_initializer=lambda *a: w3, max_n_spike=10, name="fc")
    d4 = LIFLayer(d3, n_neurons=10, tau_s=TAU_S, theta=0.3, delta_theta=0.3,
                  weight_initializer=lambda *a: w4, max_n_spike=30, name="output")
    for lay in (c1, p1, c2, p2, d3, d4):
        net.add_layer(lay)

    # oracle forward
    r1 = oconv.conv_forward(t.reshape(B, 784, 1), w1, (28, 28, 1), TAU_S, 0.04, 0.04, 30,
                            tmax=T_MAX)
    _, pt1, src1, shp1 = oconv.pool_forward(r1.n, r1.t, (24, 24, 15))
    r2 = oconv.conv_forward(pt1, w2, shp1, TAU_S, 0.8, 0.8, 3, tmax=T_MAX)
    _, pt2, src2, _ = oconv.pool_forward(r2.n, r2.t, (8, 8, 40))
    r3 = olif.lif_forward(pt2, w3, TAU_S, 0.6, 0.6, 10, tmax=T_MAX)
    r4 = olif.lif_forward(r3.t, w4, TAU_S, 0.3, 0.3, 30, tmax=T_MAX)
    tg = oloss.count_targets(r4.n, labels, 15, 3)
    ref_loss, ref_err = oloss.count_loss_and_errors(r4.n, tg, 30)

    net.forward(t.reshape(B, 784, 1), n_in, max_simulation=T_MAX, training=True)
    torch.cuda.synchronize()
    _check_forward(c1, r1, "conv1")
    assert np.array_equal(_np(p1.spike_trains[0]), pt1), "pool1 times"
    assert pt1.shape[2] == 120
    _check_forward(c2, r2, "conv2")
    assert np.array_equal(_np(p2.spike_trains[0]), pt2), "pool2 times"
    _check_forward(d3, r3, "fc")
    _check_forward(d4, r4, "output")
    spikes = [int(r.n.sum()) // B for r in (r1, r2, r3, r4)]
    print(f"[conv] spikes/sample conv1 {spikes[0]}, conv2 {spikes[1]}, fc {spikes[2]}, "
          f"output {spikes[3]}", flush=True)
    assert min(spikes) > 0

    ot, on = net.output_spike_trains
    loss, err = SpikeCountClassLoss(target_true=15, target_false=3).compute_loss_and_errors(
        ot, on, torch.as_tensor(labels, device="cuda"))
    assert np.array_equal(_np(loss), ref_loss)
    assert np.array_equal(_np(err), ref_err)

    # the reverse loop of Network.backward, keeping a copy of every layer's outgoing errors
    # (the layer below runs its reset chain on them in place, as the reference does)
    gpu_grads, gpu_errs, e = {}, {}, err
    for lay in reversed(net.layers[1:]):
        g, e = lay.backward(e)
        gpu_grads[lay.name] = g
        gpu_errs[lay.name] = None if e is None else e.clone()
    torch.cuda.synchronize()
    assert gpu_errs["conv1"] is None
    _, g4, pe4 = olif.lif_backward(r4, ref_err)
    _, g3, pe3 = olif.lif_backward(r3, pe4)
    pe_p2 = oconv.pool_backward(pe3, src2, 2560, 3)
    _, g2, pe2 = oconv.conv_backward(r2, pe_p2, pre_trainable=True)
    pe_p1 = oconv.pool_backward(pe2, src1, 8640, 30)
    _, g1, _ = oconv.conv_backward(r1, pe_p1, pre_trainable=False)
    for name, ref in (("output", pe4), ("fc", pe3), ("pool2", pe_p2), ("conv2", pe2),
                      ("pool1", pe_p1)):
        got = _np(gpu_errs[name])
        assert got.shape == ref.shape, (name, got.shape, ref.shape)
        assert_close_rel(got, ref, 1e-4, 1e-6, f"{name} outgoing errors")
    for name, ref in (("output", g4), ("fc", g3), ("conv2", g2), ("conv1", g1)):
        got = _np(gpu_grads[name][0])
        assert_close_rel(got, olif.batch_mean(ref), 1e-4, 1e-6, f"{name} gradient")
    grads = [None] + [gpu_grads[l.name] for l in net.layers[1:]]
    assert grads[2] is None and grads[4] is None      # the pooling layers have no weights
    opt, ref_opt = AdamOptimizer(learning_rate=LR), oopt.Adam(LR)
    W_before = [None if l.weights is None else _np(l.weights) for l in net.layers]
    deltas = opt.step([None if g is None else g[0] for g in grads])
    net.apply_deltas(deltas)
    torch.cuda.synchronize()
    _check_adam(opt, ref_opt, grads, deltas)
    for lay, W, d in zip(net.layers, W_before, deltas):
        if d is not None:
            assert np.array_equal(_np(lay.weights), (W + _np(d).reshape(W.shape)).astype(F32))


def test_lif_layer_descends_to_target_times():
    """tests/Layers/LIFLayer_tests.py:91-132 on the GPU: 7 inputs -> 5 LIF neurons, all weights
    1, errors = first spike time - target, gradient descent at lr 3e3 on the batch-mean gradient
    for 200 epochs; the first spike times reach the targets within 5e-4. The first epoch's
    gradient is also compared with the oracle's."""
    from bats_amd.Layers import InputLayer, LIFLayer
    tau_s, theta, K, inf = 10e-3, 0.01, 8, np.inf
    t_in = np.array([[[0.0, inf], [3e-3, inf], [6e-3, inf], [8e-3, inf], [9e-3, inf],
                      [11e-3, inf], [17e-3, inf]]], dtype=F32)
    n_in = np.array([[1, 1, 1, 1, 1, 1, 1]], dtype=np.int32)
    targets = torch.tensor([[4e-3, 6e-3, 8e-3, 12e-3, 16e-3]], dtype=torch.float32, device="cuda")
    inp = InputLayer(n_neurons=7)
    hid = LIFLayer(previous_layer=inp, n_neurons=5, tau_s=tau_s, theta=theta, delta_theta=theta,
                   weight_initializer=lambda a, b: torch.full((a, b), 1.0, device="cuda"),
                   max_n_spike=K)
    inp.set_spike_trains(t_in, n_in)
    for epoch in range(200):
        hid.reset()
        hid.forward(F32(0.02))
        spk, _ = hid.spike_trains
        errors = torch.zeros_like(spk)
        errors[:, :, 0] = spk[:, :, 0] - targets
        if epoch == 0:
            st = olif.lif_forward(t_in, np.full((5, 7), 1.0, F32), tau_s, theta, theta, K,
                                  tmax=0.02)
            e0 = _np(errors)
        grad, pre = hid.backward(errors)
        assert pre is None
        if epoch == 0:
            _, g, _ = olif.lif_backward(st, e0, pre_trainable=False)
            assert_close_rel(_np(grad[0]), olif.batch_mean(g), 1e-4, 1e-6, "first epoch gradient")
        hid.add_deltas(-3e3 * grad.mean(dim=0))
    got = _np(spk[:, :, 0])
    assert np.allclose(got, _np(targets), atol=5e-4), got
