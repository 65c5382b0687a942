"""The deep dense network (784-800-800-800-10, K = 30 everywhere, thresholds 1.0 / 1.3,
W ~ N(0, 1), latency-coded inputs, spike-count loss with targets 15 / 3) end to end against
the FP64 oracle at batch 64, layer by layer, through the Network API and through chained
autograd Functions.

Each layer's backward receives the errors the GPU propagated from the layer above (not the
oracle's), so the comparison covers the error carried through the whole chain. Bar, for every
layer: spike counts identical; spike times within 1e-5 relative; weight gradients within
1e-4 |ref| + 1e-6 max|ref| elementwise. The per-layer errors are printed."""
import numpy as np
import pytest
import torch

from oracle import lif as olif
from oracle import losses as oloss
from tests.helpers import assert_close_rel, latency_inputs, spike_slots

pytestmark = pytest.mark.gpu
F32 = np.float32
SIZES = [784, 800, 800, 800, 10]
THETAS = [1.0, 1.0, 1.0, 1.3]
TAU_S, K, T_MAX, B = 0.130, 30, 1.0, 64


def _problem():
    rng = np.random.default_rng(21)
    t, n = latency_inputs(rng, B, SIZES[0])
    Ws = [rng.normal(0.0, 1.0, size=(SIZES[k + 1], SIZES[k])).astype(F32) for k in range(4)]
    labels = rng.integers(0, 10, size=B)
    return t, n, Ws, labels


_CACHE = {}


def _reference():
    """The problem and its oracle forward, computed once for both tests (about 10 s)."""
    if not _CACHE:
        t, n, Ws, labels = _problem()
        _CACHE["v"] = (t, n, Ws, labels) + _oracle(t, Ws, labels)
    return _CACHE["v"]


def _oracle(t, Ws, labels):
    states, cur = [], t
    for W, th in zip(Ws, THETAS):
        s = olif.lif_forward(cur, W, TAU_S, th, th, K, tmax=T_MAX)
        states.append(s)
        cur = s.t
    tg = oloss.count_targets(states[-1].n, labels, 15, 3)
    _, err = oloss.count_loss_and_errors(states[-1].n, tg, K)
    return states, err


def _time_err(got, ref):
    m = spike_slots(ref.n, K)
    g, r = got[m].astype(np.float64), ref.t[m].astype(np.float64)
    return float(np.max(np.abs(g - r) / np.abs(r))) if r.size else 0.0


def _grad_report(g, r):
    scale = float(np.max(np.abs(r)))
    excess = np.abs(g - r) / (1e-4 * np.abs(r) + 1e-6 * scale)
    return float(np.max(np.abs(g - r)) / scale), float(excess.max())


def test_deep_network_per_layer_against_oracle():
    from bats_amd import Network
    from bats_amd.Layers import InputLayer, LIFLayer
    from bats_amd.Losses import SpikeCountClassLoss

    t, n, Ws, labels, states, err_out = _reference()
    # oracle backward, errors carried through the chain in the oracle itself
    ref_grads, err = [None] * 4, err_out
    for k in range(3, -1, -1):
        _, g, pe = olif.lif_backward(states[k], err, pre_trainable=k > 0)
        ref_grads[k] = olif.batch_mean(g)
        err = pe

    net = Network()
    inp = InputLayer(n_neurons=SIZES[0], name="input")
    net.add_layer(inp, input=True)
    prev, lays = inp, []
    for k, (W, th) in enumerate(zip(Ws, THETAS)):
        lay = LIFLayer(previous_layer=prev, n_neurons=SIZES[k + 1], tau_s=TAU_S, theta=th,
                       delta_theta=th, weight_initializer=lambda a, b, W=W: W,
                       max_n_spike=K, name=f"l{k}")
        net.add_layer(lay)
        lays.append(lay)
        prev = lay
    net.forward(t, n, max_simulation=T_MAX, training=True)
    ot, on = net.output_spike_trains
    _, errors = SpikeCountClassLoss(target_true=15, target_false=3).compute_loss_and_errors(
        ot, on, labels)
    grads = net.backward(errors)
    torch.cuda.synchronize()
    print()
    for k, lay in enumerate(lays):
        tk, nk = (x.cpu().numpy() for x in lay.spike_trains)
        assert np.array_equal(nk, states[k].n), f"layer {k}: spike counts differ"
        te = _time_err(tk, states[k])
        g = grads[k + 1][0].cpu().numpy().astype(np.float64)
        r = ref_grads[k].astype(np.float64)
        rel, excess = _grad_report(g, r)
        print(f"[deep] layer {k} ({SIZES[k]}->{SIZES[k + 1]}): spikes/sample "
              f"{states[k].n.sum() / B:.1f}, time rel err {te:.2e}, grad max|d|/max|ref| "
              f"{rel:.2e}, worst elementwise use of the bar {excess:.3f}", flush=True)
        assert te <= 1e-5, f"layer {k}: spike time error {te:.3e}"
        assert_close_rel(g, r, 1e-4, 1e-6, f"layer {k} weight gradient")


def test_deep_network_autograd_chain():
    from bats_amd.autograd import lif

    t, _, Ws, labels, states, err_out = _reference()
    ref_grads, err = [None] * 4, err_out
    for k in range(3, -1, -1):
        _, g, pe = olif.lif_backward(states[k], err, pre_trainable=k > 0)
        ref_grads[k] = g.astype(np.float64).sum(axis=0)  # autograd: gradient of a summed loss
        err = pe
    Wd = [torch.as_tensor(W, device="cuda").requires_grad_(True) for W in Ws]
    x = torch.as_tensor(t, device="cuda")
    outs = []
    for k, (W, th) in enumerate(zip(Wd, THETAS)):
        # the widest layers keep only their inputs and rerun the forward in the backward
        x, nk = lif(x, W, tau_s=TAU_S, theta=th, delta_theta=th, max_n_spike=K, t_max=T_MAX,
                    recompute=(k in (1, 2)))
        outs.append(nk)
    e = torch.as_tensor(err_out, device="cuda")
    L = torch.where(torch.isfinite(x), x * e, torch.zeros_like(x)).sum()
    L.backward()
    torch.cuda.synchronize()
    for k in range(4):
        assert np.array_equal(outs[k].cpu().numpy(), states[k].n), f"layer {k}: counts differ"
        g = Wd[k].grad.cpu().numpy().astype(np.float64)
        assert_close_rel(g, ref_grads[k], 1e-4, 1e-6, f"autograd layer {k} weight gradient")
