"""The deep dense network (784-800-800-800-10, K = 30 everywhere, thresholds 1.0 / 1.3,
W ~ N(0, 1), latency-coded inputs, spike-count loss with targets 15 / 3) end to end against
the FP64 oracle at batch 64, layer by layer, through the Network API and through chained
autograd Functions.

Each layer's backward receives the errors the GPU propagated from the layer above (not the
oracle's), so the comparison covers the error carried through the whole chain. Bar, for every
layer: spike counts identical; spike times within 1e-5 relative; weight gradients within
1e-4 |ref| + 1e-6 max|ref| elementwise. The per-layer errors are printed. A third test checks
the chained autograd gradients against central finite differences on 64 weights of the four
layers."""
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


def test_deep_network_finite_differences():
    """Central finite differences on 64 random weights of the deep network (16 per layer, drawn
    among the weights whose gradient is at least 1 % of their layer's largest,
    h = 1e-4 (|w| + 0.1)), through the chained autograd Functions, of a smooth loss of every
    layer's spike times: L = sum over layers and slots of e * t with fixed random e (float64
    sum of the float32 times), batch 8.

    A stencil is compared when every layer's spike counts are those of the unperturbed
    network on both sides, some float32 spike time moved (where none does, g h is below the
    times' resolution), and the two one-sided slopes agree (a spike that moves across an
    input spike time anywhere downstream makes the loss only piecewise smooth; the central
    difference then averages two slopes). The bar is the single-layer test's: 2e-2 |g| plus
    the rounding the float32 times carry into the difference, sum |e| ulp(t) / (2h) over the
    spikes whose times differ between the two sides."""
    from bats_amd.autograd import lif

    rng = np.random.default_rng(33)
    Bf = 8
    t, _ = latency_inputs(rng, Bf, SIZES[0])
    Ws = [rng.normal(0.0, 1.0, size=(SIZES[k + 1], SIZES[k])).astype(F32) for k in range(4)]
    es = [rng.normal(size=(Bf, SIZES[k + 1], K)) for k in range(4)]
    x0 = torch.as_tensor(t, device="cuda")
    e_dev = [torch.as_tensor(e, device="cuda") for e in es]

    def run(Wlist, grad=False):
        Wd = [torch.as_tensor(W, device="cuda").requires_grad_(grad) for W in Wlist]
        x, L, times, counts = x0, 0.0, [], []
        for k, (W, th) in enumerate(zip(Wd, THETAS)):
            x, nk = lif(x, W, tau_s=TAU_S, theta=th, delta_theta=th, max_n_spike=K,
                        t_max=T_MAX)
            L = L + torch.where(torch.isfinite(x), x.double() * e_dev[k],
                                torch.zeros_like(x, dtype=torch.float64)).sum()
            times.append(x.detach().cpu().numpy())
            counts.append(nk.cpu().numpy())
        if grad:
            L.backward()
            return float(L.item()), times, counts, [W.grad.cpu().numpy() for W in Wd]
        return float(L.item()), times, counts

    L0, t0, n0, grads = run(Ws, grad=True)
    assert all(c.sum() > 0 for c in n0)
    # 16 random weights per layer among those whose gradient is at least 1 % of the layer's
    # largest (a finite difference with h = 1e-4 cannot resolve much smaller slopes through
    # float32 spike times)
    picks = []
    for k in range(4):
        ok = np.argwhere(np.abs(grads[k]) >= 1e-2 * np.abs(grads[k]).max())
        for r in rng.choice(len(ok), size=16, replace=False):
            picks.append((k, int(ok[r][0]), int(ok[r][1])))
    checked = [0] * 4
    skipped_counts = skipped_kinks = skipped_flat = 0
    for k, j, i in picks:
        W = Ws[k]
        h = np.float32(1e-4 * (abs(W[j, i]) + 0.1))
        Wp = [w.copy() for w in Ws]
        Wm = [w.copy() for w in Ws]
        Wp[k][j, i] += h
        Wm[k][j, i] -= h
        Lp, tp, np_ = run(Wp)
        Lm, tm, nm = run(Wm)
        if not all(np.array_equal(a, b) and np.array_equal(c, b)
                   for a, b, c in zip(np_, n0, nm)):
            skipped_counts += 1
            continue
        hp = float(Wp[k][j, i]) - float(W[j, i])
        hm = float(W[j, i]) - float(Wm[k][j, i])
        dw = hp + hm
        noise, moved = 0.0, 0
        for a, b, e in zip(tp, tm, es):
            ch = np.isfinite(a) & (a != b)
            moved += int(ch.sum())
            noise += float(np.sum(np.abs(e[ch]) * np.spacing(np.abs(a[ch]).astype(F32))))
        noise /= dw
        if moved == 0:
            # no float32 spike time changed: g * 2h is below the times' resolution here
            skipped_flat += 1
            continue
        sp, sm = (Lp - L0) / hp, (L0 - Lm) / hm
        g = float(grads[k][j, i])
        if abs(sp - sm) > 2e-2 * max(abs(sp), abs(sm)) + 4 * noise:
            skipped_kinks += 1
            continue
        fd = (Lp - Lm) / dw
        assert abs(fd - g) <= 2e-2 * abs(g) + noise, (k, j, i, fd, g, noise)
        checked[k] += 1
    print(f"[deep fd] checked per layer {checked}, skipped: counts changed {skipped_counts}, "
          f"kinks {skipped_kinks}, no time moved {skipped_flat}", flush=True)
    assert sum(checked) >= 48 and min(checked) >= 8
