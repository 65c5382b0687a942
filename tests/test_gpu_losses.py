"""Losses and the gradient-descent step on the GPU against the FP64 oracle.

* Spike-count loss (labels and explicit targets): loss, errors and prediction bit-exact (the
  loss of integer counts is exact in float32 at these sizes; the prediction is the first index
  of the largest count).
* TTFS softmax cross-entropy at the TTFS network's shape (batch 50, 10 outputs, K = 1,
  tau = 0.005) with all-silent rows, rows whose first spikes all underflow exp(-t/tau) in float32
  and ordinary rows: loss within 1e-5 relative (NaN where the oracle's is NaN), errors within
  1e-5 relative + 1e-6 max, NaN errors turned into 0, prediction exact; and the reference test's
  hand arrays with their closed form.
* Spike-time MSE loss: loss, errors (both the sanitised and the raw-time forms), prediction.
* Gradient descent: deltas equal the float32 product -lr * g; the fused weight update equals
  W + delta.
"""
import numpy as np
import pytest
import torch

from oracle import losses as oloss
from oracle import optim as oopt
from tests.helpers import assert_close_rel

pytestmark = pytest.mark.gpu
F32 = np.float32
F64 = np.float64


def _T(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def test_count_class_loss_bit_exact():
    from bats_amd.Losses import SpikeCountClassLoss
    rng = np.random.default_rng(31)
    B, O, K = 256, 10, 30
    n = rng.integers(0, K + 1, size=(B, O)).astype(np.int32)
    n[:8] = rng.integers(0, 2, size=(8, O))          # many tied maxima
    n[8] = 0                                         # a silent sample
    labels = rng.integers(0, O, size=B)
    t = np.full((B, O, K), np.inf, dtype=F32)
    loss_fct = SpikeCountClassLoss(target_true=15, target_false=3)
    loss, err = loss_fct.compute_loss_and_errors(_T(t), _T(n), _T(labels))
    pred = loss_fct.predict(_T(t), _T(n))
    tg = oloss.count_targets(n, labels, 15, 3)
    ref_loss, ref_err = oloss.count_loss_and_errors(n, tg, K)
    assert np.array_equal(loss.cpu().numpy(), ref_loss)
    assert np.array_equal(err.cpu().numpy(), ref_err)
    assert np.array_equal(pred.cpu().numpy(), oloss.count_predict(n))
    # the separate entry points give the same values
    assert np.array_equal(loss_fct.compute_loss(_T(t), _T(n), _T(labels)).cpu().numpy(), ref_loss)
    assert np.array_equal(loss_fct.compute_errors(_T(t), _T(n), _T(labels)).cpu().numpy(), ref_err)


def test_count_loss_explicit_targets():
    from bats_amd.Losses import SpikeCountLoss
    rng = np.random.default_rng(32)
    B, O, K = 64, 10, 7
    n = rng.integers(0, K + 1, size=(B, O)).astype(np.int32)
    tg = rng.uniform(0, 10, size=(B, O)).astype(F32)
    t = np.full((B, O, K), np.inf, dtype=F32)
    loss, err = SpikeCountLoss().compute_loss_and_errors(_T(t), _T(n), _T(tg))
    ref_loss, ref_err = oloss.count_loss_and_errors(n, tg, K)
    assert_close_rel(loss.cpu().numpy(), ref_loss, 1e-6, 0.0, "count loss (targets)")
    assert np.array_equal(err.cpu().numpy(), ref_err)


def _ttfs_times(rng, B, O):
    t = rng.uniform(0.0, 1.0, size=(B, O)).astype(F32)
    t[rng.uniform(size=(B, O)) < 0.3] = np.inf      # silent neurons
    t[0] = np.inf                                    # an all-silent sample
    t[1] = rng.uniform(0.5, 1.0, size=O)             # exp(-t/0.005) underflows for all
    t[2, :] = np.inf
    t[2, 3] = 0.01                                   # only one neuron fires
    t[3] = 0.2                                       # all tied
    t[4] = rng.uniform(0.0, 0.05, size=O)            # resolved softmax
    return t[:, :, None]


@pytest.mark.parametrize("tau", [0.005, 0.2])
def test_ttfs_loss_batch50(tau):
    from bats_amd.Losses import TTFSSoftmaxCrossEntropy
    rng = np.random.default_rng(33)
    B, O = 50, 10
    t = _ttfs_times(rng, B, O)
    labels = rng.integers(0, O, size=B)
    labels[2] = 7                                    # label on a silent neuron
    loss_fct = TTFSSoftmaxCrossEntropy(tau=tau)
    loss, err = loss_fct.compute_loss_and_errors(_T(t), None, _T(labels))
    pred = loss_fct.predict(_T(t), None)
    ref_loss, ref_err = oloss.ttfs_loss_and_errors(t, labels, tau)
    g = loss.cpu().numpy()
    assert np.array_equal(np.isnan(g), np.isnan(ref_loss)), (np.isnan(g), np.isnan(ref_loss))
    fin = np.isfinite(ref_loss)
    assert_close_rel(g[fin], ref_loss[fin], 1e-5, 0.0, "TTFS loss")
    assert np.array_equal(np.isinf(g), np.isinf(ref_loss))
    ge = err.cpu().numpy()
    assert np.all(np.isfinite(ge))
    assert_close_rel(ge, ref_err, 1e-5, 1e-6, "TTFS errors")
    assert np.all(ge[0] == 0)                        # a silent sample
    assert np.array_equal(pred.cpu().numpy(), oloss.ttfs_predict(t))
    if tau == 0.005:                                 # every exp(-t/tau) of sample 1 underflows
        assert np.all(ge[1] == 0)
        assert np.isnan(ref_loss[0]) and np.isnan(ref_loss[1])


def test_ttfs_loss_reference_hand_arrays():
    # the arrays of tests/Losses/TTFSSoftmaxCrossEntropy_tests.py:18-28 and their closed form
    from bats_amd.Losses import TTFSSoftmaxCrossEntropy
    tau = 0.2
    spk = np.array([[[0.1, 0.2, 0.3, np.inf], [0.2, np.inf, np.inf, np.inf], [np.inf] * 4,
                     [0.3, 0.4, 0.5, 0.6]],
                    [[0.4, 0.5, np.inf, np.inf], [1.0, np.inf, np.inf, np.inf],
                     [0.9, 1.1, np.inf, np.inf], [0.5, np.inf, np.inf, np.inf]]]).astype(F32)
    labels = np.array([3, 0])
    exps = np.exp(-spk[..., 0].astype(F64) / tau)
    soft = exps / exps.sum(axis=1, keepdims=True)
    true_loss = -np.log(soft[np.arange(2), labels])
    neg = -soft
    neg[np.arange(2), labels] += 1
    true_err = np.zeros(spk.shape)
    true_err[..., 0] = neg / tau
    loss, err = TTFSSoftmaxCrossEntropy(tau=tau).compute_loss_and_errors(_T(spk), None, _T(labels))
    np.testing.assert_allclose(loss.cpu().numpy(), true_loss, rtol=1e-6)
    np.testing.assert_allclose(err.cpu().numpy(), true_err, rtol=1e-5, atol=1e-6)


def test_spike_time_loss():
    from bats_amd.Losses import SpikeTimeLoss
    rng = np.random.default_rng(34)
    B, O, K = 40, 10, 3
    t = np.sort(rng.uniform(0, 1, size=(B, O, K)), axis=2).astype(F32)
    t[rng.uniform(size=(B, O, K)) < 0.3] = np.inf
    t = np.sort(t, axis=2)
    tg = rng.uniform(0, 1, size=(B, O, K)).astype(F32)
    tg[rng.uniform(size=(B, O, K)) < 0.2] = np.inf
    lf = SpikeTimeLoss()
    ref_loss, ref_err = oloss.spike_time_loss_and_errors(t, tg)
    loss, err = lf.compute_loss_and_errors(_T(t), None, _T(tg))
    assert_close_rel(loss.cpu().numpy(), ref_loss, 1e-6, 0.0, "spike-time loss")
    ge = err.cpu().numpy()
    assert np.array_equal(np.isinf(ge), np.isinf(ref_err))
    fin = np.isfinite(ref_err)
    assert np.array_equal(ge[fin], ref_err[fin])
    assert_close_rel(lf.compute_loss(_T(t), None, _T(tg)).cpu().numpy(), ref_loss, 1e-6, 0.0,
                     "spike-time loss (compute_loss)")
    # compute_errors uses the sanitised times (SpikeTimeLoss.py:30-35)
    tc = np.where(np.isinf(t), F32(1.0), t)
    want = (tc.astype(F64) - np.where(np.isinf(tg), 1.0, tg).astype(F64)).astype(F32)
    assert np.array_equal(lf.compute_errors(_T(t), None, _T(tg)).cpu().numpy(), want)
    assert np.array_equal(lf.predict(_T(t), None).cpu().numpy(), np.argmin(t[..., 0], axis=1))
    # like the reference, the caller's device targets have their +inf entries set to 1.0
    tgd = _T(tg)
    lf.compute_loss_and_errors(_T(t), None, tgd)
    assert not torch.isinf(tgd).any()
    assert np.array_equal(tgd.cpu().numpy(), np.where(np.isinf(tg), F32(1.0), tg))


def test_gradient_descent_step():
    from bats_amd import ops
    from bats_amd.Optimizers import GradientDescentOptimizer
    rng = np.random.default_rng(35)
    grads = [None, rng.normal(size=(800, 784)).astype(F32), rng.normal(size=(10, 800)).astype(F32)]
    opt = GradientDescentOptimizer(learning_rate=0.05)
    out = opt.step([None if g is None else _T(g) for g in grads])
    want = oopt.gd_step(grads, 0.05)
    assert out[0] is None
    for o, w in zip(out[1:], want[1:]):
        assert np.array_equal(o.cpu().numpy(), w)
    # fused form: delta written and added to the weights in one launch
    W = rng.normal(size=(10, 800)).astype(F32)
    Wd = _T(W)
    d = ops.sgd_step(_T(grads[2]), 0.05, weights=Wd)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), want[2])
    assert np.array_equal(Wd.cpu().numpy(), (W + want[2]).astype(F32))
