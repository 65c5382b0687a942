"""Monitors (bats/Monitors/*.py): device-accumulated values equal the reference's host
formulas on the same batches (CPU tensors here; the same code runs on device tensors)."""
import numpy as np
import torch

from bats_amd.Monitors import (AccuracyMonitor, LossMonitor, MonitorsManager, SilentLabelsMonitor,
                               SilentNeuronsMonitor, SpikeCountMonitor, ValueMonitor,
                               WeightsNormMonitor)


def batches(seed, n=3, B=17, O=10):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        counts = rng.integers(0, 4, size=(B, O)).astype(np.int32)
        counts[:, 3] = 0  # a neuron that never fires
        labels = rng.integers(0, O, size=B)
        losses = rng.normal(size=B).astype(np.float32)
        losses[0] = np.inf
        losses[1] = np.nan
        yield counts, labels, losses


def test_accuracy_loss_silent_monitors():
    acc, loss, sl, sn, sc = (AccuracyMonitor(), LossMonitor(), SilentLabelsMonitor(),
                             SilentNeuronsMonitor("out"), SpikeCountMonitor("out"))
    hits = n = 0
    lsum = lcnt = 0.0
    silent = 0
    tot = None
    spikes = 0
    for counts, labels, losses in batches(0):
        pred = counts.argmax(axis=1)
        acc.add(torch.as_tensor(pred), torch.as_tensor(labels))
        loss.add(torch.as_tensor(losses))
        sl.add(torch.as_tensor(counts), torch.as_tensor(labels))
        sn.add(torch.as_tensor(counts))
        sc.add(torch.as_tensor(counts))
        # the reference's formulas
        hits += np.sum(pred == labels)
        n += labels.shape[0]
        ok = np.isfinite(losses)
        lsum += np.sum(losses[ok].astype(np.float64))
        lcnt += np.sum(ok)
        silent += np.sum(np.take_along_axis(counts, labels[:, None], axis=1) == 0)
        tot = counts.sum(axis=0) if tot is None else tot + counts.sum(axis=0)
        spikes += counts.sum()
    mgr = MonitorsManager([acc, loss, sl, sn, sc])
    got = mgr.record(1.0)
    assert got[acc] == hits / n * 100
    np.testing.assert_allclose(got[loss], lsum / lcnt, rtol=1e-12)
    assert got[sl] == silent / n * 100
    assert got[sn] == np.mean(tot == 0) * 100
    assert got[sn] >= 10.0
    assert got[sc] == spikes / n
    # record() resets the running sums
    for counts, labels, losses in batches(1, n=1):
        acc.add(torch.as_tensor(counts.argmax(1)), torch.as_tensor(labels))
    assert acc.record(2.0) == np.mean(counts.argmax(1) == labels) * 100
    assert acc.epochs == [1.0, 2.0]


def test_weights_norm_and_value_monitors(tmp_path):
    w = np.random.default_rng(2).normal(size=(10, 800)).astype(np.float32)
    m = WeightsNormMonitor("hidden", export_path=tmp_path / "norm")
    m.add(torch.as_tensor(w))
    np.testing.assert_allclose(m.record(0.5), np.linalg.norm(w.astype(np.float64)), rtol=1e-12)
    v = ValueMonitor(name="lr")
    v.add(torch.tensor(0.25))
    assert v.record(1) == 0.25
    m.export()
    data = np.load(tmp_path / "norm.npz")
    assert list(data["epochs"]) == [0.5]
