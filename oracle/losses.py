"""FP64 restatement of the reference losses.  Test infrastructure only."""
from __future__ import annotations

import numpy as np

F32 = np.float32
F64 = np.float64


def count_targets(n, labels, target_true, target_false):
    """bats/Losses/SpikeCountClassLoss.py:15-18."""
    n = np.asarray(n)
    tgt = np.full(n.shape, F32(target_false), dtype=F32)
    tgt[np.arange(len(labels)), np.asarray(labels)] = F32(target_true)
    return tgt


def count_loss_and_errors(n, targets, max_n_spike):
    """bats/Losses/SpikeCountLoss.py:31-38: loss_b = 0.5 * sum_j (target - n)^2,
    errors[b,j,k] = target - n for every slot k. Returns (loss [B] f32, errors [B,O,K] f32)."""
    n64 = np.asarray(n, dtype=F64)
    t64 = np.asarray(targets, dtype=F64)
    d = t64 - n64
    loss = (0.5 * (d * d).sum(axis=1)).astype(F32)
    err = np.repeat(d.astype(F32)[:, :, None], int(max_n_spike), axis=2)
    return loss, err


def count_predict(n):
    """bats/Losses/SpikeCountLoss.py:17-18 (first index of the maximum)."""
    return np.argmax(np.asarray(n), axis=1)


def ttfs_loss_and_errors(t, labels, tau):
    """bats/Losses/TTFSSoftmaxCrossEntropy.py:41-56 on the first spike time t[...,0]:
    exps = exp(-t0/tau) (f32 array), sums (f32 array), loss = -log(exps[label]/sums),
    errors[...,0] = nan_to_num((onehot - exps/sums)/tau), errors[...,1:] = 0."""
    t = np.asarray(t)
    B, O, K = t.shape
    labels = np.asarray(labels)
    tau64 = F64(F32(tau))
    with np.errstate(all="ignore"):
        exps = np.exp(-t[..., 0].astype(F64) / tau64).astype(F32)
        sums = exps.astype(F64).sum(axis=1).astype(F32)
        lab = exps[np.arange(B), labels]
        loss = (-np.log(lab.astype(F64) / sums.astype(F64))).astype(F32)
        neg = (-exps.astype(F64) / sums.astype(F64)[:, None]).astype(F32)
        neg = neg.astype(F64)
        neg[np.arange(B), labels] += 1.0
        neg = (neg.astype(F32).astype(F64) / tau64).astype(F32)
    err = np.zeros((B, O, K), dtype=F32)
    err[..., 0] = np.nan_to_num(neg)
    return loss, err


def ttfs_predict(t):
    """bats/Losses/TTFSSoftmaxCrossEntropy.py:20-21 (argmin over neurons, flattened)."""
    return np.argmin(np.asarray(t), axis=1).flatten()


def spike_time_loss_and_errors(t, targets):
    """bats/Losses/SpikeTimeLoss.py:38-47: infinite spike times and targets become 1.0 for
    the loss; the errors use the un-sanitised times (t - targets)."""
    t = np.asarray(t, dtype=F32)
    tg = np.array(targets, dtype=F32, copy=True)
    tg[np.isinf(tg)] = F32(1.0)
    tc = t.copy()
    tc[np.isinf(t)] = F32(1.0)
    err = (t.astype(F64) - tg.astype(F64)).astype(F32)
    d = tg[..., 0].astype(F64) - tc[..., 0].astype(F64)
    loss = (0.5 * (d * d).sum(axis=1)).astype(F32)
    return loss, err
