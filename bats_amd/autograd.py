"""torch.autograd integration of the dense LIF layer and the reference's losses.

``LIFFunction`` runs the HIP forward in ``forward`` and the BATS backward in ``backward``:
the gradient flowing into the spike times is the reference's ``errors`` (dL/dt per spike
slot), the weight gradient returned is the batch sum (the gradient of a summed loss), and the
gradient of the input spike times is the reference's pre-spike errors. With
``recompute=True`` the forward keeps only its inputs and the backward re-runs the forward to
regenerate a, x, z and the exps (deterministic, so the values are identical), for layers whose
spike buffers would not fit in memory twice.

``spike_count_loss`` is the reference's spike-count surrogate as a Function: its backward
emits ``target - n`` for every slot of every output neuron (bats/Losses/SpikeCountLoss.py:25-29),
which is a surrogate and not a true derivative of the count.
"""
from __future__ import annotations

import numpy as np
import torch

from . import ops
from ._constants import lif_constants


class LIFFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pre_times, weights, tau_s, theta, delta_theta, max_n_spike, t_max,
                recompute=False):
        consts = lif_constants(tau_s, theta, delta_theta)
        pre = pre_times.detach().contiguous()
        w = weights.detach().contiguous()
        n, t, a, x, z, es, e = ops.lif_forward(pre, w, *consts, t_max, int(max_n_spike))
        ctx.consts = consts
        ctx.K = int(max_n_spike)
        ctx.t_max = t_max
        ctx.recompute = bool(recompute)
        if recompute:
            ctx.save_for_backward(pre, w)
        else:
            ctx.save_for_backward(pre, w, n, t, a, x, z, es, e)
        ctx.mark_non_differentiable(n)
        return t, n

    @staticmethod
    def backward(ctx, grad_t, grad_n):
        if ctx.recompute:
            pre, w = ctx.saved_tensors
            n, t, a, x, z, es, e = ops.lif_forward(pre, w, *ctx.consts, ctx.t_max, ctx.K)
        else:
            pre, w, n, t, a, x, z, es, e = ctx.saved_tensors
        if grad_t is None:
            grad_t = torch.zeros_like(t)
        errors = grad_t.to(torch.float32).contiguous().clone()
        errors = torch.nan_to_num(errors, nan=0.0, posinf=0.0, neginf=0.0)
        need_pre = ctx.needs_input_grad[0]
        grad_w, grad_pre = ops.lif_backward(n, t, a, x, z, errors, pre, es, e, w, *ctx.consts,
                                            per_sample=False, pre_trainable=need_pre,
                                            batch_sum=True)
        return grad_pre, grad_w, None, None, None, None, None, None


def lif(pre_times, weights, tau_s, theta, delta_theta, max_n_spike, t_max=np.inf,
        recompute=False):
    """Differentiable dense LIF layer: returns (spike times [B,J,K], counts [B,J])."""
    return LIFFunction.apply(pre_times, weights, tau_s, theta, delta_theta, max_n_spike, t_max,
                             recompute)


class SpikeCountLossFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, t, n, labels, target_true, target_false):
        loss, errors, _ = ops.count_loss(n, t.shape[2], labels, target_true, target_false)
        ctx.save_for_backward(errors)
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        (errors,) = ctx.saved_tensors
        return errors * grad_loss[:, None, None], None, None, None, None


def spike_count_loss(t, n, labels, target_true, target_false):
    """Per-sample spike-count loss [B] whose gradient into ``t`` is the reference's surrogate."""
    return SpikeCountLossFunction.apply(t, n, labels, target_true, target_false)
