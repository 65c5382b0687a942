"""FP64 restatement of the convolutional LIF layer and 2x2 spike pooling.  Test
infrastructure only.

Neuron index of a (x, y, c) position is (x*Y + y)*C + c
(bats/CudaKernels/CuFiles/Inference/aggregate_spikes_conv.cu:2-4); conv weights are
[C_out, fx, fy, C_in] (bats/Layers/ConvLIFLayer.py:28,35-37); valid, stride-1.

The conv forward walks every input event of the sample and looks the weight up through the
receptive field, 0 outside it (compute_spike_times_conv.cu:93-115,155-173). An event with a
zero weight leaves a and b unchanged, and the crossing found after it is the same root that
the previous receptive-field event's check already tested against a longer interval, so the
layer is the dense layer with the densified weight matrix -- the equivalence the reference's
own test uses (tests/CudaKernels/Inference/compute_spike_times_conv_tests.py:82-93).
"""
from __future__ import annotations

import numpy as np

from . import lif as olif

F32 = np.float32
F64 = np.float64


def conv_out_shape(pre_shape, filters_shape):
    """bats/Layers/ConvLIFLayer.py:20-24: filters_shape = (fx, fy, C_out)."""
    X, Y, _ = pre_shape
    fx, fy, co = filters_shape
    return (X - fx + 1, Y - fy + 1, co)


def densify(W, pre_shape, post_shape):
    """Conv filters [C_out, fx, fy, C_in] -> dense [N_post, N_pre] (0 outside the field)."""
    Co, fx, fy, Ci = W.shape
    X, Y, C = pre_shape
    Xo, Yo, _ = post_shape
    D = np.zeros((Xo * Yo * Co, X * Y * C), dtype=W.dtype)
    for x in range(Xo):
        for y in range(Yo):
            for c in range(Co):
                j = (x * Yo + y) * Co + c
                for dx in range(fx):
                    for dy in range(fy):
                        base = ((x + dx) * Y + (y + dy)) * C
                        D[j, base:base + Ci] = W[c, dx, dy, :]
    return D


def conv_forward(pre_times, W, pre_shape, tau_s, theta, delta_theta, max_n_post,
                 tmax=np.inf, store=F32):
    """bats/Layers/ConvLIFLayer.py:77-107 (K1 + sort + K5)."""
    post_shape = conv_out_shape(pre_shape, (W.shape[1], W.shape[2], W.shape[0]))
    D = densify(np.asarray(W, dtype=store), pre_shape, post_shape)
    st = olif.lif_forward(pre_times, D, tau_s, theta, delta_theta, max_n_post, tmax=tmax,
                          store=store)
    st.post_shape = post_shape
    st.pre_shape = tuple(pre_shape)
    st.Wconv = np.asarray(W, dtype=store)
    return st


def conv_backward(state, errors, pre_trainable=True, store=F32):
    """bats/Layers/ConvLIFLayer.py:109-134: reset chain, factors, the per-sample filter
    gradient [B, C_out, fx, fy, C_in] (compute_weights_gradient_conv.cu:31-60: the dense
    per-sample gradient summed over the positions that share each filter weight) and the
    pre-spike errors (propagate_errors_to_pre_spikes_conv.cu + the sum over post neurons)."""
    e = olif.recurrent_errors(state.x, state.z, errors, state.dtt, store=store)
    f1, f2 = olif.factors(state.t, state.a, state.x, state.z, state.c, state.tau, store=store)
    G = olif.weights_gradient(f1, f2, state.t, state.pre_times, state.pre_exp_s,
                              state.pre_exp, e, store=F64)
    Co, fx, fy, Ci = state.Wconv.shape
    X, Y, C = state.pre_shape
    Xo, Yo, _ = state.post_shape
    B = G.shape[0]
    grad = np.zeros((B, Co, fx, fy, Ci))
    for x in range(Xo):
        for y in range(Yo):
            for c in range(Co):
                j = (x * Yo + y) * Co + c
                for dx in range(fx):
                    for dy in range(fy):
                        base = ((x + dx) * Y + (y + dy)) * C
                        grad[:, c, dx, dy, :] += G[:, j, base:base + Ci]
    pre_err = None
    if pre_trainable:
        pre_err = olif.pre_spike_errors(f1, f2, state.t, state.pre_times, state.pre_exp_s,
                                        state.pre_exp, state.W, e, state.tau_s, state.tau,
                                        store=store)
    return e, grad.astype(store), pre_err


def pool_forward(n, t, pre_shape):
    """2x2 spike pooling, bats/Layers/PoolingLayer.py:49-53 with
    aggregate_spikes_conv.cu:13-41 and the argsort of
    bats/CudaKernels/Wrappers/Inference/aggregate_spikes_conv.py:24-26.

    n [B, N_pre], t [B, N_pre, K]. Returns (n_out [B, N_post], t_out [B, N_post, 4K] sorted,
    src [B, N_post, 4K] flat indices into t (sample offset included), post_shape)."""
    B, N, K = t.shape
    X, Y, C = pre_shape
    Xo, Yo = X // 2, Y // 2
    No = Xo * Yo * C
    n_out = np.zeros((B, No), dtype=np.int32)
    t_cat = np.empty((B, No, 4 * K), dtype=F32)
    src = np.empty((B, No, 4 * K), dtype=np.int64)
    flat_idx = np.arange(B * N * K).reshape(B, N, K)
    for x in range(Xo):
        for y in range(Yo):
            for c in range(C):
                j = (x * Yo + y) * C + c
                slot = 0
                for dx in range(2):
                    for dy in range(2):
                        i = ((2 * x + dx) * Y + (2 * y + dy)) * C + c
                        n_out[:, j] += n[:, i]
                        t_cat[:, j, slot:slot + K] = t[:, i, :]
                        src[:, j, slot:slot + K] = flat_idx[:, i, :]
                        slot += K
    order = np.argsort(t_cat, axis=2, kind="stable")
    return (n_out, np.take_along_axis(t_cat, order, axis=2),
            np.take_along_axis(src, order, axis=2), (Xo, Yo, C))


def pool_backward(errors, src, n_pre, K_pre):
    """bats/Layers/PoolingLayer.py:55-66: scatter the errors back to their source slots."""
    B = errors.shape[0]
    out = np.zeros(B * n_pre * K_pre, dtype=errors.dtype)
    out[src.reshape(-1)] = errors.reshape(-1)
    return out.reshape(B, n_pre, K_pre)
