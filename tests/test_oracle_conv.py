"""CPU pins of the conv / pooling oracle: the conv forward against the brute-force membrane
integrator run on each neuron's receptive field, the filter gradient against float64 central
differences, and the pooling against its invariants (counts add up, merged lists ascending,
every source slot used once and holding the same time)."""
import numpy as np
import pytest

from oracle import conv as oconv
from oracle import integrator
from oracle import lif as olif

F32 = np.float32
F64 = np.float64


def _inputs(rng, B, shape, K, tmax=0.6, p=0.5):
    X, Y, C = shape
    N = X * Y * C
    n = (rng.uniform(size=(B, N)) < p) * rng.integers(1, K + 1, size=(B, N))
    t = np.sort(rng.uniform(0, tmax, size=(B, N, K)), axis=2).astype(F32)
    t[np.arange(K)[None, None, :] >= n[:, :, None]] = np.inf
    return t


def test_conv_forward_matches_integrator_on_receptive_fields():
    rng = np.random.default_rng(0)
    pre_shape = (6, 6, 2)
    W = rng.normal(0.8, 1.0, size=(3, 3, 3, 2)).astype(F32)  # C_out=3, 3x3, C_in=2
    t = _inputs(rng, 2, pre_shape, 2)
    params = (0.05, 0.08, 0.06)
    st = oconv.conv_forward(t, W, pre_shape, *params, 4, tmax=0.6, store=F64)
    ts32, tau32, c32, dtt32 = olif.lif_constants(*params)
    Xo, Yo, Co = st.post_shape
    X, Y, C = pre_shape
    checked = 0
    for b in range(2):
        for x in range(Xo):
            for y in range(Yo):
                for c in range(Co):
                    tin, win = [], []
                    for dx in range(3):
                        for dy in range(3):
                            for ci in range(C):
                                i = ((x + dx) * Y + (y + dy)) * C + ci
                                for tv in t[b, i]:
                                    if np.isfinite(tv):
                                        tin.append(F64(tv))
                                        win.append(F64(W[c, dx, dy, ci]))
                    ref = integrator.simulate_neuron(tin, win, F64(ts32), F64(tau32),
                                                     F64(c32) * F64(tau32),
                                                     F64(dtt32) * F64(tau32), F64(F32(0.6)), 4)
                    j = (x * Yo + y) * Co + c
                    assert len(ref) == st.n[b, j]
                    if len(ref):
                        assert np.max(np.abs(ref - st.t[b, j, :len(ref)])) < 1e-9
                    checked += len(ref)
    assert checked > 20


def test_conv_filter_gradient_central_differences():
    rng = np.random.default_rng(1)
    pre_shape = (5, 5, 1)
    W = rng.normal(1.0, 0.8, size=(2, 3, 3, 1))
    t = _inputs(rng, 2, pre_shape, 2, p=0.7).astype(F64)
    params = (0.05, 0.06, 0.04)
    K = 6
    st0 = oconv.conv_forward(t, W, pre_shape, *params, K, tmax=0.6, store=F64)
    assert st0.n.sum() > 5
    e = rng.normal(size=st0.t.shape)
    _, grad, _ = oconv.conv_backward(st0, e.copy(), pre_trainable=False, store=F64)
    g = grad.sum(axis=0)

    def L(Wp):
        st = oconv.conv_forward(t, Wp, pre_shape, *params, K, tmax=0.6, store=F64)
        if not np.array_equal(st.n, st0.n):
            return None
        return np.where(np.isfinite(st.t), st.t * e, 0.0).sum()

    checked = 0
    for idx in np.ndindex(W.shape):
        h = 1e-6 * (abs(W[idx]) + 1.0)
        Wp, Wm = W.copy(), W.copy()
        Wp[idx] += h
        Wm[idx] -= h
        lp, lm = L(Wp), L(Wm)
        if lp is None or lm is None:
            continue
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, g[idx])
        checked += 1
    assert checked >= 12


def test_pooling_invariants_and_backward():
    rng = np.random.default_rng(2)
    pre_shape = (6, 4, 3)
    K = 3
    t = _inputs(rng, 2, pre_shape, K)
    n = np.isfinite(t).sum(axis=2).astype(np.int32)
    n_out, t_out, src, post_shape = oconv.pool_forward(n, t, pre_shape)
    assert post_shape == (3, 2, 3)
    flat = t.reshape(-1)
    assert np.array_equal(np.sort(src.reshape(-1)), np.arange(flat.size))
    assert np.array_equal(flat[src], t_out)
    assert np.all(np.diff(np.where(np.isinf(t_out), 1e9, t_out), axis=2) >= 0)
    assert np.array_equal(n_out, np.isfinite(t_out).sum(axis=2))
    err = rng.normal(size=t_out.shape).astype(F32)
    back = oconv.pool_backward(err, src, t.shape[1], K)
    assert np.array_equal(back.reshape(-1)[src.reshape(-1)], err.reshape(-1))
