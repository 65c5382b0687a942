"""Shared input generators and comparison helpers for the parity tests."""
from __future__ import annotations

import numpy as np

F32 = np.float32


def latency_inputs(rng, B, I, threshold=0.1):
    """x ~ U[0,1] per (sample, input); active iff x > threshold; t = f32(1 - x); 1 slot."""
    x = rng.uniform(0.0, 1.0, size=(B, I))
    t = np.where(x > threshold, (1.0 - x), np.inf).astype(F32)
    return t[:, :, None], np.isfinite(t).astype(np.int32)


def multi_spike_inputs(rng, B, I, K, t_max=1.0, p_silent=0.3):
    """Up to K ascending spikes per input neuron, +inf padded."""
    n = rng.integers(0, K + 1, size=(B, I))
    n[rng.uniform(size=(B, I)) < p_silent] = 0
    t = np.sort(rng.uniform(0.0, t_max, size=(B, I, K)), axis=2).astype(F32)
    t[np.arange(K)[None, None, :] >= n[:, :, None]] = np.inf
    return t, n.astype(np.int32)


def spike_slots(n, K):
    return np.arange(K)[None, None, :] < n[:, :, None]


def assert_forward_match(n_g, t_g, a_g, x_g, z_g, ref, rtol=1e-5):
    """Spike counts identical; t, a, x, z of every emitted spike within rtol (relative)."""
    n_r = ref.n
    bad = np.argwhere(n_g != n_r)
    assert bad.size == 0, f"spike counts differ at {bad[:5].tolist()} gpu={n_g[tuple(bad[0])]} ref={n_r[tuple(bad[0])]}"
    K = t_g.shape[2]
    m = spike_slots(n_r, K)
    for name, g, r in (("t", t_g, ref.t), ("a", a_g, ref.a), ("x", x_g, ref.x), ("z", z_g, ref.z)):
        gv = g[m].astype(np.float64)
        rv = r[m].astype(np.float64)
        err = np.abs(gv - rv) / np.maximum(np.abs(rv), 1e-30)
        assert err.size == 0 or err.max() <= rtol, f"{name}: max rel err {err.max():.3e}"
    assert np.all(np.isinf(t_g[~m])), "empty slots of t must be +inf"


def assert_close_rel(g, r, rtol, atol_frac, what):
    """|g - r| <= rtol*|r| + atol_frac*max|r| elementwise."""
    g = np.asarray(g, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    scale = np.max(np.abs(r)) if r.size else 0.0
    tol = rtol * np.abs(r) + atol_frac * scale
    diff = np.abs(g - r)
    ok = diff <= tol
    if not ok.all():
        k = np.argmax(diff - tol)
        raise AssertionError(f"{what}: |diff|={diff.flat[k]:.3e} > tol={tol.flat[k]:.3e} "
                             f"(gpu={g.flat[k]:.6e}, ref={r.flat[k]:.6e})")
