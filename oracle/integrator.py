"""Independent brute-force LIF membrane integrator used to pin the oracle's solver.

It does not use the (a, b, z) quadratic of the reference solver. It evaluates the
membrane potential directly from the neuron model that the reference tests simulate
with brian2 (tests/CudaKernels/Inference/compute_spike_times_tests.py:15-22):
du/dt = -u/tau + v, dv/dt = -v/tau_s, an input spike adds w to v, threshold u >= theta,
reset u -= delta_theta. With tau = 2 tau_s each input contributes
w*tau*(exp(-(t-ti)/tau) - exp(-(t-ti)/tau_s)) and each output spike -delta_theta*exp(-(t-tk)/tau).
Crossings are located on a dense grid between input events and refined by bisection.
Test infrastructure only.
"""
from __future__ import annotations

import numpy as np


def membrane(t, in_t, in_w, out_t, tau_s, tau, delta_theta):
    t = np.atleast_1d(np.asarray(t, dtype=np.float64))
    u = np.zeros_like(t)
    for ti, wi in zip(in_t, in_w):
        s = t - ti
        m = s >= 0
        u[m] += wi * tau * (np.exp(-s[m] / tau) - np.exp(-s[m] / tau_s))
    for tk in out_t:
        s = t - tk
        m = s >= 0
        u[m] -= delta_theta * np.exp(-s[m] / tau)
    return u


def simulate_neuron(in_t, in_w, tau_s, tau, theta, delta_theta, tmax, max_spikes,
                    grid=400):
    """Spike times of one neuron (float64), inputs sorted by time."""
    in_t = np.asarray(in_t, dtype=np.float64)
    in_w = np.asarray(in_w, dtype=np.float64)
    order = np.argsort(in_t, kind="stable")
    in_t = in_t[order]
    in_w = in_w[order]
    out = []
    bounds = list(in_t) + [np.inf]
    for q in range(len(in_t)):
        lo = in_t[q]
        hi = min(bounds[q + 1], tmax)
        if not np.isfinite(hi):
            hi = lo + 40.0 * tau
        if hi <= lo:
            continue
        vis_t = in_t[: q + 1]
        vis_w = in_w[: q + 1]
        start = lo
        while len(out) < max_spikes:
            ts = np.linspace(start, hi, grid)
            u = membrane(ts, vis_t, vis_w, out, tau_s, tau, delta_theta) - theta
            above = np.nonzero(u[1:] >= 0)[0]
            if above.size == 0:
                break
            k = above[0] + 1
            a_, b_ = ts[k - 1], ts[k]
            if u[k - 1] >= 0:  # already above at the left end (only right after a reset)
                break
            for _ in range(200):
                mid = 0.5 * (a_ + b_)
                if membrane(mid, vis_t, vis_w, out, tau_s, tau, delta_theta)[0] - theta >= 0:
                    b_ = mid
                else:
                    a_ = mid
                if b_ - a_ < 1e-14:
                    break
            out.append(b_)
            start = b_
        if len(out) >= max_spikes:
            break
    return np.array(out)
