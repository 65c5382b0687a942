"""Writes the golden fixtures of tests/golden/ from the FP64 oracle.

The reference repository ships no numeric fixtures for this path (its tests draw fresh
random inputs or use brian2 at run time); the hand-written arrays its tests do contain are
used directly as known-answer tests in tests/test_oracle_pins.py. These fixtures freeze the
oracle's outputs on the reference tests' seeded inputs (regenerated with NumPy's legacy
RandomState, seed 42) and on small multi-spike cases, so that GPU parity can be checked
against stored vectors and any change of the oracle is caught.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from oracle import events as oev  # noqa: E402
from oracle import lif as olif  # noqa: E402

F32 = np.float32


def seed42_kernel_case():
    # inputs of tests/CudaKernels/Inference/compute_spike_times_tests.py:44-55
    np.random.seed(42)
    idx = np.random.randint(10, size=(1, 100))
    tt = np.sort(np.random.uniform(0.0, 0.5, size=(1, 100)), axis=1)
    tt[0, -28:] = np.inf
    W = np.random.normal(0, 1.0, size=(5, 10)).astype(F32)
    ii = np.where(np.isfinite(tt), idx, -1).astype(np.int32)
    per, _ = oev.events_to_per_neuron(ii, tt.astype(F32), 10)
    return dict(pre_times=per, W=W, tau_s=0.1, theta=0.1, delta_theta=0.2, K=128, tmax=0.5)


def seed42_layer_case():
    # inputs of tests/Layers/LIFLayer_tests.py:44-58
    np.random.seed(42)
    n_in = np.random.randint(0, 5, size=(1, 1000))
    spk = np.full((1, 1000, 5), np.inf, dtype=F32)
    for s, ns in zip(spk, n_in):
        for row, k in zip(s, ns):
            row[:k] = np.sort(np.random.uniform(0.0, 0.5, size=(k,)))
    W = np.random.normal(0, 1.0, size=(50, 1000)).astype(F32)
    return dict(pre_times=spk, W=W, tau_s=0.1, theta=0.1, delta_theta=0.2, K=128, tmax=0.5)


def multi_spike_case(seed):
    rng = np.random.default_rng(seed)
    B, I, J, Kp, K = 4, 40, 12, 3, 6
    n = rng.integers(0, Kp + 1, size=(B, I))
    t = np.sort(rng.uniform(0, 0.8, size=(B, I, Kp)), axis=2).astype(F32)
    t[np.arange(Kp)[None, None, :] >= n[:, :, None]] = np.inf
    W = rng.normal(0.5, 1.0, size=(J, I)).astype(F32)
    return dict(pre_times=t, W=W, tau_s=0.05, theta=0.2, delta_theta=0.15, K=K, tmax=0.7)


def write(name, case, with_backward, seed=0):
    st = olif.lif_forward(case["pre_times"], case["W"], case["tau_s"], case["theta"],
                          case["delta_theta"], case["K"], tmax=case["tmax"])
    out = dict(case)
    out.update(n=st.n, t=st.t, a=st.a, x=st.x, z=st.z, pre_exp_s=st.pre_exp_s,
               pre_exp=st.pre_exp)
    if with_backward:
        rng = np.random.default_rng(100 + seed)
        errors = rng.normal(size=st.t.shape).astype(F32)
        e, g, p = olif.lif_backward(st, errors, pre_trainable=True)
        out.update(errors_in=errors, errors_out=e, grad=g, pre_errors=p)
    np.savez_compressed(HERE / f"{name}.npz", **out)
    print(f"{name}: {int(st.n.sum())} spikes")


if __name__ == "__main__":
    write("lif_seed42_kernel", seed42_kernel_case(), with_backward=True)
    write("lif_seed42_layer", seed42_layer_case(), with_backward=True, seed=1)
    write("lif_multi_spike", multi_spike_case(0), with_backward=True, seed=2)
