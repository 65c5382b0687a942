"""The golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle must reproduce them exactly (CPU), and the HIP path must match them at the
parity tolerances (GPU): identical spike counts, spike times / a / x / z within 1e-5
relative, gradients and errors within 1e-4 |ref| + 1e-6 max|ref|."""
from pathlib import Path

import numpy as np
import pytest

from oracle import lif as olif
from tests.helpers import assert_close_rel, spike_slots

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = ["lif_seed42_kernel", "lif_seed42_layer", "lif_multi_spike"]


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(name):
    g = load(name)
    st = olif.lif_forward(g["pre_times"], g["W"], float(g["tau_s"]), float(g["theta"]),
                          float(g["delta_theta"]), int(g["K"]), tmax=float(g["tmax"]))
    assert np.array_equal(st.n, g["n"])
    m = spike_slots(g["n"], int(g["K"]))
    for k in ("t", "a", "x", "z"):
        assert np.array_equal(getattr(st, k)[m], g[k][m])
    e, gr, p = olif.lif_backward(st, g["errors_in"], pre_trainable=True)
    assert np.array_equal(e, g["errors_out"])
    np.testing.assert_array_equal(gr, g["grad"])
    np.testing.assert_array_equal(p, g["pre_errors"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_fixture(name):
    import torch
    from bats_amd import ops
    from bats_amd._constants import lif_constants
    g = load(name)
    print(f"[golden] {name}", flush=True)
    ts, tau, c, dtt = lif_constants(float(g["tau_s"]), float(g["theta"]), float(g["delta_theta"]))
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    K = int(g["K"])
    n, t, a, x, z, es, e = ops.lif_forward(T(g["pre_times"]), T(g["W"]), ts, tau, c, dtt,
                                           float(g["tmax"]), K)
    err = T(g["errors_in"])
    grad, pre = ops.lif_backward(n, t, a, x, z, err, T(g["pre_times"]), es, e, T(g["W"]), ts,
                                 tau, c, dtt, per_sample=True, pre_trainable=True)
    torch.cuda.synchronize()
    assert np.array_equal(n.cpu().numpy(), g["n"])
    m = spike_slots(g["n"], K)
    for name_, arr in (("t", t), ("a", a), ("x", x), ("z", z)):
        gv = arr.cpu().numpy()[m].astype(np.float64)
        rv = g[name_][m].astype(np.float64)
        assert np.all(np.abs(gv - rv) <= 1e-5 * np.abs(rv)), name_
    assert_close_rel(err.cpu().numpy(), g["errors_out"], 1e-4, 1e-6, "errors")
    assert_close_rel(grad.cpu().numpy(), g["grad"], 1e-4, 1e-6, "grad")
    assert_close_rel(pre.cpu().numpy(), g["pre_errors"], 1e-4, 1e-6, "pre errors")
