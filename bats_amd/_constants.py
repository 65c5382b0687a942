"""Float32 layer constants, computed as bats/Layers/LIFLayer.py:16-19 does."""
from __future__ import annotations

import numpy as np


def lif_constants(tau_s, theta, delta_theta):
    """Returns (tau_s, tau, c, delta_theta_tau) as numpy.float32.

    tau = f32(2*tau_s); c = f32(theta/tau) and delta_theta_tau = f32(delta_theta/tau), the
    divisions evaluated in float64 on the float32 tau (NumPy's legacy promotion of
    ``python float / numpy.float32``) and rounded to float32.
    """
    tau_s32 = np.float32(tau_s)
    tau32 = np.float32(2 * np.float64(tau_s))
    c32 = np.float32(np.float64(theta) / np.float64(tau32))
    dtt32 = np.float32(np.float64(delta_theta) / np.float64(tau32))
    return tau_s32, tau32, c32, dtt32
