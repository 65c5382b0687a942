"""Adam (bats/Optimizers/AdamOptimizer.py:22-55) on the fused HIP Adam kernel."""
from __future__ import annotations

from typing import List, Optional

import numpy as np
import torch

from .. import ops
from ..AbstractOptimizer import AbstractOptimizer


class AdamOptimizer(AbstractOptimizer):
    _nadam = False

    def __init__(self, beta_1=0.9, beta_2=0.999, epsilon=1e-8, capturable: bool = False,
                 **kwargs):
        """``capturable=True`` keeps the step counter on the device so that a HIP graph of the
        whole training step can be replayed (the reference keeps it on the host)."""
        super().__init__(**kwargs)
        self._capturable = bool(capturable)
        self._t_dev: Optional[torch.Tensor] = None
        self._beta_1 = np.float32(beta_1)
        self._one_minus_beta_1 = np.float32(1.0 - beta_1)
        self._beta_2 = np.float32(beta_2)
        self._one_minus_beta_2 = np.float32(1.0 - beta_2)
        self._epsilon = np.float32(epsilon)
        self._m: Optional[List[Optional[torch.Tensor]]] = None
        self._v: Optional[List[Optional[torch.Tensor]]] = None
        self._t = 0

    @property
    def state(self):
        """(m, v, t) -- the moment estimates and the step counter."""
        t = int(self._t_dev.item()) if self._t_dev is not None else self._t
        return self._m, self._v, t

    def load_state(self, m, v, t) -> None:
        self._m, self._v, self._t = m, v, int(t)

    def _bias(self):
        b1 = 1.0 - np.float64(self._beta_1) ** self._t
        b2 = 1.0 - np.float64(self._beta_2) ** self._t
        return b1, b2

    def step(self, gradient: List[Optional[torch.Tensor]], weights=None) -> List[Optional[torch.Tensor]]:
        """Returns the deltas (the reference's contract). If ``weights`` (a list of weight
        tensors aligned with ``gradient``) is given, the deltas are also added to them in the
        same kernel."""
        self._t += 1
        if self._m is None:
            self._m = [None if g is None else torch.zeros_like(g, dtype=torch.float32) for g in gradient]
            self._v = [None if g is None else torch.zeros_like(g, dtype=torch.float32) for g in gradient]
        step_dev = None
        if self._capturable:
            if self._t_dev is None:
                dev = next(g.device for g in gradient if g is not None)
                self._t_dev = torch.zeros((), dtype=torch.int64, device=dev)
            self._t_dev.add_(1)
            step_dev = self._t_dev
        b1, b2 = self._bias()
        out = []
        for k, g in enumerate(gradient):
            if g is None:
                out.append(None)
                continue
            g = ops.as_device(g, torch.float32, g.device if isinstance(g, torch.Tensor) else None)
            w = None if weights is None else weights[k]
            out.append(ops.adam_step(g, self._m[k], self._v[k], self._learning_rate,
                                     self._beta_1, self._one_minus_beta_1, self._beta_2,
                                     self._one_minus_beta_2, self._epsilon, b1, b2,
                                     nadam=self._nadam, weights=w, step_dev=step_dev))
        return out
