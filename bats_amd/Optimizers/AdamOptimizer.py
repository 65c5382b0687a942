"""Adam (bats/Optimizers/AdamOptimizer.py:22-55) on the fused HIP Adam kernel."""
from __future__ import annotations

from pathlib import Path
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
        if self._t_dev is not None:
            self._t_dev.fill_(int(t))

    STATE_FILE = "adam_state.npz"

    def store(self, dir_path) -> None:
        """Writes `adam_state.npz`: the step counter, the hyperparameters and the float32
        moment estimates m_<k>, v_<k> of every layer k that has them (k indexes the
        gradient list of `step`), so that training resumes exactly where it stopped."""
        m, v, t = self.state
        arrays = {"t": np.int64(t), "n": np.int64(0 if m is None else len(m)),
                  "learning_rate": np.float32(self._learning_rate),
                  "beta_1": self._beta_1, "beta_2": self._beta_2, "epsilon": self._epsilon,
                  "nadam": np.int8(self._nadam)}
        for k, (mk, vk) in enumerate(zip(m or [], v or [])):
            if mk is not None:
                arrays[f"m_{k}"] = mk.detach().cpu().numpy()
                arrays[f"v_{k}"] = vk.detach().cpu().numpy()
        np.savez(Path(dir_path) / self.STATE_FILE, **arrays)

    def restore(self, dir_path, device=None) -> None:
        """Reads `adam_state.npz` if present (no pickled objects are loaded); the moments go
        to `device` (the current GPU by default, the CPU when none is visible)."""
        path = Path(dir_path) / self.STATE_FILE
        if not path.exists():
            return
        if device is None:
            device = torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")
        with np.load(path, allow_pickle=False) as f:
            if int(f["nadam"]) != int(self._nadam):
                raise ValueError(f"{path} holds the state of another optimizer")
            n = int(f["n"])
            m = [None] * n
            v = [None] * n
            for k in range(n):
                if f"m_{k}" in f:
                    m[k] = torch.as_tensor(f[f"m_{k}"], device=device).clone()
                    v[k] = torch.as_tensor(f[f"v_{k}"], device=device).clone()
            self.learning_rate = float(f["learning_rate"])
            self.load_state(m if n else None, v if n else None, int(f["t"]))

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
                # starts from the host count, so that a restored state resumes at t + 1
                dev = next(g.device for g in gradient if g is not None)
                self._t_dev = torch.full((), self._t - 1, dtype=torch.int64, device=dev)
            self._t_dev.add_(1)
            step_dev = self._t_dev
        b1, b2 = self._bias()
        # every layer's update in one launch
        keys = [k for k, g in enumerate(gradient) if g is not None]
        gs = [ops.as_device(gradient[k], torch.float32,
                            gradient[k].device if isinstance(gradient[k], torch.Tensor) else None)
              for k in keys]
        ws = None if weights is None else [weights[k] for k in keys]
        deltas = ops.adam_step_multi(gs, [self._m[k] for k in keys], [self._v[k] for k in keys],
                                     self._learning_rate, self._beta_1, self._one_minus_beta_1,
                                     self._beta_2, self._one_minus_beta_2, self._epsilon, b1,
                                     b2, nadam=self._nadam, weights=ws, step_dev=step_dev)
        out: List[Optional[torch.Tensor]] = [None] * len(gradient)
        for k, d in zip(keys, deltas):
            out[k] = d
        return out
