"""Training monitors (bats/AbstractMonitor.py:9-45), accumulating on the device.

The reference monitors pull every batch's statistics to the host (`.get()` in each `add`,
e.g. bats/Monitors/SpikeCountMonitor.py:18), which synchronises the stream once per batch
and monitor. Here `add` only enqueues device reductions into running float64 sums; the
single device-to-host read happens in `record` (once per epoch or reporting interval).
Inputs may be torch tensors (any device), CuPy-free NumPy arrays or Python numbers.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from pathlib import Path
from typing import Optional

import numpy as np
import torch


def as_tensor(x, device=None) -> torch.Tensor:
    """A tensor view of `x` (torch tensors are used as they are)."""
    if isinstance(x, torch.Tensor):
        return x if device is None else x.to(device)
    return torch.as_tensor(np.asarray(x), device=device)


class DeviceSum:
    """A running float64 sum that lives on the device of the first value added."""

    def __init__(self):
        self._value: Optional[torch.Tensor] = None

    def add(self, t: torch.Tensor) -> None:
        t = t.to(torch.float64)
        if self._value is None:
            self._value = t.clone()
        else:
            self._value += t

    def take(self):
        """The sum as a host value (float for a scalar, ndarray otherwise), then reset."""
        v = self._value
        self._value = None
        if v is None:
            return 0.0
        v = v.cpu()
        return float(v) if v.dim() == 0 else v.numpy()

    @property
    def empty(self) -> bool:
        return self._value is None


class AbstractMonitor(ABC):
    def __init__(self, name: str, export_path: Optional[Path] = None, decimal: int = 2):
        self._name: str = name
        self._export_path: Optional[Path] = export_path
        self._decimal = decimal
        self._epochs = []
        self._values = []

    def __str__(self) -> str:
        return self._name + ": " + self._format()

    def _format(self) -> str:
        return str(np.around(self._values[-1], self._decimal))

    def reset(self) -> None:
        self._epochs = []
        self._values = []

    def _record(self, epoch: float, value: float) -> None:
        self._epochs.append(epoch)
        self._values.append(value)

    @property
    def epochs(self):
        return list(self._epochs)

    @property
    def values(self):
        return list(self._values)

    @abstractmethod
    def record(self, epoch: float) -> float:
        pass

    def export(self) -> None:
        """Writes `<export_path>.npz` (epochs, values) and, when matplotlib is importable, a
        plot next to it (bats/AbstractMonitor.py:36-45)."""
        if self._export_path is None:
            return
        path = Path(self._export_path)
        np.savez(path, epochs=self._epochs, values=self._values)
        try:
            import matplotlib
            matplotlib.use("Agg")
            import matplotlib.pyplot as plt
        except Exception:  # plotting is optional
            return
        plt.plot(self._epochs, self._values)
        plt.xlabel("Epoch")
        plt.ylabel(self._name)
        plt.savefig(path.with_suffix(".png"))
        plt.close()
