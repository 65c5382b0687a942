"""The reference's monitors (bats/Monitors/*.py) with device-side running sums."""
from __future__ import annotations

import time
from typing import Dict, List

import numpy as np
import torch

from .AbstractMonitor import AbstractMonitor, DeviceSum, as_tensor


class AccuracyMonitor(AbstractMonitor):
    """Percentage of correct predictions (bats/Monitors/AccuracyMonitor.py:4-19)."""

    def __init__(self, **kwargs):
        super().__init__("Accuracy (%)", **kwargs)
        self._hits = DeviceSum()
        self._n_samples = 0

    def add(self, predictions, targets) -> None:
        p = as_tensor(predictions)
        t = as_tensor(targets, p.device)
        self._hits.add((p == t).sum())
        self._n_samples += int(t.shape[0])

    def record(self, epoch) -> float:
        accuracy = self._hits.take() / self._n_samples * 100
        super()._record(epoch, accuracy)
        self._n_samples = 0
        return accuracy


class LossMonitor(AbstractMonitor):
    """Mean of the finite losses (bats/Monitors/LossMonitor.py:7-27)."""

    def __init__(self, **kwargs):
        super().__init__("Loss", **kwargs)
        self._losses_cumul = DeviceSum()
        self._n_valid_loss = DeviceSum()

    def add(self, losses) -> None:
        ls = as_tensor(losses)
        valid = torch.isfinite(ls)
        self._losses_cumul.add(torch.where(valid, ls, torch.zeros_like(ls)).to(torch.float64).sum())
        self._n_valid_loss.add(valid.sum())

    def record(self, epoch) -> float:
        loss = self._losses_cumul.take() / self._n_valid_loss.take()
        super()._record(epoch, loss)
        return loss


class TimeMonitor(AbstractMonitor):
    """Wall time since `start` (bats/Monitors/TimeMonitor.py:7-20)."""

    def __init__(self, **kwargs):
        super().__init__("Time (s)", **kwargs)
        self._start_time = 0.0

    def start(self) -> None:
        self._start_time = time.time()

    def record(self, epoch) -> float:
        elapsed_time = time.time() - self._start_time
        super()._record(epoch, elapsed_time)
        return elapsed_time


class SilentLabelsMonitor(AbstractMonitor):
    """Percentage of samples whose true-class output neuron never fired
    (bats/Monitors/SilentLabelsMonitor.py:7-25)."""

    def __init__(self, **kwargs):
        super().__init__("Silent labels (%)", **kwargs)
        self._silent_labels = DeviceSum()
        self._n_samples = 0

    def add(self, n_out_spikes, targets) -> None:
        n = as_tensor(n_out_spikes)
        t = as_tensor(targets, n.device).to(torch.int64)
        counts = torch.take_along_dim(n, t[:, None], dim=1)
        self._silent_labels.add((counts == 0).sum())
        self._n_samples += int(t.shape[0])

    def record(self, epoch) -> float:
        silent_ratio = self._silent_labels.take() / self._n_samples * 100
        super()._record(epoch, silent_ratio)
        self._n_samples = 0
        return silent_ratio


class SilentNeuronsMonitor(AbstractMonitor):
    """Percentage of a layer's neurons that never fired over the recorded batches
    (bats/Monitors/SilentNeuronsMonitor.py:7-25)."""

    def __init__(self, layer_name: str, **kwargs):
        super().__init__(layer_name + " silents (%)", **kwargs)
        self._total_counts = DeviceSum()

    def add(self, n_spikes) -> None:
        self._total_counts.add(as_tensor(n_spikes).sum(dim=0))

    def record(self, epoch) -> float:
        counts = self._total_counts.take()
        silent_ratio = float(np.mean(np.asarray(counts) == 0) * 100)
        super()._record(epoch, silent_ratio)
        return silent_ratio


class SpikeCountMonitor(AbstractMonitor):
    """Mean number of spikes of a layer per sample (bats/Monitors/SpikeCountMonitor.py:11-28)."""

    def __init__(self, layer_name: str, **kwargs):
        super().__init__(layer_name + " spike counts", **kwargs)
        self._counts = DeviceSum()
        self._n_samples = 0.0

    def add(self, n_spike_per_neuron) -> None:
        n = as_tensor(n_spike_per_neuron)
        self._counts.add(n.sum())
        self._n_samples += n.shape[0]

    def record(self, epoch) -> float:
        avg_count = self._counts.take() / self._n_samples
        super()._record(epoch, avg_count)
        self._n_samples = 0.0
        return avg_count


class WeightsNormMonitor(AbstractMonitor):
    """Frobenius norm of a layer's weights (bats/Monitors/WeightsNormMonitor.py:6-18)."""

    def __init__(self, layer_name: str, **kwargs):
        super().__init__(layer_name + " weights norm", **kwargs)
        self.__norm = None

    def add(self, weights) -> None:
        w = as_tensor(weights)
        self.__norm = torch.linalg.vector_norm(w.to(torch.float64))  # stays on the device

    def record(self, epoch) -> float:
        norm = float(self.__norm.cpu()) if self.__norm is not None else float("nan")
        super()._record(epoch, norm)
        return norm


class ValueMonitor(AbstractMonitor):
    """Records the last value given (bats/Monitors/ValueMonitor.py:6-16)."""

    def __init__(self, **kwargs):
        super().__init__(**kwargs)
        self.__value = 0.0

    def add(self, value) -> None:
        self.__value = value

    def record(self, epoch) -> float:
        v = self.__value
        v = float(v.cpu()) if isinstance(v, torch.Tensor) else float(v)
        self._record(epoch, v)
        return v


class MonitorsManager:
    """Groups monitors for printing, recording and export (bats/Monitors/MonitorsManager.py:7-28)."""

    def __init__(self, monitors: List[AbstractMonitor] = None, print_prefix: str = ""):
        self._monitors: List[AbstractMonitor] = list(monitors) if monitors else []
        self._print_prefix = print_prefix

    def add(self, monitor: AbstractMonitor) -> None:
        self._monitors.append(monitor)

    def print(self, epoch: float, decimal: int = 1) -> None:
        to_print = self._print_prefix + f"Epoch {np.around(epoch, decimal)}"
        for i, monitor in enumerate(self._monitors):
            to_print += (" | " if i == 0 else ", ") + str(monitor)
        print(to_print)

    def record(self, epoch: float) -> Dict:
        return {monitor: monitor.record(epoch) for monitor in self._monitors}

    def export(self) -> None:
        for monitor in self._monitors:
            monitor.export()
