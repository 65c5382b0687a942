"""Spike-time MSE loss (bats/Losses/SpikeTimeLoss.py:7-47). It is in no configuration of the
hot path; it runs as torch elementwise operations on the device."""
from __future__ import annotations

from typing import Tuple

import torch

from .. import ops
from ..AbstractLoss import AbstractLoss


def _sanitise(x: torch.Tensor) -> torch.Tensor:
    return torch.where(torch.isinf(x), torch.ones_like(x), x)


class SpikeTimeLoss(AbstractLoss):
    def predict(self, spikes_per_neuron, n_spike_per_neuron) -> torch.Tensor:
        return torch.argmin(spikes_per_neuron[..., 0], dim=1)

    def compute_loss(self, spikes_per_neuron, n_spike_per_neuron, targets) -> torch.Tensor:
        t = _sanitise(spikes_per_neuron)
        tg = _sanitise(ops.as_device(targets, torch.float32, t.device))
        d = tg[..., 0].double() - t[..., 0].double()
        return (0.5 * (d * d).sum(dim=1)).float()

    def compute_errors(self, spikes_per_neuron, n_spike_per_neuron, targets) -> torch.Tensor:
        t = _sanitise(spikes_per_neuron)
        tg = _sanitise(ops.as_device(targets, torch.float32, t.device))
        return (t.double() - tg.double()).float()

    def compute_loss_and_errors(self, spikes_per_neuron, n_spike_per_neuron,
                                targets) -> Tuple[torch.Tensor, torch.Tensor]:
        # the reference uses the un-sanitised times for the errors here (SpikeTimeLoss.py:45)
        tg = _sanitise(ops.as_device(targets, torch.float32, spikes_per_neuron.device))
        errors = (spikes_per_neuron.double() - tg.double()).float()
        return self.compute_loss(spikes_per_neuron, n_spike_per_neuron, targets), errors
