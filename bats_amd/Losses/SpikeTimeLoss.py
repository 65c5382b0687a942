"""Spike-time MSE loss (bats/Losses/SpikeTimeLoss.py:7-47). It is in no configuration of the
hot path; it runs as torch elementwise operations on the device."""
from __future__ import annotations

from typing import Tuple

import torch

from .. import ops
from ..AbstractLoss import AbstractLoss


def _sanitise(x: torch.Tensor) -> torch.Tensor:
    return torch.where(torch.isinf(x), torch.ones_like(x), x)


def _targets(targets, device) -> torch.Tensor:
    """The reference replaces the infinite entries of the caller's targets by 1.0 in place
    (SpikeTimeLoss.py:25,32,41); a float32 device tensor given here is changed the same way.
    Returns the sanitised float32 targets on ``device``."""
    if isinstance(targets, torch.Tensor) and targets.dtype == torch.float32 \
            and targets.device == device:
        targets.masked_fill_(torch.isinf(targets), 1.0)
        return targets
    return _sanitise(ops.as_device(targets, torch.float32, device))


class SpikeTimeLoss(AbstractLoss):
    def predict(self, spikes_per_neuron, n_spike_per_neuron) -> torch.Tensor:
        return torch.argmin(spikes_per_neuron[..., 0], dim=1)

    def compute_loss(self, spikes_per_neuron, n_spike_per_neuron, targets) -> torch.Tensor:
        t = _sanitise(spikes_per_neuron)
        tg = _targets(targets, t.device)
        d = tg[..., 0].double() - t[..., 0].double()
        return (0.5 * (d * d).sum(dim=1)).float()

    def compute_errors(self, spikes_per_neuron, n_spike_per_neuron, targets) -> torch.Tensor:
        t = _sanitise(spikes_per_neuron)
        tg = _targets(targets, t.device)
        return (t.double() - tg.double()).float()

    def compute_loss_and_errors(self, spikes_per_neuron, n_spike_per_neuron,
                                targets) -> Tuple[torch.Tensor, torch.Tensor]:
        # the reference uses the un-sanitised times for the errors here (SpikeTimeLoss.py:45)
        tg = _targets(targets, spikes_per_neuron.device)
        errors = (spikes_per_neuron.double() - tg.double()).float()
        return self.compute_loss(spikes_per_neuron, n_spike_per_neuron, tg), errors
