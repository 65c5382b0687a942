"""Spike-count loss (bats/Losses/SpikeCountLoss.py:7-38) on the HIP count-loss kernel."""
from __future__ import annotations

from typing import Tuple

import torch

from .. import ops
from ..AbstractLoss import AbstractLoss


class SpikeCountLoss(AbstractLoss):
    """Targets are given per output neuron, [B, O]."""

    def predict(self, spikes_per_neuron, n_spike_per_neuron) -> torch.Tensor:
        _, _, pred = ops.count_loss(n_spike_per_neuron, 0, targets=torch.zeros_like(
            n_spike_per_neuron, dtype=torch.float32), want_loss=False, want_errors=False,
            want_pred=True)
        return pred

    def compute_loss(self, spikes_per_neuron, n_spike_per_neuron, targets) -> torch.Tensor:
        loss, _, _ = ops.count_loss(n_spike_per_neuron, spikes_per_neuron.shape[2],
                                    targets=targets, want_errors=False)
        return loss

    def compute_errors(self, spikes_per_neuron, n_spike_per_neuron, targets) -> torch.Tensor:
        _, err, _ = ops.count_loss(n_spike_per_neuron, spikes_per_neuron.shape[2],
                                   targets=targets, want_loss=False)
        return err

    def compute_loss_and_errors(self, spikes_per_neuron, n_spike_per_neuron,
                                targets) -> Tuple[torch.Tensor, torch.Tensor]:
        loss, err, _ = ops.count_loss(n_spike_per_neuron, spikes_per_neuron.shape[2],
                                      targets=targets)
        return loss, err
