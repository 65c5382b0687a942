"""Spike-count class loss (bats/Losses/SpikeCountClassLoss.py:8-33): target_true spikes for
the label neuron, target_false for the others."""
from __future__ import annotations

from typing import Tuple

import numpy as np
import torch

from .. import ops
from .SpikeCountLoss import SpikeCountLoss


class SpikeCountClassLoss(SpikeCountLoss):
    def __init__(self, target_true: float, target_false: float):
        super().__init__()
        self.__target_true = np.float32(target_true)
        self.__target_false = np.float32(target_false)

    def compute_loss(self, spikes_per_neuron, n_spike_per_neuron, labels) -> torch.Tensor:
        loss, _, _ = ops.count_loss(n_spike_per_neuron, spikes_per_neuron.shape[2], labels,
                                    self.__target_true, self.__target_false, want_errors=False)
        return loss

    def compute_errors(self, spikes_per_neuron, n_spike_per_neuron, labels) -> torch.Tensor:
        _, err, _ = ops.count_loss(n_spike_per_neuron, spikes_per_neuron.shape[2], labels,
                                   self.__target_true, self.__target_false, want_loss=False)
        return err

    def compute_loss_and_errors(self, spikes_per_neuron, n_spike_per_neuron,
                                labels) -> Tuple[torch.Tensor, torch.Tensor]:
        loss, err, _ = ops.count_loss(n_spike_per_neuron, spikes_per_neuron.shape[2], labels,
                                      self.__target_true, self.__target_false)
        return loss, err
