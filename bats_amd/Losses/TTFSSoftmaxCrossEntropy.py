"""Softmax cross-entropy on the time to first spike
(bats/Losses/TTFSSoftmaxCrossEntropy.py:7-56) on the HIP TTFS kernel."""
from __future__ import annotations

from typing import Tuple

import numpy as np
import torch

from .. import ops
from ..AbstractLoss import AbstractLoss


class TTFSSoftmaxCrossEntropy(AbstractLoss):
    def __init__(self, tau: float):
        self.__tau = np.float32(tau)

    def predict(self, spikes_per_neuron, n_spike_per_neuron) -> torch.Tensor:
        _, _, pred = ops.ttfs_loss(spikes_per_neuron, None, self.__tau, want_loss=False,
                                   want_errors=False, want_pred=True)
        return pred

    def compute_loss(self, spikes_per_neuron, n_spike_per_neuron, labels) -> torch.Tensor:
        loss, _, _ = ops.ttfs_loss(spikes_per_neuron, labels, self.__tau, want_errors=False)
        return loss

    def compute_errors(self, spikes_per_neuron, n_spike_per_neuron, labels) -> torch.Tensor:
        _, err, _ = ops.ttfs_loss(spikes_per_neuron, labels, self.__tau, want_loss=False)
        return err

    def compute_loss_and_errors(self, spikes_per_neuron, n_spike_per_neuron,
                                labels) -> Tuple[torch.Tensor, torch.Tensor]:
        loss, err, _ = ops.ttfs_loss(spikes_per_neuron, labels, self.__tau)
        return loss, err
