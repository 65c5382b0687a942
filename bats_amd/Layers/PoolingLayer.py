"""2x2 spike pooling (bats/Layers/PoolingLayer.py:13-69): each output neuron's spike train
is the time-ordered union of its four children's."""
from __future__ import annotations

from typing import Tuple

import numpy as np
import torch

from .. import ops
from ..AbstractConvLayer import AbstractConvLayer


class PoolingLayer(AbstractConvLayer):
    def __init__(self, previous_layer: AbstractConvLayer, **kwargs):
        prev_x, prev_y, prev_c = (int(v) for v in previous_layer.neurons_shape)
        super().__init__(neurons_shape=np.array([prev_x // 2, prev_y // 2, prev_c], dtype=np.int32),
                         **kwargs)
        self.__prev = previous_layer
        self.reset()

    @property
    def trainable(self) -> bool:
        return True

    @property
    def spike_trains(self) -> Tuple[torch.Tensor, torch.Tensor]:
        return self.__t, self.__n

    @property
    def weights(self):
        return None

    @weights.setter
    def weights(self, weights) -> None:
        pass

    def reset(self) -> None:
        self.__n = self.__t = self.__src = None

    def forward(self, max_simulation: float, training: bool = False) -> None:
        pre_t, pre_n = self.__prev.spike_trains
        self.__n, self.__t, self.__src = ops.pool_forward(pre_n, pre_t, self.__prev.neurons_shape)

    def backward(self, errors: torch.Tensor):
        if not self.__prev.trainable:
            return None, None
        pre_t, _ = self.__prev.spike_trains
        errors = ops.as_device(errors, torch.float32, pre_t.device)
        return None, ops.pool_backward(errors, self.__src, self.__prev.neurons_shape,
                                       pre_t.shape[2])

    def add_deltas(self, delta_weights) -> None:
        pass
