"""Convolutional LIF layer (bats/Layers/ConvLIFLayer.py:15-137) on the HIP kernels:
valid, stride 1, filters [C_out, fx, fy, C_in]."""
from __future__ import annotations

from typing import Callable, Optional, Tuple

import numpy as np
import torch

from .. import ops
from .._constants import lif_constants
from ..AbstractConvLayer import AbstractConvLayer


class ConvLIFLayer(AbstractConvLayer):
    """Same constructor as the reference: ``filters_shape = (fx, fy, C_out)``,
    ``weight_initializer(C_out, fx, fy, C_in)``. ``backward`` returns the batch-mean filter
    gradient with a unit batch axis, [1, C_out, fx, fy, C_in], unless
    ``per_sample_gradient=True`` ([B, ...], as the reference)."""

    def __init__(self, previous_layer: AbstractConvLayer, filters_shape, tau_s: float,
                 theta: float, delta_theta: float, weight_initializer: Callable = None,
                 max_n_spike: int = 32, per_sample_gradient: bool = False, **kwargs):
        prev_x, prev_y, prev_c = (int(v) for v in previous_layer.neurons_shape)
        fx, fy, fc = (int(v) for v in filters_shape)
        neurons_shape = np.array([prev_x - fx + 1, prev_y - fy + 1, fc], dtype=np.int32)
        super().__init__(neurons_shape=neurons_shape, **kwargs)
        self.__filters = (fx, fy, fc)
        self.__prev = previous_layer
        self.__tau_s, self.__tau, self.__c, self.__dtt = lif_constants(tau_s, theta, delta_theta)
        if weight_initializer is None:
            self.__weights = torch.zeros((fc, fx, fy, prev_c), dtype=torch.float32, device="cuda")
        else:
            self.__weights = ops.as_device(weight_initializer(fc, fx, fy, prev_c), torch.float32)
        self.__K = int(max_n_spike)
        self.per_sample_gradient = per_sample_gradient
        self.reset()

    @property
    def trainable(self) -> bool:
        return True

    @property
    def spike_trains(self) -> Tuple[torch.Tensor, torch.Tensor]:
        return self.__t, self.__n

    @property
    def weights(self) -> Optional[torch.Tensor]:
        return self.__weights

    @weights.setter
    def weights(self, weights) -> None:
        self.__weights = ops.as_device(weights, torch.float32).clone()

    def reset(self) -> None:
        self.__n = self.__t = self.__a = self.__x = self.__z = None
        self.__es = self.__e = None

    def forward(self, max_simulation: float, training: bool = False) -> None:
        pre_t, _ = self.__prev.spike_trains
        (self.__n, self.__t, self.__a, self.__x, self.__z, self.__es,
         self.__e) = ops.conv_lif_forward(pre_t, self.__prev.neurons_shape, self.__weights,
                                          self.__tau_s, self.__tau, self.__c, self.__dtt,
                                          max_simulation, self.__K)

    def backward(self, errors: torch.Tensor):
        pre_t, _ = self.__prev.spike_trains
        if not (isinstance(errors, torch.Tensor) and errors.is_cuda
                and errors.dtype == torch.float32 and errors.is_contiguous()):
            errors = ops.as_device(errors, torch.float32)
        grad, pre = ops.conv_lif_backward(
            self.__n, self.__t, self.__a, self.__x, self.__z, errors, pre_t,
            self.__prev.neurons_shape, self.__es, self.__e, self.__weights, self.__tau_s,
            self.__tau, self.__c, self.__dtt, per_sample=self.per_sample_gradient,
            pre_trainable=self.__prev.trainable)
        if not self.per_sample_gradient:
            grad = grad.unsqueeze(0)
        return grad, pre

    def add_deltas(self, delta_weights) -> None:
        ops.add_deltas(self.__weights, ops.as_device(delta_weights, torch.float32))
