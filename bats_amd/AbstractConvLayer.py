"""Convolutional layer base class (bats/AbstractConvLayer.py:11-19)."""
from __future__ import annotations

import numpy as np

from .AbstractLayer import AbstractLayer


class AbstractConvLayer(AbstractLayer):
    def __init__(self, neurons_shape, name: str = ""):
        shape = np.asarray(neurons_shape, dtype=np.int32)
        n_neurons = int(shape[0]) * int(shape[1]) * int(shape[2])
        super().__init__(n_neurons, name)
        self._neurons_shape: np.ndarray = shape

    @property
    def neurons_shape(self) -> np.ndarray:
        """(x, y, c); a host array (the reference keeps a device copy and reads it back)."""
        return self._neurons_shape
