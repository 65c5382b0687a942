"""Optimizer base class (bats/AbstractOptimizer.py:6-20)."""
from __future__ import annotations

from abc import ABC, abstractmethod
from typing import List

import numpy as np


class AbstractOptimizer(ABC):
    def __init__(self, learning_rate: float):
        self._learning_rate: np.float32 = np.float32(learning_rate)

    @property
    def learning_rate(self) -> np.float32:
        return self._learning_rate

    @learning_rate.setter
    def learning_rate(self, value: float) -> None:
        self._learning_rate = np.float32(value)

    @abstractmethod
    def step(self, gradient: List) -> List:
        pass

    def store(self, dir_path) -> None:
        """Writes the optimizer's state next to the layers' weight files (stateless
        optimizers write nothing). The reference keeps no optimizer state on disk."""

    def restore(self, dir_path) -> None:
        """Reads what `store` wrote (nothing for stateless optimizers)."""
