"""Layer base class, the API of bats/AbstractLayer.py:11-70 over torch device tensors."""
from __future__ import annotations

from abc import ABC, abstractmethod
from pathlib import Path
from typing import Optional, Tuple

import numpy as np
import torch

WEIGHTS_FILE_SUFFIX = "_weights.npy"


class AbstractLayer(ABC):
    def __init__(self, n_neurons: int, name: str = ""):
        self._n_neurons: int = int(n_neurons)
        self._name: str = name

    @property
    def n_neurons(self) -> int:
        return self._n_neurons

    @property
    def name(self) -> str:
        return self._name

    @property
    @abstractmethod
    def trainable(self) -> bool:
        pass

    @property
    @abstractmethod
    def spike_trains(self) -> Tuple[torch.Tensor, torch.Tensor]:
        pass

    @property
    @abstractmethod
    def weights(self) -> Optional[torch.Tensor]:
        pass

    @weights.setter
    @abstractmethod
    def weights(self, weights) -> None:
        pass

    @abstractmethod
    def reset(self) -> None:
        pass

    @abstractmethod
    def forward(self, max_simulation: float, training: bool = False) -> None:
        pass

    @abstractmethod
    def backward(self, errors: torch.Tensor) -> Optional[Tuple[torch.Tensor, torch.Tensor]]:
        pass

    @abstractmethod
    def add_deltas(self, delta_weights: torch.Tensor) -> None:
        pass

    def store(self, dir_path: Path) -> None:
        """Writes <name>_weights.npy, the reference's checkpoint format (bats/AbstractLayer.py:61-65)."""
        weights = self.weights
        if weights is not None:
            filename = Path(dir_path) / (self._name + WEIGHTS_FILE_SUFFIX)
            np.save(filename, weights.detach().cpu().numpy())

    def restore(self, dir_path: Path) -> None:
        """Reads <name>_weights.npy if present (bats/AbstractLayer.py:67-70)."""
        filename = Path(dir_path) / (self._name + WEIGHTS_FILE_SUFFIX)
        if filename.exists():
            self.weights = np.load(filename, allow_pickle=False)
