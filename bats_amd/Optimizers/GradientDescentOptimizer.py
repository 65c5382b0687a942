"""Gradient descent (bats/Optimizers/GradientDescentOptimizer.py:7-9)."""
from __future__ import annotations

from typing import List, Optional

import torch

from .. import ops
from ..AbstractOptimizer import AbstractOptimizer


class GradientDescentOptimizer(AbstractOptimizer):
    def step(self, gradient: List[Optional[torch.Tensor]]) -> List[Optional[torch.Tensor]]:
        return [None if g is None else ops.sgd_step(ops.as_device(g, torch.float32),
                                                    self._learning_rate) for g in gradient]
