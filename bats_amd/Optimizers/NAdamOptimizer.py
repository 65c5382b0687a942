"""NAdam (bats/Optimizers/NAdamOptimizer.py:23-58) on the fused HIP kernel."""
from __future__ import annotations

from .AdamOptimizer import AdamOptimizer


class NAdamOptimizer(AdamOptimizer):
    _nadam = True
