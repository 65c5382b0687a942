"""Training checkpoints: the reference's per-layer weight files plus the optimizer state.

    save(dir_path, network, optimizer=None)
    load(dir_path, network, optimizer=None)

Weights are written as `<layer name>_weights.npy` exactly as the reference's
`Network.store` does (bats/Network.py:56-59, bats/AbstractLayer.py:61-65), so checkpoints of
either implementation load in the other. The optimizer state (Adam / NAdam moments and step
counter, which the reference does not save) goes to `adam_state.npz` in the same directory;
restoring both resumes training bit for bit.
"""
from __future__ import annotations

from pathlib import Path


def save(dir_path, network, optimizer=None) -> None:
    dir_path = Path(dir_path)
    network.store(dir_path)
    if optimizer is not None:
        optimizer.store(dir_path)


def load(dir_path, network, optimizer=None) -> None:
    dir_path = Path(dir_path)
    network.restore(dir_path)
    if optimizer is not None:
        optimizer.restore(dir_path)
