"""Sequential network (bats/Network.py:10-63)."""
from __future__ import annotations

from pathlib import Path
from typing import List, Optional, Tuple

import numpy as np
import torch

from . import ops
from .AbstractLayer import AbstractLayer


def _plain_add_methods():
    """The add_deltas methods that are a plain weights += delta (batched in apply_deltas)."""
    from .Layers.ConvLIFLayer import ConvLIFLayer
    from .Layers.LIFLayer import LIFLayer
    return (LIFLayer.add_deltas, ConvLIFLayer.add_deltas)


class Network:
    def __init__(self):
        self.__layers: List[AbstractLayer] = []
        self.__input_layer = None

    @property
    def layers(self) -> List[AbstractLayer]:
        return self.__layers

    @property
    def output_spike_trains(self) -> Tuple[torch.Tensor, torch.Tensor]:
        return self.__layers[-1].spike_trains

    def add_layer(self, layer: AbstractLayer, input: bool = False) -> None:
        self.__layers.append(layer)
        if input:
            self.__input_layer = layer

    def reset(self):
        for layer in self.__layers:
            layer.reset()

    def forward(self, spikes_per_neuron, n_spikes_per_neuron, max_simulation: float = np.inf,
                training: bool = False) -> None:
        self.__input_layer.set_spike_trains(spikes_per_neuron, n_spikes_per_neuron)
        from .Layers.LIFLayer import LIFLayer
        layers = self.__layers
        for k, layer in enumerate(layers):
            nxt = layers[k + 1] if k + 1 < len(layers) else None
            # a dense LIF layer feeding the next dense LIF layer makes that layer's event
            # lists in its own forward (bats_lif_forward_chain) when its solver can: same
            # results, no sort pass
            if (type(layer) is LIFLayer and type(nxt) is LIFLayer
                    and nxt.previous_layer is layer
                    and ops.chain_fuses(len(spikes_per_neuron), layer.n_neurons)):
                layer.forward(max_simulation, training, next_layer=nxt)
            else:
                layer.forward(max_simulation, training)

    def backward(self, output_errors: torch.Tensor) -> List[Optional[torch.Tensor]]:
        """Reverse layer loop, stopping at the first non-trainable layer (bats/Network.py:38-48)."""
        errors = output_errors
        gradient = []
        for layer in reversed(self.__layers):
            if not layer.trainable:
                gradient.insert(0, None)
                break
            weights_grad, errors = layer.backward(errors)
            gradient.insert(0, weights_grad)
        return gradient

    def apply_deltas(self, deltas: List[Optional[torch.Tensor]]) -> None:
        """Adds each layer's delta to its weights (bats/Network.py:50-54). The LIF and conv
        LIF layers' additions run in one launch; other layers use their own add_deltas."""
        plain = _plain_add_methods()
        ws, ds = [], []
        for layer, delta in zip(self.__layers, deltas):
            if delta is None:
                continue
            if type(layer).add_deltas in plain and isinstance(delta, torch.Tensor) \
                    and delta.is_cuda and delta.dtype == torch.float32:
                ws.append(layer.weights)
                ds.append(delta)
            else:
                layer.add_deltas(delta)
        if ws:
            ops.add_deltas_multi(ws, ds)

    def store(self, dir_path: Path) -> None:
        dir_path = Path(dir_path)
        dir_path.mkdir(parents=True, exist_ok=True)
        for it in self.__layers:
            it.store(dir_path)

    def restore(self, dir_path: Path) -> None:
        for it in self.__layers:
            it.restore(Path(dir_path))
