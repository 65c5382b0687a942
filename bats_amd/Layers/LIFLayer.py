"""Dense LIF layer (bats/Layers/LIFLayer.py:11-113) on the HIP forward/backward."""
from __future__ import annotations

from typing import Callable, Optional, Tuple

import numpy as np
import torch

from .. import ops
from .._constants import lif_constants
from ..AbstractLayer import AbstractLayer


class LIFLayer(AbstractLayer):
    """Same constructor and methods as the reference layer.

    ``backward`` returns the weight gradient already averaged over the batch, shaped
    [1, J, I], so the reference's caller-side ``mean(axis=0)`` leaves it unchanged; set
    ``per_sample_gradient=True`` to get the reference's [B, J, I] per-sample gradients.
    """

    def __init__(self, previous_layer: AbstractLayer, tau_s: float, theta: float,
                 delta_theta: float, weight_initializer: Callable[[int, int], object] = None,
                 max_n_spike: int = 32, per_sample_gradient: bool = False, **kwargs):
        super().__init__(**kwargs)
        self.__previous_layer = previous_layer
        self.__tau_s, self.__tau, self.__c, self.__delta_theta_tau = lif_constants(
            tau_s, theta, delta_theta)
        if weight_initializer is None:
            self.__weights = torch.zeros((self.n_neurons, previous_layer.n_neurons),
                                         dtype=torch.float32, device="cuda")
        else:
            self.__weights = ops.as_device(weight_initializer(self.n_neurons,
                                                              previous_layer.n_neurons),
                                           torch.float32)
        self.__max_n_spike = int(max_n_spike)
        self.per_sample_gradient = per_sample_gradient
        self.reset()

    @property
    def trainable(self) -> bool:
        return True

    @property
    def spike_trains(self) -> Tuple[torch.Tensor, torch.Tensor]:
        return self.__spike_times_per_neuron, self.__n_spike_per_neuron

    @property
    def weights(self) -> Optional[torch.Tensor]:
        return self.__weights

    @weights.setter
    def weights(self, weights) -> None:
        self.__weights = ops.as_device(weights, torch.float32).clone()

    @property
    def max_n_spike(self) -> int:
        return self.__max_n_spike

    @property
    def constants(self):
        """(tau_s, tau, c = theta/tau, delta_theta/tau) as float32."""
        return self.__tau_s, self.__tau, self.__c, self.__delta_theta_tau

    @property
    def previous_layer(self) -> AbstractLayer:
        return self.__previous_layer

    def reset(self) -> None:
        self.__n_spike_per_neuron = None
        self.__spike_times_per_neuron = None
        self.__a = None
        self.__x = None
        self.__post_exp_tau = None
        self.__pre_exp_tau_s = None
        self.__pre_exp_tau = None
        self.__event_lists = None

    def accept_event_lists(self, lists: Optional[ops.EventLists]) -> None:
        """Pre exps and event lists of this layer's next inputs, made by the previous layer's
        forward (used once, if they still match the inputs and constants)."""
        self.__event_lists = lists

    def forward(self, max_simulation: float, training: bool = False,
                next_layer: Optional["LIFLayer"] = None) -> None:
        """``next_layer`` (given by Network.forward when the next layer is a dense LIF layer
        fed by this one): this call also makes that layer's pre exps and event lists, so its
        forward needs no per-sample sort; the results are the same."""
        pre_times, _ = self.__previous_layer.spike_trains
        lists, self.__event_lists = self.__event_lists, None
        consts = (self.__tau_s, self.__tau, self.__c, self.__delta_theta_tau)
        chain = None
        if next_layer is not None:
            nts, nt, nc, _ = next_layer.constants
            chain = (nts, nt, nc, max_simulation)
        nxt = None
        key = (float(self.__tau_s), float(self.__tau), float(self.__c),
               float(np.float32(max_simulation)))
        if lists is not None and lists.matches(pre_times, key):
            res = ops.lif_forward_events(lists, self.__weights, *consts, self.__max_n_spike,
                                         chain=chain)
            if chain is not None:
                res, nxt = res
        elif chain is not None:
            res, nxt = ops.lif_forward_chain(pre_times, self.__weights, *consts, max_simulation,
                                             self.__max_n_spike, chain)
        else:
            res = ops.lif_forward(pre_times, self.__weights, *consts, max_simulation,
                                  self.__max_n_spike)
        (self.__n_spike_per_neuron, self.__spike_times_per_neuron, self.__a, self.__x,
         self.__post_exp_tau, self.__pre_exp_tau_s, self.__pre_exp_tau) = res
        if next_layer is not None:
            next_layer.accept_event_lists(nxt)

    def backward(self, errors: torch.Tensor) -> Optional[Tuple[torch.Tensor, torch.Tensor]]:
        """``errors`` [B,J,K] is updated in place with the reset-chain errors (as the
        reference's propagate_recurrent_errors does) when it is a contiguous float32
        device tensor; otherwise a converted copy is used."""
        pre_times, _ = self.__previous_layer.spike_trains
        if not (isinstance(errors, torch.Tensor) and errors.is_cuda
                and errors.dtype == torch.float32 and errors.is_contiguous()):
            errors = ops.as_device(errors, torch.float32)
        grad, pre_errors = ops.lif_backward(
            self.__n_spike_per_neuron, self.__spike_times_per_neuron, self.__a, self.__x,
            self.__post_exp_tau, errors, pre_times, self.__pre_exp_tau_s, self.__pre_exp_tau,
            self.__weights, self.__tau_s, self.__tau, self.__c, self.__delta_theta_tau,
            per_sample=self.per_sample_gradient,
            pre_trainable=self.__previous_layer.trainable)
        if not self.per_sample_gradient:
            grad = grad.unsqueeze(0)
        return grad, pre_errors

    def add_deltas(self, delta_weights: torch.Tensor) -> None:
        ops.add_deltas(self.__weights, ops.as_device(delta_weights, torch.float32))
