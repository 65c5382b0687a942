"""The plain FP32 straight-translation form of the dense path -- the measured baseline.

It keeps the reference's structure step for step: the same layer-level sequence as
bats/Layers/LIFLayer.py:63-110 (exps kernel, a device-to-host read of the largest spike
count, a per-sample sort of all slots, gathers, the materialised [B, J, P] spike weights,
the one-thread-per-sample solver, the per-sample [B, J, I] gradient, the [B, I, K_pre, J]
pre-spike error terms summed afterwards), the reference's loss and Adam as separate
elementwise operations, and the caller's batch mean. Its kernels are in
csrc/plain_form.hip. It is used by bench.py to measure the optimised path against the
reference's structure on the same GPU, never by the optimised path itself.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib
from ._constants import lif_constants
from .AbstractLayer import AbstractLayer
from .ops import _ptr, _stream, as_device

_vp, _i, _f = C.c_void_p, C.c_int, C.c_float
PLAIN_SIGNATURES = {
    "bats_plain_pre_exps": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _f, _f]),
    "bats_plain_spike_weights": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp]),
    "bats_plain_solve": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _vp, _vp,
                              _vp, _vp, _vp]),
    "bats_plain_recurrent": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _f]),
    "bats_plain_factors": (_i, [_vp, _i, _i, _i, _vp, _vp, _f, _vp, _vp, _f, _vp, _vp]),
    "bats_plain_wgrad": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "bats_plain_pre_errors": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                   _vp, _vp, _f, _f]),
}


def _lib_plain():
    lib = _lib.load()
    if not getattr(lib, "_plain_declared", False):
        for name, (res, args) in PLAIN_SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib._plain_declared = True
    return lib


class PlainLIFLayer(AbstractLayer):
    """Same constructor and methods as the reference's LIFLayer; returns the reference's
    per-sample [B, J, I] gradients."""

    def __init__(self, previous_layer: AbstractLayer, tau_s: float, theta: float,
                 delta_theta: float, weight_initializer=None, max_n_spike: int = 32, **kwargs):
        super().__init__(**kwargs)
        self.__prev = previous_layer
        self.__tau_s, self.__tau, self.__c, self.__dtt = lif_constants(tau_s, theta, delta_theta)
        if weight_initializer is None:
            self.__w = torch.zeros((self.n_neurons, previous_layer.n_neurons),
                                   dtype=torch.float32, device="cuda")
        else:
            self.__w = as_device(weight_initializer(self.n_neurons, previous_layer.n_neurons),
                                 torch.float32)
        self.__K = int(max_n_spike)
        self.reset()

    @property
    def trainable(self) -> bool:
        return True

    @property
    def spike_trains(self):
        return self.__t, self.__n

    @property
    def weights(self):
        return self.__w

    @weights.setter
    def weights(self, w) -> None:
        self.__w = as_device(w, torch.float32).clone()

    def reset(self) -> None:
        self.__n = self.__t = self.__a = self.__x = self.__z = None
        self.__es = self.__e = None

    def forward(self, max_simulation: float, training: bool = False) -> None:
        lib = _lib_plain()
        pre_t, pre_n = self.__prev.spike_trains
        B, N, Kp = pre_t.shape
        J, K = self.n_neurons, self.__K
        dev = pre_t.device
        self.__es = torch.empty_like(pre_t)
        self.__e = torch.empty_like(pre_t)
        _lib.check(lib.bats_plain_pre_exps(_stream(), B, N, Kp, _ptr(pre_t), _ptr(self.__es),
                                           _ptr(self.__e), float(self.__tau_s), float(self.__tau)))
        flat = pre_t.reshape(B, N * Kp)
        P = int(pre_n.sum(dim=1).max().item())  # the reference's device-to-host read
        shape = (B, J, K)
        if P == 0:
            self.__n = torch.zeros((B, J), dtype=torch.int32, device=dev)
            self.__t = torch.full(shape, np.inf, dtype=torch.float32, device=dev)
            self.__z = torch.full(shape, np.inf, dtype=torch.float32, device=dev)
            self.__a = torch.full(shape, np.inf, dtype=torch.float32, device=dev)
            self.__x = torch.full(shape, np.inf, dtype=torch.float32, device=dev)
            return
        order = torch.argsort(flat, dim=1, stable=True)[:, :P].contiguous()
        sidx = (order // Kp).to(torch.int32).contiguous()
        st = torch.take_along_dim(flat, order, dim=1).contiguous()
        ses = torch.take_along_dim(self.__es.reshape(B, N * Kp), order, dim=1).contiguous()
        se = torch.take_along_dim(self.__e.reshape(B, N * Kp), order, dim=1).contiguous()
        sw = torch.empty((B, J, P), dtype=torch.float32, device=dev)
        _lib.check(lib.bats_plain_spike_weights(_stream(), B, P, J, N, _ptr(self.__w), _ptr(sidx),
                                                _ptr(sw)))
        self.__n = torch.zeros((B, J), dtype=torch.int32, device=dev)
        self.__a = torch.empty(shape, dtype=torch.float32, device=dev)
        self.__x = torch.empty(shape, dtype=torch.float32, device=dev)
        self.__t = torch.full(shape, np.inf, dtype=torch.float32, device=dev)
        self.__z = torch.full(shape, np.inf, dtype=torch.float32, device=dev)
        _lib.check(lib.bats_plain_solve(_stream(), B, P, J, K, _ptr(st), _ptr(ses), _ptr(se),
                                        _ptr(sw), float(self.__c), float(self.__dtt),
                                        float(self.__tau), float(np.float32(max_simulation)),
                                        _ptr(self.__n), _ptr(self.__a), _ptr(self.__x),
                                        _ptr(self.__t), _ptr(self.__z)))

    def backward(self, errors: torch.Tensor):
        lib = _lib_plain()
        pre_t, _ = self.__prev.spike_trains
        B, J, K = errors.shape
        _, I, Kp = pre_t.shape
        _lib.check(lib.bats_plain_recurrent(_stream(), B, J, K, _ptr(self.__x), _ptr(self.__z),
                                            _ptr(errors), float(self.__dtt)))
        f1 = torch.empty_like(errors)
        f2 = torch.empty_like(errors)
        _lib.check(lib.bats_plain_factors(_stream(), B, J, K, _ptr(self.__t), _ptr(self.__a),
                                          float(self.__c), _ptr(self.__x), _ptr(self.__z),
                                          float(self.__tau), _ptr(f1), _ptr(f2)))
        grad = torch.zeros((B, J, I), dtype=torch.float32, device=errors.device)
        _lib.check(lib.bats_plain_wgrad(_stream(), B, J, I, K, Kp, _ptr(f1), _ptr(f2),
                                        _ptr(self.__t), _ptr(pre_t), _ptr(self.__es),
                                        _ptr(self.__e), _ptr(errors), _ptr(grad)))
        pre = None
        if self.__prev.trainable:
            terms = torch.zeros((B, I, Kp, J), dtype=torch.float32, device=errors.device)
            _lib.check(lib.bats_plain_pre_errors(_stream(), B, I, Kp, J, K, _ptr(f1), _ptr(f2),
                                                 _ptr(self.__t), _ptr(pre_t), _ptr(self.__es),
                                                 _ptr(self.__e), _ptr(self.__w), _ptr(errors),
                                                 _ptr(terms), float(self.__tau_s),
                                                 float(self.__tau)))
            pre = terms.sum(dim=3)
        return grad, pre

    def add_deltas(self, delta) -> None:
        self.__w += delta


class PlainSpikeCountClassLoss:
    """bats/Losses/SpikeCountClassLoss.py + SpikeCountLoss.py as separate float32 elementwise
    and reduction operations."""

    def __init__(self, target_true, target_false):
        self.tt = float(np.float32(target_true))
        self.tf = float(np.float32(target_false))

    def compute_loss_and_errors(self, spikes, n, labels):
        B, O, K = spikes.shape
        targets = torch.full((B, O), self.tf, dtype=torch.float32, device=n.device)
        targets[torch.arange(B, device=n.device), labels] = self.tt
        d = targets - n.to(torch.float32)
        loss = (d * d).sum(dim=1) / 2
        errors = d[:, :, None].repeat(1, 1, K).contiguous()
        return loss, errors


class PlainAdam:
    """bats/Optimizers/AdamOptimizer.py:35-55 as separate float32 elementwise operations."""

    def __init__(self, learning_rate, beta_1=0.9, beta_2=0.999, epsilon=1e-8):
        self.lr, self.b1, self.b2, self.eps = learning_rate, beta_1, beta_2, epsilon
        self.m = self.v = None
        self.t = 0

    def step(self, grads):
        self.t += 1
        if self.m is None:
            self.m = [None if g is None else torch.zeros_like(g) for g in grads]
            self.v = [None if g is None else torch.zeros_like(g) for g in grads]
        self.m = [None if g is None else self.b1 * m + (1 - self.b1) * g for m, g in zip(self.m, grads)]
        self.v = [None if g is None else self.b2 * v + (1 - self.b2) * g * g for v, g in zip(self.v, grads)]
        s1 = 1 - np.float64(np.float32(self.b1)) ** self.t
        s2 = 1 - np.float64(np.float32(self.b2)) ** self.t
        mh = [None if m is None else m / s1 for m in self.m]
        vh = [None if v is None else v / s2 for v in self.v]
        return [None if g is None else -(self.lr * m / (torch.sqrt(v) + self.eps))
                for g, m, v in zip(grads, mh, vh)]
