"""The CuPy names the reference's training scripts use, over torch device tensors.

The reference's scripts read results back with ``.get()`` (e.g.
experiments/mnist/train_spike_count.py:156-158), average per-sample gradients with
``cp.mean(g, axis=0)`` (:167) and draw initial weights with ``cp.random.uniform`` (:61).
With their imports changed to

    import bats_amd.cupy_compat as cp
    from bats_amd.Monitors import *
    from bats_amd.Layers import InputLayer, LIFLayer
    from bats_amd.Losses import *
    from bats_amd.Network import Network
    from bats_amd.Optimizers import *

the code below the imports runs as written: this module provides the array functions such a
script calls, acting on torch tensors on the current GPU, and importing it gives torch tensors
cupy.ndarray's ``get()`` (a copy into a NumPy array). The library itself never imports it.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

# dtypes are NumPy's, as CuPy's are
float16, float32, float64 = np.float16, np.float32, np.float64
int8, int16, int32, int64, uint8, bool_ = np.int8, np.int16, np.int32, np.int64, np.uint8, np.bool_
inf, nan, pi, e = np.inf, np.nan, np.pi, np.e
ndarray = torch.Tensor

_TORCH_DTYPES = {
    np.dtype(np.float16): torch.float16, np.dtype(np.float32): torch.float32,
    np.dtype(np.float64): torch.float64, np.dtype(np.int8): torch.int8,
    np.dtype(np.int16): torch.int16, np.dtype(np.int32): torch.int32,
    np.dtype(np.int64): torch.int64, np.dtype(np.uint8): torch.uint8,
    np.dtype(np.bool_): torch.bool,
}


def _device() -> torch.device:
    return torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() \
        else torch.device("cpu")


def _dtype(dtype) -> Optional[torch.dtype]:
    if dtype is None or isinstance(dtype, torch.dtype):
        return dtype
    return _TORCH_DTYPES[np.dtype(dtype)]


def _get(self: torch.Tensor, stream=None, order="C", out: Optional[np.ndarray] = None):
    """cupy.ndarray.get: the tensor's values as a NumPy array (into ``out`` if given)."""
    host = self.detach().cpu().numpy()
    if out is not None:
        out[...] = host
        return out
    return np.array(host, order=order, copy=True) if order != "K" else host


if not hasattr(torch.Tensor, "get"):
    torch.Tensor.get = _get


# ---------------------------------------------------------------- creation and transfer
def asarray(a, dtype=None) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        a = a.to(_device())
        return a if dtype is None else a.to(_dtype(dtype))
    return torch.as_tensor(np.asarray(a, dtype=dtype), device=_device())


def array(a, dtype=None, copy=True) -> torch.Tensor:
    t = asarray(a, dtype)
    return t.clone() if copy and isinstance(a, torch.Tensor) else t


def asnumpy(a, stream=None) -> np.ndarray:
    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def _shape(shape):
    return (shape,) if isinstance(shape, (int, np.integer)) else tuple(shape)


def zeros(shape, dtype=float64) -> torch.Tensor:
    return torch.zeros(_shape(shape), dtype=_dtype(dtype), device=_device())


def ones(shape, dtype=float64) -> torch.Tensor:
    return torch.ones(_shape(shape), dtype=_dtype(dtype), device=_device())


def empty(shape, dtype=float64) -> torch.Tensor:
    return torch.empty(_shape(shape), dtype=_dtype(dtype), device=_device())


def full(shape, fill_value, dtype=None) -> torch.Tensor:
    dt = _dtype(dtype) if dtype is not None else _dtype(np.asarray(fill_value).dtype)
    return torch.full(_shape(shape), fill_value, dtype=dt, device=_device())


def arange(*args, dtype=None) -> torch.Tensor:
    return asarray(np.arange(*args, dtype=dtype))


# ---------------------------------------------------------------- reductions and elementwise
def _t(a) -> torch.Tensor:
    return a if isinstance(a, torch.Tensor) else asarray(a)


def _reduce(fn, a, axis, keepdims, dtype=None):
    t = _t(a)
    if dtype is not None:
        t = t.to(_dtype(dtype))
    if axis is None:
        r = fn(t)
        return r.reshape((1,) * t.dim()) if keepdims else r
    return fn(t, dim=axis, keepdim=keepdims)


def mean(a, axis=None, dtype=None, keepdims=False) -> torch.Tensor:
    t = _t(a)
    if dtype is None and not (t.is_floating_point() or t.is_complex()):
        dtype = float64  # as NumPy and CuPy: the mean of integers is float64
    return _reduce(torch.mean, t, axis, keepdims, dtype)


def sum(a, axis=None, dtype=None, keepdims=False) -> torch.Tensor:  # noqa: A001
    return _reduce(torch.sum, a, axis, keepdims, dtype)


def max(a, axis=None, keepdims=False) -> torch.Tensor:  # noqa: A001
    if axis is None:
        return _reduce(torch.amax, a, None, keepdims)
    return torch.amax(_t(a), dim=axis, keepdim=keepdims)


def min(a, axis=None, keepdims=False) -> torch.Tensor:  # noqa: A001
    if axis is None:
        return _reduce(torch.amin, a, None, keepdims)
    return torch.amin(_t(a), dim=axis, keepdim=keepdims)


def argmax(a, axis=None) -> torch.Tensor:
    return torch.argmax(_t(a), dim=axis)


def abs(a) -> torch.Tensor:  # noqa: A001
    return torch.abs(_t(a))


def sqrt(a) -> torch.Tensor:
    return torch.sqrt(_t(a))


def exp(a) -> torch.Tensor:
    return torch.exp(_t(a))


def log(a) -> torch.Tensor:
    return torch.log(_t(a))


def isfinite(a) -> torch.Tensor:
    return torch.isfinite(_t(a))


def isinf(a) -> torch.Tensor:
    return torch.isinf(_t(a))


def isnan(a) -> torch.Tensor:
    return torch.isnan(_t(a))


def where(cond, x, y) -> torch.Tensor:
    c = _t(cond)
    x = x if isinstance(x, torch.Tensor) else torch.as_tensor(x, device=c.device)
    y = y if isinstance(y, torch.Tensor) else torch.as_tensor(y, device=c.device)
    return torch.where(c, x, y)


def maximum(a, b) -> torch.Tensor:
    a, b = _t(a), b if isinstance(b, torch.Tensor) else torch.as_tensor(b, device=_t(a).device)
    return torch.maximum(a, b.to(a.dtype))


def minimum(a, b) -> torch.Tensor:
    a, b = _t(a), b if isinstance(b, torch.Tensor) else torch.as_tensor(b, device=_t(a).device)
    return torch.minimum(a, b.to(a.dtype))


# ---------------------------------------------------------------- random numbers
class _Random:
    """cupy.random's module-level functions over one torch generator on the current GPU."""

    def __init__(self):
        self._gen: Optional[torch.Generator] = None

    def _generator(self) -> torch.Generator:
        if self._gen is None:
            self._gen = torch.Generator(device=_device())
            self._gen.seed()
        return self._gen

    def seed(self, seed: Optional[int] = None) -> None:
        g = self._generator()
        if seed is None:
            g.seed()
        else:
            g.manual_seed(int(seed))

    def uniform(self, low=0.0, high=1.0, size=None, dtype=float64) -> torch.Tensor:
        t = torch.empty(_shape(size) if size is not None else (), dtype=_dtype(dtype),
                        device=_device())
        return t.uniform_(float(low), float(high), generator=self._generator())

    def normal(self, loc=0.0, scale=1.0, size=None, dtype=float64) -> torch.Tensor:
        t = torch.empty(_shape(size) if size is not None else (), dtype=_dtype(dtype),
                        device=_device())
        return t.normal_(float(loc), float(scale), generator=self._generator())

    def rand(self, *size, dtype=float64) -> torch.Tensor:
        return self.uniform(0.0, 1.0, size=size, dtype=dtype)

    def randn(self, *size, dtype=float64) -> torch.Tensor:
        return self.normal(0.0, 1.0, size=size, dtype=dtype)

    def randint(self, low, high=None, size=None, dtype=int64) -> torch.Tensor:
        if high is None:
            low, high = 0, low
        return torch.randint(int(low), int(high), _shape(size) if size is not None else (),
                             generator=self._generator(), dtype=_dtype(dtype), device=_device())


random = _Random()
