"""Functional operators over torch device tensors, each one C-ABI call into libbats_amd.

These are the replacements of the reference's kernel wrappers
(bats/CudaKernels/Wrappers/**). Every operator allocates its outputs with torch's caching
allocator, launches on torch's current stream and returns without synchronising.
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib

F32 = torch.float32
I32 = torch.int32
I64 = torch.int64


def _require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError("bats_amd operators need a ROCm GPU; none is visible")


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _ptr(t: Optional[torch.Tensor], dtype=None) -> Optional[int]:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


def as_device(x, dtype, device=None) -> torch.Tensor:
    """Host/any tensor -> contiguous device tensor of ``dtype`` (no copy when already so)."""
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device).contiguous()


def _workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


# ---------------------------------------------------------------- dense LIF
def lif_forward(pre_times: torch.Tensor, weights: torch.Tensor, tau_s: np.float32,
                tau: np.float32, c: np.float32, delta_theta_tau: np.float32,
                max_simulation: float, max_n_post: int, solver_mode: int = 0,
                solver_blocks: int = 0):
    """Dense LIF forward (bats/Layers/LIFLayer.py:63-91).

    Returns (n [B,J] i32, t [B,J,K] f32 +inf padded, a, x, z [B,J,K], pre_exp_tau_s,
    pre_exp_tau [B,I,K_pre]). ``solver_mode`` / ``solver_blocks`` force a solver form and the
    lane solver's workgroups per sample for this call (0: automatic; see
    bats_lif_forward_solver); every form gives the same results.
    """
    _require_gpu()
    B, I, Kp = pre_times.shape
    J, I2 = weights.shape
    if I2 != I:
        raise ValueError(f"weights are [{J},{I2}] but the previous layer has {I} neurons")
    K = int(max_n_post)
    dev = pre_times.device
    n = torch.empty((B, J), dtype=I32, device=dev)
    t = torch.empty((B, J, K), dtype=F32, device=dev)
    a = torch.empty((B, J, K), dtype=F32, device=dev)
    x = torch.empty((B, J, K), dtype=F32, device=dev)
    z = torch.empty((B, J, K), dtype=F32, device=dev)
    es = torch.empty((B, I, Kp), dtype=F32, device=dev)
    e = torch.empty((B, I, Kp), dtype=F32, device=dev)
    lib = _lib.load()
    ws_bytes = lib.bats_lif_forward_workspace_size(B, I, Kp, J)
    ws = _workspace(ws_bytes, dev)
    _lib.check(lib.bats_lif_forward_solver(
        _stream(), B, I, Kp, J, K, _ptr(pre_times, F32), _ptr(weights, F32),
        float(tau_s), float(tau), float(c), float(delta_theta_tau), float(np.float32(max_simulation)),
        _ptr(n), _ptr(t), _ptr(a), _ptr(x), _ptr(z), _ptr(es), _ptr(e), int(solver_mode),
        int(solver_blocks), _ptr(ws), ws_bytes))
    return n, t, a, x, z, es, e


_CU_COUNT = {}


def chain_fuses(batch: int, n_neurons: int) -> bool:
    """Whether a chained dense layer of ``n_neurons`` builds the next layer's event lists inside
    its solver (one workgroup per sample: at least as many samples as the device has compute
    units, at most 1024 neurons); otherwise the chained call runs the per-sample sort, which is
    what two plain calls do."""
    dev = torch.cuda.current_device()
    if dev not in _CU_COUNT:
        _CU_COUNT[dev] = torch.cuda.get_device_properties(dev).multi_processor_count
    return batch >= _CU_COUNT[dev] and n_neurons <= 1024


class EventLists:
    """The next dense layer's pre exps and time-ordered event lists, made by the layer before
    it (:func:`lif_forward_chain`, or :func:`lif_forward_events` with ``chain``) from its
    spike trains ``source`` [B, I, K_pre]; consumed by :func:`lif_forward_events`. It belongs
    to those spike trains as they were: :meth:`matches` is False once they were written to
    (their tensor version changed) or for other constants."""

    __slots__ = ("buf", "nbytes", "B", "I", "K_pre", "es", "e", "source", "version", "key")

    def __init__(self, buf, nbytes, source, es, e, key):
        self.buf, self.nbytes = buf, int(nbytes)
        self.B, self.I, self.K_pre = (int(v) for v in source.shape)
        self.es, self.e = es, e
        self.source, self.version = source, source._version
        self.key = key

    def matches(self, pre_times: torch.Tensor, key) -> bool:
        return (pre_times is self.source and pre_times._version == self.version
                and key == self.key)


def _chain_args(chain, B, J, K, dev):
    """(next tau_s, tau, c, max_simulation, next exps es/e, list buffer, bytes, key) of the
    layer after a [B, J, K] layer; ``chain`` = (tau_s, tau, c, max_simulation) of that layer."""
    tau_s2, tau2, c2, tmax2 = (float(np.float32(v)) for v in chain)
    es2 = torch.empty((B, J, K), dtype=F32, device=dev)
    e2 = torch.empty((B, J, K), dtype=F32, device=dev)
    lib = _lib.load()
    nbytes = int(lib.bats_event_lists_size(B, J * K))
    buf = _workspace(nbytes, dev)
    return tau_s2, tau2, c2, tmax2, es2, e2, buf, nbytes, (tau_s2, tau2, c2, tmax2)


def lif_forward_chain(pre_times: torch.Tensor, weights: torch.Tensor, tau_s, tau, c,
                      delta_theta_tau, max_simulation: float, max_n_post: int, chain):
    """:func:`lif_forward` of a layer whose spike trains feed a dense layer with constants
    ``chain`` = (tau_s, tau, c, max_simulation): also makes that layer's pre exps and event
    lists (bats_lif_forward_chain). Returns the 7 outputs of :func:`lif_forward` and an
    :class:`EventLists` for :func:`lif_forward_events`."""
    _require_gpu()
    B, I, Kp = pre_times.shape
    J, I2 = weights.shape
    if I2 != I:
        raise ValueError(f"weights are [{J},{I2}] but the previous layer has {I} neurons")
    K = int(max_n_post)
    dev = pre_times.device
    n = torch.empty((B, J), dtype=I32, device=dev)
    t = torch.empty((B, J, K), dtype=F32, device=dev)
    a = torch.empty((B, J, K), dtype=F32, device=dev)
    x = torch.empty((B, J, K), dtype=F32, device=dev)
    z = torch.empty((B, J, K), dtype=F32, device=dev)
    es = torch.empty((B, I, Kp), dtype=F32, device=dev)
    e = torch.empty((B, I, Kp), dtype=F32, device=dev)
    tau_s2, tau2, c2, tmax2, es2, e2, buf, nbytes, key = _chain_args(chain, B, J, K, dev)
    lib = _lib.load()
    ws_bytes = lib.bats_lif_forward_workspace_size(B, I, Kp, J)
    ws = _workspace(ws_bytes, dev)
    _lib.check(lib.bats_lif_forward_chain(
        _stream(), B, I, Kp, J, K, _ptr(pre_times, F32), _ptr(weights, F32), float(tau_s),
        float(tau), float(c), float(delta_theta_tau), float(np.float32(max_simulation)),
        _ptr(n), _ptr(t), _ptr(a), _ptr(x), _ptr(z), _ptr(es), _ptr(e), tau_s2, tau2, c2, tmax2,
        _ptr(es2), _ptr(e2), _ptr(buf), nbytes, _ptr(ws), ws_bytes))
    return (n, t, a, x, z, es, e), EventLists(buf, nbytes, t, es2, e2, key)


def lif_forward_events(lists: EventLists, weights: torch.Tensor, tau_s, tau, c,
                       delta_theta_tau, max_n_post: int, chain=None):
    """Dense LIF forward on event lists made by the layer before it (bats_lif_forward_events).
    Returns the 7 outputs of :func:`lif_forward` (the pre exps are ``lists.es`` / ``lists.e``)
    and, with ``chain`` (the next layer's constants, as for :func:`lif_forward_chain`), the
    next layer's :class:`EventLists` as an 8th value."""
    _require_gpu()
    B, I, Kp = lists.B, lists.I, lists.K_pre
    J, I2 = weights.shape
    if I2 != I:
        raise ValueError(f"weights are [{J},{I2}] but the previous layer has {I} neurons")
    K = int(max_n_post)
    dev = weights.device
    n = torch.empty((B, J), dtype=I32, device=dev)
    t = torch.empty((B, J, K), dtype=F32, device=dev)
    a = torch.empty((B, J, K), dtype=F32, device=dev)
    x = torch.empty((B, J, K), dtype=F32, device=dev)
    z = torch.empty((B, J, K), dtype=F32, device=dev)
    lib = _lib.load()
    ws_bytes = lib.bats_lif_forward_events_workspace_size(I, J)
    ws = _workspace(ws_bytes, dev)
    if chain is not None:
        tau_s2, tau2, c2, tmax2, es2, e2, buf, nbytes, key = _chain_args(chain, B, J, K, dev)
        nxt = (tau_s2, tau2, c2, tmax2, _ptr(es2), _ptr(e2), _ptr(buf), nbytes)
    else:
        nxt = (1.0, 1.0, 0.0, 0.0, None, None, None, 0)
    _lib.check(lib.bats_lif_forward_events(
        _stream(), B, I, Kp, J, K, _ptr(lists.buf), lists.nbytes, _ptr(weights, F32),
        float(tau_s), float(tau), float(c), float(delta_theta_tau), _ptr(n), _ptr(t), _ptr(a),
        _ptr(x), _ptr(z), *nxt, _ptr(ws), ws_bytes))
    out = (n, t, a, x, z, lists.es, lists.e)
    if chain is not None:
        return out, EventLists(buf, nbytes, t, es2, e2, key)
    return out


def lif_forward_packed(row_offsets: torch.Tensor, indices: torch.Tensor, times: torch.Tensor,
                       P_max: int, weights: torch.Tensor, tau_s, tau, c, delta_theta_tau,
                       max_simulation: float, max_n_post: int, out=None, chain=None):
    """Dense LIF forward of a first layer fed by packed, time-ordered event lists (inference):
    sample b's events are ``indices/times[row_offsets[b]:row_offsets[b+1]]``, at most
    ``P_max`` each; an index outside [0, I) or a non-finite time ends a list. Returns
    (n, t, a, x, z) as :func:`lif_forward` (no pre exps). ``out`` may hold those five tensors
    (written in place, for graph capture). With ``chain`` (the next dense layer's constants, as
    for :func:`lif_forward_chain`; bats_lif_forward_packed_chain) it also returns that layer's
    :class:`EventLists`."""
    _require_gpu()
    dev = weights.device
    off = as_device(row_offsets, I32, dev)
    idx = as_device(indices, I32, dev)
    tim = as_device(times, F32, dev)
    B = int(off.shape[0]) - 1
    J, I = int(weights.shape[0]), int(weights.shape[1])
    K = int(max_n_post)
    if out is None:
        out = (torch.empty((B, J), dtype=I32, device=dev),
               torch.empty((B, J, K), dtype=F32, device=dev),
               torch.empty((B, J, K), dtype=F32, device=dev),
               torch.empty((B, J, K), dtype=F32, device=dev),
               torch.empty((B, J, K), dtype=F32, device=dev))
    n, t, a, x, z = out
    lib = _lib.load()
    ws_bytes = lib.bats_lif_forward_packed_workspace_size(B, I, int(P_max), J)
    ws = _workspace(ws_bytes, dev)
    if chain is not None:
        tau_s2, tau2, c2, tmax2, es2, e2, buf, nbytes, key = _chain_args(chain, B, J, K, dev)
        _lib.check(lib.bats_lif_forward_packed_chain(
            _stream(), B, I, J, K, int(P_max), _ptr(off), _ptr(idx), _ptr(tim),
            _ptr(weights, F32), float(tau_s), float(tau), float(c), float(delta_theta_tau),
            float(max_simulation), _ptr(n), _ptr(t), _ptr(a), _ptr(x), _ptr(z), tau_s2, tau2,
            c2, tmax2, _ptr(es2), _ptr(e2), _ptr(buf), nbytes, _ptr(ws), ws_bytes))
        return (n, t, a, x, z), EventLists(buf, nbytes, t, es2, e2, key)
    _lib.check(lib.bats_lif_forward_packed(
        _stream(), B, I, J, K, int(P_max), _ptr(off), _ptr(idx), _ptr(tim),
        _ptr(weights, F32), float(tau_s), float(tau), float(c), float(delta_theta_tau),
        float(max_simulation), _ptr(n), _ptr(t), _ptr(a), _ptr(x), _ptr(z), _ptr(ws),
        ws_bytes))
    return n, t, a, x, z


def lif_backward(n, t, a, x, z, errors, pre_times, pre_es, pre_e, weights, tau_s, tau, c,
                 delta_theta_tau, per_sample: bool = False, pre_trainable: bool = True,
                 batch_sum: bool = False):
    """Dense LIF BATS backward (bats/Layers/LIFLayer.py:93-110). ``errors`` is updated in
    place with the reset-chain errors. Returns (grad, pre_errors | None); grad is the batch
    mean [J,I] (the batch sum with ``batch_sum``) unless ``per_sample`` ([B,J,I])."""
    _require_gpu()
    B, J, K = t.shape
    _, I, Kp = pre_times.shape
    dev = t.device
    grad = torch.empty((B, J, I) if per_sample else (J, I), dtype=F32, device=dev)
    mode = 1 if per_sample else (2 if batch_sum else 0)
    pre_err = torch.empty((B, I, Kp), dtype=F32, device=dev) if pre_trainable else None
    lib = _lib.load()
    ws_bytes = lib.bats_lif_backward_workspace_size(B, I, Kp, J, K)
    ws = _workspace(ws_bytes, dev)
    _lib.check(lib.bats_lif_backward(
        _stream(), B, I, Kp, J, K, _ptr(n, I32), _ptr(t, F32), _ptr(a, F32), _ptr(x, F32),
        _ptr(z, F32), _ptr(errors, F32), _ptr(pre_times, F32), _ptr(pre_es, F32),
        _ptr(pre_e, F32), _ptr(weights, F32), float(tau_s), float(tau), float(c),
        float(delta_theta_tau), _ptr(grad), mode, _ptr(pre_err), _ptr(ws),
        ws_bytes))
    return grad, pre_err


# ---------------------------------------------------------------- losses
def count_loss(n: torch.Tensor, K: int, labels=None, target_true=None, target_false=None,
               targets=None, want_loss=True, want_errors=True, want_pred=False):
    """Spike-count loss (bats/Losses/SpikeCountLoss.py:17-38); targets either from labels
    (SpikeCountClassLoss.py:15-18) or given [B,O]."""
    _require_gpu()
    B, O = n.shape
    dev = n.device
    loss = torch.empty((B,), dtype=F32, device=dev) if want_loss else None
    err = torch.empty((B, O, int(K)), dtype=F32, device=dev) if want_errors else None
    pred = torch.empty((B,), dtype=I64, device=dev) if want_pred else None
    lib = _lib.load()
    if targets is None:
        lab = as_device(labels, I64, dev)
        _lib.check(lib.bats_count_loss(_stream(), B, O, int(K), _ptr(n, I32), _ptr(lab),
                                       float(np.float32(target_true)),
                                       float(np.float32(target_false)), _ptr(loss), _ptr(err),
                                       _ptr(pred)))
    else:
        tg = as_device(targets, F32, dev)
        _lib.check(lib.bats_count_loss_targets(_stream(), B, O, int(K), _ptr(n, I32), _ptr(tg),
                                               _ptr(loss), _ptr(err), _ptr(pred)))
    return loss, err, pred


def ttfs_loss(t: torch.Tensor, labels=None, tau: float = 1.0, want_loss=True,
              want_errors=True, want_pred=False):
    """Softmax cross-entropy on first spike times (bats/Losses/TTFSSoftmaxCrossEntropy.py:20-56)."""
    _require_gpu()
    B, O, K = t.shape
    dev = t.device
    lab = as_device(labels, I64, dev) if labels is not None else None
    loss = torch.empty((B,), dtype=F32, device=dev) if want_loss else None
    err = torch.empty((B, O, K), dtype=F32, device=dev) if want_errors else None
    pred = torch.empty((B * K,), dtype=I64, device=dev) if want_pred else None
    lib = _lib.load()
    _lib.check(lib.bats_ttfs_loss(_stream(), B, O, K, _ptr(t, F32), _ptr(lab),
                                  float(np.float32(tau)), _ptr(loss), _ptr(err), _ptr(pred)))
    return loss, err, pred


# ---------------------------------------------------------------- optimizers
def adam_step(grad, m, v, lr, beta_1, omb1, beta_2, omb2, eps, bias1, bias2, nadam=False,
              weights=None, want_delta=True, step_dev: Optional[torch.Tensor] = None):
    """Adam / NAdam deltas (bats/Optimizers/AdamOptimizer.py:35-55); with ``step_dev`` (an
    int64 device scalar holding t) the bias corrections are computed on the device."""
    _require_gpu()
    delta = torch.empty_like(grad) if want_delta else None
    lib = _lib.load()
    w = _ptr(weights, F32) if weights is not None else None
    if step_dev is not None:
        _lib.check(lib.bats_adam_step_dev(_stream(), grad.numel(), _ptr(grad, F32), _ptr(m, F32),
                                          _ptr(v, F32), _ptr(delta), w, float(lr), float(beta_1),
                                          float(omb1), float(beta_2), float(omb2), float(eps),
                                          _ptr(step_dev, I64), 1 if nadam else 0))
        return delta
    fn = lib.bats_nadam_step if nadam else lib.bats_adam_step
    _lib.check(fn(_stream(), grad.numel(), _ptr(grad, F32), _ptr(m, F32), _ptr(v, F32),
                  _ptr(delta), w, float(lr), float(beta_1), float(omb1), float(beta_2),
                  float(omb2), float(eps), float(bias1), float(bias2)))
    return delta


def timing(enable: Optional[bool] = None, reset: bool = False):
    """Library-side HIP-event timing of the main kernels (see bats_timing_* in bats_amd.h)."""
    lib = _lib.load()
    if reset:
        lib.bats_timing_reset()
    if enable is not None:
        lib.bats_timing_enable(1 if enable else 0)


def timing_query(kernel: str):
    """(total milliseconds, launches) recorded for ``kernel`` since the last reset."""
    import ctypes
    lib = _lib.load()
    ms = ctypes.c_double(0.0)
    cnt = ctypes.c_longlong(0)
    lib.bats_timing_query(kernel.encode(), ctypes.byref(ms), ctypes.byref(cnt))
    return ms.value, cnt.value


def sgd_step(grad, lr, weights=None, want_delta=True):
    _require_gpu()
    delta = torch.empty_like(grad) if want_delta else None
    lib = _lib.load()
    _lib.check(lib.bats_sgd_step(_stream(), grad.numel(), _ptr(grad, F32), _ptr(delta),
                                 _ptr(weights, F32) if weights is not None else None, float(lr)))
    return delta


def _ptr_array(tensors):
    import ctypes
    return (ctypes.c_void_p * len(tensors))(*[None if t is None else t.data_ptr() for t in tensors])


MULTI_MAX = 16  # tensors per multi-tensor launch


def adam_step_multi(grads, ms, vs, lr, beta_1, omb1, beta_2, omb2, eps, bias1, bias2,
                    nadam=False, weights=None, want_delta=True,
                    step_dev: Optional[torch.Tensor] = None):
    """:func:`adam_step` over several tensors in one launch per 16 tensors (the per-layer
    loop of bats/Optimizers/AdamOptimizer.py:35-55). Returns the list of deltas (None
    entries when ``want_delta`` is False)."""
    import ctypes
    _require_gpu()
    lib = _lib.load()
    for t in list(grads) + list(ms) + list(vs) + ([] if weights is None else list(weights)):
        _ptr(t, F32)
    deltas = [torch.empty_like(g) if want_delta else None for g in grads]
    for k0 in range(0, len(grads), MULTI_MAX):
        sl = slice(k0, k0 + MULTI_MAX)
        g, m, v = grads[sl], ms[sl], vs[sl]
        sizes = (ctypes.c_int64 * len(g))(*[t.numel() for t in g])
        _lib.check(lib.bats_adam_step_multi(
            _stream(), len(g), sizes, _ptr_array(g), _ptr_array(m), _ptr_array(v),
            _ptr_array(deltas[sl]) if want_delta else None,
            _ptr_array(weights[sl]) if weights is not None else None, float(lr), float(beta_1),
            float(omb1), float(beta_2), float(omb2), float(eps), float(bias1), float(bias2),
            _ptr(step_dev) if step_dev is not None else None, 1 if nadam else 0))
    return deltas


def add_deltas_multi(weights, deltas) -> None:
    """:func:`add_deltas` over several (weights, delta) pairs in one launch per 16 pairs."""
    import ctypes
    _require_gpu()
    lib = _lib.load()
    ds = []
    for w, d in zip(weights, deltas):
        if d.numel() != w.numel():
            raise ValueError("delta and weights differ in size")
        _ptr(w, F32)
        dc = d.reshape(w.shape).contiguous()
        _ptr(dc, F32)
        ds.append(dc)
    for k0 in range(0, len(weights), MULTI_MAX):
        w, d = weights[k0:k0 + MULTI_MAX], ds[k0:k0 + MULTI_MAX]
        sizes = (ctypes.c_int64 * len(w))(*[t.numel() for t in w])
        _lib.check(lib.bats_add_deltas_multi(_stream(), len(w), sizes, _ptr_array(w),
                                             _ptr_array(d)))


def add_deltas(weights: torch.Tensor, delta: torch.Tensor) -> None:
    _require_gpu()
    if delta.numel() != weights.numel():
        raise ValueError("delta and weights differ in size")
    d = delta.reshape(weights.shape).contiguous()
    lib = _lib.load()
    _lib.check(lib.bats_add_deltas(_stream(), weights.numel(), _ptr(weights, F32), _ptr(d, F32)))


# ---------------------------------------------------------------- event lists
def events_to_per_neuron(indices, times, n_neurons: int, max_per_neuron: Optional[int] = None):
    """Event list -> per-neuron buffers (bats/CudaKernels/Wrappers/Inference/
    convert_to_spike_per_neuron.py:20-33). When ``max_per_neuron`` is None the buffer depth
    is the largest per-neuron multiplicity, which needs one device-to-host read (the
    reference does a per-sample host loop for the same number)."""
    _require_gpu()
    idx = as_device(indices, I32)
    tim = as_device(times, F32, idx.device)
    B, P = idx.shape
    N = int(n_neurons)
    if max_per_neuron is None:
        valid = idx >= 0
        ends = torch.where(valid.all(dim=1), torch.full((B,), P, device=idx.device),
                           (~valid).int().argmax(dim=1))
        live = torch.arange(P, device=idx.device)[None, :] < ends[:, None]
        flat = torch.where(live, idx.long() + torch.arange(B, device=idx.device)[:, None] * N,
                           torch.full_like(idx, -1, dtype=torch.long))
        flat = flat[flat >= 0]
        K = int(torch.bincount(flat, minlength=B * N).max().item()) if flat.numel() else 0
    else:
        K = int(max_per_neuron)
    out_t = torch.empty((B, N, K), dtype=F32, device=idx.device)
    out_n = torch.empty((B, N), dtype=I32, device=idx.device)
    lib = _lib.load()
    _lib.check(lib.bats_events_to_per_neuron(_stream(), B, P, N, K, _ptr(idx), _ptr(tim),
                                             _ptr(out_t), _ptr(out_n)))
    return out_t, out_n


# largest input layer of the packed conversion (its per-neuron counters and lane masks live in
# one wave's LDS)
PACKED_EVENTS_MAX_NEURONS = 4096


def events_csr_to_per_neuron(row_offsets, indices, times, n_neurons: int, max_per_neuron: int,
                             out=None):
    """Packed event lists -> per-neuron buffers: sample b's time-ordered events are
    ``indices/times[row_offsets[b]:row_offsets[b+1]]`` (same result as
    :func:`events_to_per_neuron` on the padded [B,P] form; reference:
    bats/CudaKernels/Wrappers/Inference/convert_to_spike_per_neuron.py:20-33).
    Returns (times [B,N,K], counts [B,N], row_max [B]); row_max is each sample's largest
    per-neuron count, events dropped beyond K included. ``out`` may hold those three tensors
    (written in place, for graph capture)."""
    _require_gpu()
    off = as_device(row_offsets, I32)
    idx = as_device(indices, I32, off.device)
    tim = as_device(times, F32, off.device)
    B = int(off.shape[0]) - 1
    N, K = int(n_neurons), int(max_per_neuron)
    if out is None:
        out = (torch.empty((B, N, K), dtype=F32, device=off.device),
               torch.empty((B, N), dtype=I32, device=off.device),
               torch.empty((B,), dtype=I32, device=off.device))
    out_t, out_n, row_max = out
    assert tuple(out_t.shape) == (B, N, K) and tuple(out_n.shape) == (B, N)
    assert tuple(row_max.shape) == (B,)
    lib = _lib.load()
    _lib.check(lib.bats_events_csr_to_per_neuron(_stream(), B, N, K, _ptr(off), _ptr(idx),
                                                 _ptr(tim), _ptr(out_t), _ptr(out_n),
                                                 _ptr(row_max)))
    return out_t, out_n, row_max


# ---------------------------------------------------------------- input encoding
def latency_encode(pixels, time_window: float = 100e-3, max_value: float = 255.0):
    """Latency code of a batch of images on the device (experiments/mnist/Dataset.py:66-71):
    pixels [B, ...] uint8 or float -> (times [B, N, 1] float32, counts [B, N] int32) with
    N = the product of the trailing dims, t = time_window * (1 - px / max_value) and +inf
    where px = 0."""
    _require_gpu()
    if not isinstance(pixels, torch.Tensor):
        pixels = torch.as_tensor(np.asarray(pixels))
    dev = pixels.device if pixels.is_cuda else torch.device("cuda")
    kind = 0 if pixels.dtype == torch.uint8 else 1
    px = pixels.to(dev, torch.uint8 if kind == 0 else F32).contiguous()
    B = int(px.shape[0])
    N = int(px.numel() // max(B, 1))
    times = torch.empty((B, N, 1), dtype=F32, device=dev)
    counts = torch.empty((B, N), dtype=I32, device=dev)
    lib = _lib.load()
    _lib.check(lib.bats_latency_encode(_stream(), px.numel(), _ptr(px), kind,
                                       float(time_window), float(max_value), _ptr(times),
                                       _ptr(counts)))
    return times, counts


# ---------------------------------------------------------------- conv LIF and pooling
def _shape3(v):
    import ctypes
    arr = (ctypes.c_int * 3)(*[int(u) for u in v])
    return arr


def conv_lif_forward(pre_times, pre_shape, weights, tau_s, tau, c, delta_theta_tau,
                     max_simulation, max_n_post):
    """Conv LIF forward (bats/Layers/ConvLIFLayer.py:77-107). weights [C_out, fx, fy, C_in];
    pre_shape (X, Y, C_in). Returns (n, t, a, x, z, pre_exp_tau_s, pre_exp_tau)."""
    _require_gpu()
    B, Npre, Kp = pre_times.shape
    Co, fx, fy, Ci = weights.shape
    X, Y, Cin = (int(v) for v in pre_shape)
    if X * Y * Cin != Npre or Cin != Ci:
        raise ValueError("pre_shape / weights do not match the input spike trains")
    Xo, Yo = X - fx + 1, Y - fy + 1
    J = Xo * Yo * Co
    K = int(max_n_post)
    dev = pre_times.device
    n = torch.empty((B, J), dtype=I32, device=dev)
    t = torch.empty((B, J, K), dtype=F32, device=dev)
    a = torch.empty((B, J, K), dtype=F32, device=dev)
    x = torch.empty((B, J, K), dtype=F32, device=dev)
    z = torch.empty((B, J, K), dtype=F32, device=dev)
    es = torch.empty((B, Npre, Kp), dtype=F32, device=dev)
    e = torch.empty((B, Npre, Kp), dtype=F32, device=dev)
    lib = _lib.load()
    ps, fs = _shape3((X, Y, Cin)), _shape3((fx, fy, Co))
    ws_bytes = lib.bats_conv_lif_forward_workspace_size(B, ps, Kp, fs)
    ws = _workspace(ws_bytes, dev)
    _lib.check(lib.bats_conv_lif_forward(
        _stream(), B, ps, Kp, fs, K, _ptr(pre_times, F32), _ptr(weights, F32), float(tau_s),
        float(tau), float(c), float(delta_theta_tau), float(np.float32(max_simulation)),
        _ptr(n), _ptr(t), _ptr(a), _ptr(x), _ptr(z), _ptr(es), _ptr(e), _ptr(ws), ws_bytes))
    return n, t, a, x, z, es, e


def conv_lif_backward(n, t, a, x, z, errors, pre_times, pre_shape, pre_es, pre_e, weights,
                      tau_s, tau, c, delta_theta_tau, per_sample=False, pre_trainable=True,
                      batch_sum=False):
    """Conv LIF backward (bats/Layers/ConvLIFLayer.py:109-134); ``errors`` updated in place.
    Returns (grad [C_out,fx,fy,C_in] or [B,...], pre_errors | None)."""
    _require_gpu()
    B, J, K = t.shape
    _, Npre, Kp = pre_times.shape
    Co, fx, fy, Ci = weights.shape
    dev = t.device
    grad = torch.empty(((B,) if per_sample else ()) + (Co, fx, fy, Ci), dtype=F32, device=dev)
    pre_err = torch.empty((B, Npre, Kp), dtype=F32, device=dev) if pre_trainable else None
    mode = 1 if per_sample else (2 if batch_sum else 0)
    lib = _lib.load()
    ps, fs = _shape3(pre_shape), _shape3((fx, fy, Co))
    ws_bytes = lib.bats_conv_lif_backward_workspace_size(B, ps, Kp, fs, K)
    ws = _workspace(ws_bytes, dev)
    _lib.check(lib.bats_conv_lif_backward(
        _stream(), B, ps, Kp, fs, K, _ptr(n, I32), _ptr(t, F32), _ptr(a, F32), _ptr(x, F32),
        _ptr(z, F32), _ptr(errors, F32), _ptr(pre_times, F32), _ptr(pre_es, F32),
        _ptr(pre_e, F32), _ptr(weights, F32), float(tau_s), float(tau), float(c),
        float(delta_theta_tau), _ptr(grad), mode, _ptr(pre_err), _ptr(ws), ws_bytes))
    return grad, pre_err


def pool_forward(n, t, pre_shape):
    """2x2 spike pooling (bats/Layers/PoolingLayer.py:49-53). Returns (n, t [B,N,4K], src)."""
    _require_gpu()
    B, Npre, K = t.shape
    X, Y, C = (int(v) for v in pre_shape)
    No = (X // 2) * (Y // 2) * C
    dev = t.device
    out_n = torch.empty((B, No), dtype=I32, device=dev)
    out_t = torch.empty((B, No, 4 * K), dtype=F32, device=dev)
    src = torch.empty((B, No, 4 * K), dtype=I32, device=dev)
    lib = _lib.load()
    _lib.check(lib.bats_pool2x2_forward(_stream(), B, _shape3((X, Y, C)), K, _ptr(n, I32),
                                        _ptr(t, F32), _ptr(out_n), _ptr(out_t), _ptr(src)))
    return out_n, out_t, src


def pool_backward(errors, src, pre_shape, K):
    """Scatter pooled errors back to their source slots (bats/Layers/PoolingLayer.py:55-66)."""
    _require_gpu()
    B = errors.shape[0]
    X, Y, C = (int(v) for v in pre_shape)
    out = torch.empty((B, X * Y * C, int(K)), dtype=F32, device=errors.device)
    lib = _lib.load()
    _lib.check(lib.bats_pool2x2_backward(_stream(), B, _shape3((X, Y, C)), int(K),
                                         _ptr(errors, F32), _ptr(src, I32), _ptr(out)))
    return out
