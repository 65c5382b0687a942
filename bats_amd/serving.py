"""Continuous-batching inference server for a dense LIF network on one GPU.

Requests are single samples given as event lists (input neuron index, spike time), of any
length; several spikes per input neuron are allowed. The server keeps the network's weights
and all batch buffers resident on the device and runs batches back to back on two streams:
while one batch computes, the next is assembled from whatever requests have arrived and
admitted, and launched as soon as a stream is free. A request is admitted when a batch slot is
free and its device footprint fits the memory budget; it is released when its result has been
copied out. Each batch replays a HIP graph captured for its padded shape (batch bucket,
events-per-request bucket, spikes-per-neuron bucket), so the host cost per batch is a few
copies and one graph launch.

The inference of a request is independent of the other requests of its batch (each sample is
sorted and solved on its own), so a request's spike trains are the same alone and in any mix.

Reference counterparts: bats/Network.py:32-36 (forward), bats/Losses/SpikeCountLoss.py:17-18
(predict), bats/CudaKernels/Wrappers/Inference/convert_to_spike_per_neuron.py:20-33 (event
list to per-neuron buffers). The reference has no server; this is new.
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib, ops
from ._constants import lif_constants
from .ops import _ptr


@dataclasses.dataclass
class Request:
    rid: int
    indices: np.ndarray  # int32 [P], sorted by time
    times: np.ndarray    # float32 [P], ascending
    arrival: float       # host clock (seconds)
    multiplicity: int = 0
    done: float = 0.0
    pred: int = -1
    counts: Optional[np.ndarray] = None
    spike_times: Optional[np.ndarray] = None


@dataclasses.dataclass
class LayerSpec:
    weights: np.ndarray   # [J, I] float32
    tau_s: float
    theta: float
    delta_theta: float
    max_n_spike: int


def _bucket(v: int, minimum: int) -> int:
    b = minimum
    while b < v:
        b *= 2
    return b


class _GraphBatch:
    """Static buffers + captured graph for one padded batch shape."""

    def __init__(self, server: "InferenceServer", B: int, P: int, K_in: int, stream):
        self.B, self.P, self.K_in = B, P, K_in
        dev = server.device
        self.stream = stream
        self.h_idx = torch.empty((B, P), dtype=torch.int32, pin_memory=True)
        self.h_t = torch.empty((B, P), dtype=torch.float32, pin_memory=True)
        self.h_idx.fill_(-1)
        self.h_t.fill_(float("inf"))
        self.row_len = np.zeros(B, dtype=np.int64)  # events written in each host row
        self.idx = torch.empty((B, P), dtype=torch.int32, device=dev)
        self.t = torch.empty((B, P), dtype=torch.float32, device=dev)
        self.idx.fill_(-1)
        self.t.fill_(float("inf"))
        n_out = server.layers[-1].weights.shape[0]
        self.zero_targets = torch.zeros((B, n_out), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.stream(stream):
            self._run(server)  # warm-up (allocations, lazy init) outside capture
        stream.synchronize()
        with torch.cuda.graph(self.graph, stream=stream):
            self.outputs = self._run(server)
        self.h_n = torch.empty(self.outputs[0].shape, dtype=torch.int32, pin_memory=True)
        self.h_st = torch.empty(self.outputs[1].shape, dtype=torch.float32, pin_memory=True)
        self.h_pred = torch.empty(self.outputs[2].shape, dtype=torch.int64, pin_memory=True)
        self.done = torch.cuda.Event()
        self.requests: List[Request] = []

    def _run(self, server):
        per_t, _ = ops.events_to_per_neuron(self.idx, self.t, server.n_inputs, self.K_in)
        x = per_t
        n = None
        for L, w in zip(server.layers, server.w_dev):
            n, x, *_ = ops.lif_forward(x, w, *lif_constants(L.tau_s, L.theta, L.delta_theta),
                                       server.max_simulation, L.max_n_spike)
        _, _, pred = ops.count_loss(n, x.shape[2], targets=self.zero_targets, want_loss=False,
                                    want_errors=False, want_pred=True)
        return n, x, pred


class InferenceServer:
    """Admits requests, batches them continuously and runs them on the GPU.

    ``memory_budget`` bounds the device bytes held by admitted requests (their padded share of
    the batch buffers); a request that does not fit waits in the queue.
    """

    def __init__(self, layers: Sequence[LayerSpec], n_inputs: int, max_batch: int = 1024,
                 max_simulation: float = 1.0, memory_budget: Optional[int] = None,
                 device=None, min_batch_bucket: int = 8, min_events_bucket: int = 64,
                 min_spikes_bucket: int = 1, n_streams: int = 2):
        self.device = torch.device("cuda", 0) if device is None else torch.device(device)
        self.layers = list(layers)
        self.n_inputs = int(n_inputs)
        self.max_batch = int(max_batch)
        self.max_simulation = float(max_simulation)
        self.min_batch_bucket = int(min_batch_bucket)
        # larger minimum buckets mean fewer distinct padded shapes (graphs) at the cost of
        # padding work; a server with a known request mix sets them to its upper ends
        self.min_events_bucket = int(min_events_bucket)
        self.min_spikes_bucket = int(min_spikes_bucket)
        self.w_dev = [ops.as_device(L.weights, torch.float32, self.device) for L in self.layers]
        self.memory_budget = memory_budget
        self.memory_in_use = 0
        # batches in flight at once (one per stream): with two, a batch is assembled and
        # launched while the previous one computes
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(max(1, int(n_streams)))]
        self.graphs: Dict[Tuple[int, int, int], _GraphBatch] = {}
        self.queue: collections.deque = collections.deque()
        self.inflight: List[Tuple[int, _GraphBatch]] = []
        self.completed: List[Request] = []

    # -------------------------------------------------------------- accounting
    def request_bytes(self, r: Request) -> int:
        """Device bytes a request holds while admitted: its rows of the event buffers, the
        per-neuron input buffer and every layer's outputs and workspace rows."""
        P = _bucket(len(r.indices), self.min_events_bucket)
        K_in = _bucket(max(r.multiplicity, 1), self.min_spikes_bucket)
        b = 8 * P + 4 * self.n_inputs * K_in * 3 + 64 * self.n_inputs * K_in
        for L in self.layers:
            b += 4 * L.weights.shape[0] * (1 + 5 * L.max_n_spike)
        return b

    def submit(self, r: Request) -> None:
        r.multiplicity = int(np.bincount(r.indices, minlength=self.n_inputs).max()) if len(r.indices) else 0
        self.queue.append(r)

    # -------------------------------------------------------------- batching
    def _admit(self) -> List[Request]:
        batch: List[Request] = []
        while self.queue and len(batch) < self.max_batch:
            r = self.queue[0]
            need = self.request_bytes(r)
            if self.memory_budget is not None and self.memory_in_use + need > self.memory_budget:
                break
            self.queue.popleft()
            self.memory_in_use += need
            batch.append(r)
        return batch

    def _graph_for(self, B: int, P: int, K_in: int, stream) -> _GraphBatch:
        key = (B, P, K_in, id(stream))
        g = self.graphs.get(key)
        if g is None:
            g = _GraphBatch(self, B, P, K_in, stream)
            self.graphs[key] = g
        return g

    def _launch(self, batch: List[Request], stream) -> _GraphBatch:
        Bb = min(_bucket(len(batch), self.min_batch_bucket), _bucket(self.max_batch, 1))
        Pb = _bucket(max(len(r.indices) for r in batch), self.min_events_bucket)
        Kb = _bucket(max(max(r.multiplicity for r in batch), 1), self.min_spikes_bucket)
        g = self._graph_for(Bb, Pb, Kb, stream)
        hi, ht = g.h_idx.numpy(), g.h_t.numpy()
        # only the slots written by an earlier batch are reset (padding is index -1, time +inf)
        for s in range(Bb):
            L = len(batch[s].indices) if s < len(batch) else 0
            prev = int(g.row_len[s])
            if s < len(batch):
                hi[s, :L] = batch[s].indices
                ht[s, :L] = batch[s].times
            if prev > L:
                hi[s, L:prev] = -1
                ht[s, L:prev] = np.inf
            g.row_len[s] = L
        with torch.cuda.stream(stream):
            g.idx.copy_(g.h_idx, non_blocking=True)
            g.t.copy_(g.h_t, non_blocking=True)
            g.graph.replay()
            g.h_n.copy_(g.outputs[0], non_blocking=True)
            g.h_st.copy_(g.outputs[1], non_blocking=True)
            g.h_pred.copy_(g.outputs[2], non_blocking=True)
            g.done.record(stream)
        g.requests = batch
        return g

    def _retire(self, g: _GraphBatch) -> None:
        now = time.perf_counter()
        n, st, pred = g.h_n.numpy(), g.h_st.numpy(), g.h_pred.numpy()
        for s, r in enumerate(g.requests):
            r.done = now
            r.pred = int(pred[s])
            r.counts = n[s].copy()
            r.spike_times = st[s].copy()
            self.memory_in_use -= self.request_bytes(r)
            self.completed.append(r)
        g.requests = []

    def poll(self) -> None:
        """One scheduling round: retire finished batches, then launch new ones on free streams."""
        still = []
        for si, g in self.inflight:
            if g.done.query():
                self._retire(g)
            else:
                still.append((si, g))
        self.inflight = still
        busy = {si for si, _ in self.inflight}
        for si, stream in enumerate(self.streams):
            if si in busy or not self.queue:
                continue
            batch = self._admit()
            if not batch:
                break
            self.inflight.append((si, self._launch(batch, stream)))

    def drain(self) -> None:
        while self.queue or self.inflight:
            self.poll()

    def warmup(self, shapes: Sequence[Tuple[int, int, int]]) -> None:
        """Capture the graphs of the given (batch, events, spikes-per-neuron) buckets."""
        for B, P, K in shapes:
            for s in self.streams:
                self._graph_for(B, P, K, s)
        torch.cuda.synchronize()


def run_poisson(server: InferenceServer, requests: Sequence[Request], rate: float,
                seed: int = 0) -> Dict[str, float]:
    """Feeds ``requests`` with Poisson arrivals at ``rate`` requests/s (wall clock), serves
    them to completion and returns throughput and latency percentiles (arrival -> result
    copied to the host)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(requests))
    t0 = time.perf_counter()
    arrivals = t0 + np.cumsum(gaps)
    server.completed = []
    k = 0
    while k < len(requests) or server.queue or server.inflight:
        now = time.perf_counter()
        while k < len(requests) and arrivals[k] <= now:
            requests[k].arrival = arrivals[k]
            server.submit(requests[k])
            k += 1
        server.poll()
    lat = np.array([r.done - r.arrival for r in server.completed])
    span = max(r.done for r in server.completed) - t0
    return {"rate": rate, "samples_per_s": len(server.completed) / span,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "n": len(server.completed)}


def make_requests(n: int, n_inputs: int = 784, p_min: int = 50, p_max: int = 2000,
                  seed: int = 0) -> List[Request]:
    """Synthetic requests: P ~ U{p_min..p_max} events, input index ~ U{0..n_inputs-1} with
    replacement, times ~ U[0, 1], sorted by time."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        P = int(rng.integers(p_min, p_max + 1))
        idx = rng.integers(0, n_inputs, size=P).astype(np.int32)
        tm = rng.uniform(0.0, 1.0, size=P).astype(np.float32)
        order = np.argsort(tm, kind="stable")
        out.append(Request(rid=i, indices=idx[order], times=tm[order], arrival=0.0))
    return out
