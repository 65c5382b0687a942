"""Continuous-batching inference server for a dense LIF network on one GPU.

Requests are single samples given as event lists (input neuron index, spike time), of any
length; several spikes per input neuron are allowed. The server keeps the network's weights
and all batch buffers resident on the device and runs batches back to back on two streams:
while one batch computes, the next is assembled from whatever requests have arrived and
admitted, and launched as soon as a stream is free. A request is admitted when a batch slot is
free and its device footprint fits the memory budget; it is released when its result has been
copied out. Each batch replays a HIP graph captured for its padded shape (batch bucket,
events-per-request bucket, spikes-per-neuron bucket), so the host cost per batch is one packed
copy of the batch's event lists and one graph launch.

The host does no per-event work: a batch's event lists are concatenated into one packed list
with row offsets, and the device converts it to per-neuron buffers. The spikes-per-input-neuron
bucket is the server's minimum bucket unless a request is known to need more; the device
reports every sample's largest per-neuron count, and a request whose events did not all fit
is served again in a batch with a bucket that holds them (its first result is discarded).

The inference of a request is independent of the other requests of its batch (each sample is
sorted and solved on its own), so a request's spike trains are the same alone and in any mix.

Reference counterparts: bats/Network.py:32-36 (forward), bats/Losses/SpikeCountLoss.py:17-18
(predict), bats/CudaKernels/Wrappers/Inference/convert_to_spike_per_neuron.py:20-33 (event
list to per-neuron buffers). The reference has no server; this is new.
"""
from __future__ import annotations

import collections
import dataclasses
import gc
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
    multiplicity: int = 0  # largest per-neuron count, when known (0: not known)
    held_bytes: int = 0
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
    """Static buffers + captured graph for one padded batch shape: room for B requests of up
    to P events each (packed back to back), K spikes per input neuron."""

    def __init__(self, server: "InferenceServer", B: int, P: int, K_in: int, stream):
        self.B, self.P, self.K_in = B, P, K_in
        dev = server.device
        self.stream = stream
        self.h_off = torch.zeros((B + 1,), dtype=torch.int32, pin_memory=True)
        self.h_idx = torch.full((B * P,), -1, dtype=torch.int32, pin_memory=True)
        self.h_t = torch.full((B * P,), float("inf"), dtype=torch.float32, pin_memory=True)
        self.off = torch.zeros((B + 1,), dtype=torch.int32, device=dev)
        self.idx = torch.full((B * P,), -1, dtype=torch.int32, device=dev)
        self.t = torch.full((B * P,), float("inf"), dtype=torch.float32, device=dev)
        N = server.n_inputs
        self.per = (torch.empty((B, N, K_in), dtype=torch.float32, device=dev),
                    torch.empty((B, N), dtype=torch.int32, device=dev),
                    torch.empty((B,), dtype=torch.int32, device=dev))
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
        self.h_rowmax = torch.empty((B,), dtype=torch.int32, pin_memory=True)
        self.done = torch.cuda.Event()
        self.requests: List[Request] = []
        # the first use of a graph and of its host buffers in copies carries one-time set-up
        # costs (milliseconds); one full round trip on empty lists takes them here
        with torch.cuda.stream(stream):
            self._round_trip()
        stream.synchronize()

    def _round_trip(self, total: Optional[int] = None) -> None:
        """Copies the packed lists in, replays the graph, copies the results out and records
        ``done`` (on the current stream)."""
        total = self.B * self.P if total is None else total
        self.off.copy_(self.h_off, non_blocking=True)
        if total:
            self.idx[:total].copy_(self.h_idx[:total], non_blocking=True)
            self.t[:total].copy_(self.h_t[:total], non_blocking=True)
        self.graph.replay()
        self.h_n.copy_(self.outputs[0], non_blocking=True)
        self.h_st.copy_(self.outputs[1], non_blocking=True)
        self.h_pred.copy_(self.outputs[2], non_blocking=True)
        self.h_rowmax.copy_(self.outputs[3], non_blocking=True)
        self.done.record()

    def _run(self, server):
        per_t, _, row_max = ops.events_csr_to_per_neuron(self.off, self.idx, self.t,
                                                         server.n_inputs, self.K_in,
                                                         out=self.per)
        x = per_t
        n = None
        for L, w in zip(server.layers, server.w_dev):
            n, x, *_ = ops.lif_forward(x, w, *lif_constants(L.tau_s, L.theta, L.delta_theta),
                                       server.max_simulation, L.max_n_spike)
        _, _, pred = ops.count_loss(n, x.shape[2], targets=self.zero_targets, want_loss=False,
                                    want_errors=False, want_pred=True)
        return n, x, pred, row_max


class InferenceServer:
    """Admits requests, batches them continuously and runs them on the GPU.

    ``memory_budget`` bounds the device bytes held by admitted requests (their padded share of
    the batch buffers); a request that does not fit waits in the queue.
    """

    def __init__(self, layers: Sequence[LayerSpec], n_inputs: int, max_batch: int = 1024,
                 max_simulation: float = 1.0, memory_budget: Optional[int] = None,
                 device=None, min_batch_bucket: int = 8, min_events_bucket: int = 64,
                 min_spikes_bucket: int = 8, n_streams: int = 2):
        self.device = torch.device("cuda", 0) if device is None else torch.device(device)
        self.layers = list(layers)
        self.n_inputs = int(n_inputs)
        if not 1 <= self.n_inputs <= ops.PACKED_EVENTS_MAX_NEURONS:
            raise ValueError(f"n_inputs must be in [1, {ops.PACKED_EVENTS_MAX_NEURONS}] (the "
                             f"device conversion of packed event lists), got {self.n_inputs}")
        self.max_batch = int(max_batch)
        self.max_simulation = float(max_simulation)
        self.min_batch_bucket = int(min_batch_bucket)
        # larger minimum buckets mean fewer distinct padded shapes (graphs) at the cost of
        # padding work; a server with a known request mix sets them to its upper ends. A
        # request with more spikes on one input neuron than the spikes bucket is served twice
        # (the second time in a bucket that holds them).
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
        """Device bytes a request holds while admitted: its share of the packed event
        buffers, the per-neuron input buffer and every layer's outputs and workspace rows."""
        P = _bucket(len(r.indices), self.min_events_bucket)
        K_in = _bucket(max(r.multiplicity, 1), self.min_spikes_bucket)
        b = 8 * P + 4 * self.n_inputs * K_in * 3 + 64 * self.n_inputs * K_in
        for L in self.layers:
            b += 4 * L.weights.shape[0] * (1 + 5 * L.max_n_spike)
        return b

    def submit(self, r: Request) -> None:
        self.queue.append(r)

    # -------------------------------------------------------------- batching
    def _admit(self) -> List[Request]:
        batch: List[Request] = []
        while self.queue and len(batch) < self.max_batch:
            r = self.queue[0]
            need = self.request_bytes(r)
            if self.memory_budget is not None and need > self.memory_budget:
                raise ValueError(f"request {r.rid} needs {need} bytes, more than the whole "
                                 f"memory budget ({self.memory_budget})")
            if self.memory_budget is not None and self.memory_in_use + need > self.memory_budget:
                break
            self.queue.popleft()
            self.memory_in_use += need
            r.held_bytes = need
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
        lens = [len(r.indices) for r in batch]
        Pb = _bucket(max(lens), self.min_events_bucket)
        Kb = _bucket(max(max(r.multiplicity for r in batch), 1), self.min_spikes_bucket)
        g = self._graph_for(Bb, Pb, Kb, stream)
        off = g.h_off.numpy()
        off[0] = 0
        np.cumsum(lens, out=off[1:len(batch) + 1])
        total = int(off[len(batch)])
        off[len(batch) + 1:] = total  # unused slots: empty lists
        if total:
            np.concatenate([r.indices for r in batch], out=g.h_idx.numpy()[:total])
            np.concatenate([r.times for r in batch], out=g.h_t.numpy()[:total])
        with torch.cuda.stream(stream):
            g._round_trip(total)
        g.requests = batch
        return g

    def _retire(self, g: _GraphBatch) -> None:
        now = time.perf_counter()
        nb = len(g.requests)
        n = g.h_n.numpy()[:nb].copy()
        st = g.h_st.numpy()[:nb].copy()
        pred = g.h_pred.numpy()[:nb].tolist()
        row_max = g.h_rowmax.numpy()[:nb]
        over = np.nonzero(row_max > g.K_in)[0]
        redo = []
        for s in over.tolist():
            # the request's events did not all fit its bucket: serve it again with one that
            # holds them (its slot's results above are discarded)
            r = g.requests[s]
            r.multiplicity = int(row_max[s])
            redo.append(r)
        skip = set(over.tolist())
        for s, r in enumerate(g.requests):
            self.memory_in_use -= r.held_bytes
            r.held_bytes = 0
            if s in skip:
                continue
            r.done = now
            r.pred = pred[s]
            r.counts = n[s]
            r.spike_times = st[s]
            self.completed.append(r)
        self.queue.extendleft(reversed(redo))
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
    # earlier results are released (and their memory collected) before the clock starts
    server.completed = []
    gc.collect()
    t0 = time.perf_counter()
    arrivals = t0 + np.cumsum(gaps)
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
