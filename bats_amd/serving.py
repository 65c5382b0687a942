"""Continuous-batching inference server for a dense LIF network on one GPU.

Requests are single samples given as event lists (input neuron index, spike time), of any
length; several spikes per input neuron are allowed. The server keeps the network's weights
and all batch buffers resident on the device and runs batches back to back on two streams:
while one batch computes, the next is assembled from whatever requests have arrived and
launched as soon as a stream is free. Each batch replays a HIP graph captured for its padded
shape (batch bucket, events-per-request bucket), so the host cost per batch is one packed copy
of the batch's event lists and one graph launch.

Batching is length-aware. ``length_classes`` cut the event counts into classes, each with its
own queue. A batch takes the class of the oldest waiting request and fills up with requests of
that class and shorter ones, oldest first: a batch's duration is set by its longest sample (one
workgroup walks a sample's events), so short requests that are the oldest do not wait behind
long ones. Inside a batch the requests are ordered longest first, so that the workgroups of the
longest samples start first. A graph's shape does not depend on the classes (the event
conversion and the solver work on the real events only), and a graph is replayed on whichever
stream is free: a second copy of a shape is made only when two batches of it are in flight.

Device memory: ``memory_budget`` bounds the bytes of the captured graphs the server holds
(their static buffers and private memory pools, measured when a graph is made). A batch shape
whose graph is not resident evicts idle graphs, least recently used first, to make room; when
it still does not fit, the batch is made smaller.

The host does no per-event work in Python: a request's events live in the server's pinned
host arena from submission (``new_request`` hands out arena space for a producer to fill;
``submit`` copies other arrays in), and a batch's event lists are gathered back to back into
the batch's pinned buffer by the library's host thread pool (``bats_pack_event_lists``). The
first layer reads the packed lists directly (``bats_lif_forward_packed``: a request's events
are already in time order, so no per-neuron buffers and no sort), so any number of events per
input neuron is served in one pass.

The inference of a request is independent of the other requests of its batch (each sample is
sorted and solved on its own), so a request's spike trains are the same alone and in any mix.

Reference counterparts: bats/Network.py:32-36 (forward), bats/Losses/SpikeCountLoss.py:17-18
(predict), bats/CudaKernels/Wrappers/Inference/convert_to_spike_per_neuron.py:20-33 (the event
lists the first layer reads). The reference has no server; this is new.
"""
from __future__ import annotations

import bisect
import collections
import dataclasses
import gc
import heapq
import operator
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib, ops
from ._constants import lif_constants
from .ops import _ptr


class Request:
    """One inference request: ``indices`` (int32 [P]) and ``times`` (float32 [P], ascending)
    of its input events, its ``arrival`` on the host clock (seconds). Once served: ``counts``
    (int32 [N_out]), ``spike_times`` (float32 [N_out, K], +inf past the count), ``pred`` and
    ``done`` (host clock), read from its batch's result arrays (a request keeps them alive);
    ``indices`` and ``times`` are then None (their arena space is reused). While the events
    are in a server's arena, ``indices`` and ``times`` are views of it, made on each access."""

    __slots__ = ("rid", "arrival", "n_events", "block", "cls", "_data", "_out", "_slot")

    def __init__(self, rid: int, indices: np.ndarray, times: np.ndarray, arrival: float = 0.0):
        self.rid = rid
        self.arrival = arrival
        self.n_events = len(indices)
        self.block: Optional[list] = None  # [start, end, free] of its events in the arena
        self.cls = -1
        # (indices, times) of the caller, or the arena that holds them (at ``block``)
        self._data = (indices, times)
        self._out: Optional[tuple] = None  # (counts, spike_times, preds, done) of its batch
        self._slot = 0

    @property
    def indices(self) -> Optional[np.ndarray]:
        d = self._data
        if d is None or d.__class__ is tuple:
            return None if d is None else d[0]
        return d.idx[self.block[0]:self.block[1]]

    @property
    def times(self) -> Optional[np.ndarray]:
        d = self._data
        if d is None or d.__class__ is tuple:
            return None if d is None else d[1]
        return d.t[self.block[0]:self.block[1]]

    @property
    def done(self) -> float:
        return 0.0 if self._out is None else self._out[3]

    @property
    def pred(self) -> int:
        return -1 if self._out is None else self._out[2][self._slot]

    @property
    def counts(self) -> Optional[np.ndarray]:
        return None if self._out is None else self._out[0][self._slot]

    @property
    def spike_times(self) -> Optional[np.ndarray]:
        return None if self._out is None else self._out[1][self._slot]


@dataclasses.dataclass
class LayerSpec:
    weights: np.ndarray   # [J, I] float32
    tau_s: float
    theta: float
    delta_theta: float
    max_n_spike: int


_n_events = operator.attrgetter("n_events")
_NO_EVENTS = (np.zeros(0, np.int32), np.zeros(0, np.float32))


def _bucket(v: int, minimum: int) -> int:
    b = minimum
    while b < v:
        b *= 2
    return b


class _Arena:
    """Pinned host memory holding the event lists of the requests a server has accepted, used
    as a ring: a request's events take one contiguous block, handed out in order; a block is
    freed when its request completes, and its space is reused once every older block is free
    (requests complete roughly in order)."""

    def __init__(self, capacity: int):
        self.capacity = int(capacity)
        self._idx = torch.empty((self.capacity,), dtype=torch.int32, pin_memory=True)
        self._t = torch.empty((self.capacity,), dtype=torch.float32, pin_memory=True)
        self.idx, self.t = self._idx.numpy(), self._t.numpy()
        self.idx_ptr, self.t_ptr = self._idx.data_ptr(), self._t.data_ptr()
        self.blocks: collections.deque = collections.deque()  # [start, end, free], oldest first

    def alloc(self, n: int) -> Optional[list]:
        """A block of n events, or None when the ring has no contiguous room for it."""
        if n > self.capacity:
            return None
        if not self.blocks:
            start = 0
        else:
            old, new = self.blocks[0], self.blocks[-1]
            if new[0] >= old[0]:  # live region not wrapped: room after it, then before it
                if new[1] + n <= self.capacity:
                    start = new[1]
                elif n <= old[0]:
                    start = 0
                else:
                    return None
            elif new[1] + n <= old[0]:
                start = new[1]
            else:
                return None
        blk = [start, start + n, False]
        self.blocks.append(blk)
        return blk

    def free(self, blk: list) -> None:
        blk[2] = True
        self.release()

    def release(self) -> None:
        """Returns the space of the oldest blocks that are marked free."""
        blocks = self.blocks
        while blocks and blocks[0][2]:
            blocks.popleft()

    @property
    def in_use(self) -> int:
        """Events held by live blocks."""
        return sum(b[1] - b[0] for b in self.blocks if not b[2])


class _GraphBatch:
    """Static buffers + captured graph for one padded batch shape: room for B requests of up
    to P events each (packed back to back)."""

    def __init__(self, server: "InferenceServer", B: int, P: int, stream):
        self.B, self.P = B, P
        dev = server.device
        self.stream = stream
        self.h_off = torch.zeros((B + 1,), dtype=torch.int32, pin_memory=True)
        self.h_idx = torch.full((B * P,), -1, dtype=torch.int32, pin_memory=True)
        self.h_t = torch.full((B * P,), float("inf"), dtype=torch.float32, pin_memory=True)
        self.off = torch.zeros((B + 1,), dtype=torch.int32, device=dev)
        self.idx = torch.full((B * P,), -1, dtype=torch.int32, device=dev)
        self.t = torch.full((B * P,), float("inf"), dtype=torch.float32, device=dev)
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
        # the first use of a graph and of its host buffers in copies carries one-time set-up
        # costs (milliseconds); one full round trip on empty lists takes them here
        with torch.cuda.stream(stream):
            self._round_trip()
        stream.synchronize()

    def release(self) -> None:
        """Drops the graph and its device buffers (the caller returns the memory)."""
        self.graph.reset()
        self.graph = self.outputs = None
        self.off = self.idx = self.t = self.zero_targets = None

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
        self.done.record()

    def _run(self, server):
        layers = list(zip(server.layers, server.w_dev))
        tmax = server.max_simulation

        def chain_to(k):
            # the next layer's constants when layer k's solver builds its event lists
            if k + 1 >= len(layers) or not ops.chain_fuses(self.B, layers[k][1].shape[0]):
                return None
            L = layers[k + 1][0]
            return (*lif_constants(L.tau_s, L.theta, L.delta_theta)[:3], tmax)

        L0, w0 = layers[0]
        res = ops.lif_forward_packed(self.off, self.idx, self.t, self.P, w0,
                                     *lif_constants(L0.tau_s, L0.theta, L0.delta_theta),
                                     tmax, L0.max_n_spike, chain=chain_to(0))
        lists = None
        if isinstance(res[-1], ops.EventLists):
            res, lists = res
        n, x = res[0], res[1]
        for k in range(1, len(layers)):
            L, w = layers[k]
            consts = lif_constants(L.tau_s, L.theta, L.delta_theta)
            ch = chain_to(k)
            if lists is not None:
                res = ops.lif_forward_events(lists, w, *consts, L.max_n_spike, chain=ch)
            elif ch is not None:
                res = ops.lif_forward_chain(x, w, *consts, tmax, L.max_n_spike, ch)
            else:
                res = ops.lif_forward(x, w, *consts, tmax, L.max_n_spike)
            lists = None
            if ch is not None:
                res, lists = res
            n, x = res[0], res[1]
        _, _, pred = ops.count_loss(n, x.shape[2], targets=self.zero_targets, want_loss=False,
                                    want_errors=False, want_pred=True)
        return n, x, pred


class InferenceServer:
    """Admits requests, batches them continuously by length class and runs them on the GPU.

    ``memory_budget`` (bytes, optional) bounds the device memory of the captured batch graphs
    the server holds; the weights are outside it. ``batch_growth`` is the ratio of successive
    batch buckets (fewer distinct graphs with a larger ratio, more padding).
    """

    def __init__(self, layers: Sequence[LayerSpec], n_inputs: int, max_batch: int = 1024,
                 max_simulation: float = 1.0, memory_budget: Optional[int] = None,
                 device=None, min_batch_bucket: int = 8, min_events_bucket: int = 64,
                 n_streams: int = 2, batch_growth: int = 2,
                 length_classes: Sequence[int] = (), arena_events: int = 1 << 24):
        self.device = torch.device("cuda", 0) if device is None else torch.device(device)
        self.layers = list(layers)
        self.n_inputs = int(n_inputs)
        if not 1 <= self.n_inputs <= ops.PACKED_EVENTS_MAX_NEURONS:
            raise ValueError(f"n_inputs must be in [1, {ops.PACKED_EVENTS_MAX_NEURONS}] (the "
                             f"device conversion of packed event lists), got {self.n_inputs}")
        self.max_batch = int(max_batch)
        self.max_simulation = float(max_simulation)
        self.min_batch_bucket = int(min_batch_bucket)
        self.batch_growth = max(2, int(batch_growth))
        # larger minimum buckets mean fewer distinct padded shapes (graphs) at the cost of
        # larger packed buffers (the kernels work on the real events only)
        self.min_events_bucket = int(min_events_bucket)
        self.length_classes = sorted(int(v) for v in length_classes)
        # accepted requests' event lists (pinned host memory, 8 bytes per event)
        self.arena = _Arena(arena_events)
        self.w_dev = [ops.as_device(L.weights, torch.float32, self.device) for L in self.layers]
        self.memory_budget = memory_budget
        self.memory_in_use = 0  # device bytes of the resident graphs
        self.graphs_made = 0
        self.graphs_evicted = 0
        # batches in flight at once (one per stream): with two, a batch is assembled and
        # launched while the previous one computes
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(max(1, int(n_streams)))]
        # (B, P, copy) -> graph, least recently used first
        self.graphs: "collections.OrderedDict[Tuple[int, int, int, int], _GraphBatch]" = \
            collections.OrderedDict()
        self.queues: Dict[int, collections.deque] = {}
        self.inflight: List[Tuple[int, _GraphBatch]] = []
        self.completed: List[Request] = []

    # -------------------------------------------------------------- shapes and memory
    def events_class(self, r: Request) -> int:
        """The request's length class: the number of class bounds below its event count."""
        return bisect.bisect_left(self.length_classes, r.n_events)

    def batch_bucket(self, n: int) -> int:
        b = self.min_batch_bucket
        while b < n:
            b *= self.batch_growth
        return min(b, self.max_batch)

    def graph_bytes_estimate(self, B: int, P: int) -> int:
        """Device bytes of a batch graph of this shape: packed lists, every layer's outputs and
        workspace (and the later layers' input exps), the prediction (with a margin for the
        allocator's rounding)."""
        lib = _lib.load()
        N = self.n_inputs
        b = 8 * B * P + 4 * (B + 1)
        I, Kp = N, 0
        for k, L in enumerate(self.layers):
            J, K = L.weights.shape[0], L.max_n_spike
            b += 4 * B * J + 16 * B * J * K
            if k == 0:
                b += int(lib.bats_lif_forward_packed_workspace_size(B, I, P, J))
            else:
                b += 8 * B * I * Kp + int(lib.bats_lif_forward_workspace_size(B, I, Kp, J))
            I, Kp = J, K
        b += 4 * B * I + 8 * B
        return int(b * 1.25) + (8 << 20)

    def _resident_bytes(self) -> int:
        return sum(g.nbytes for g in self.graphs.values())

    def _evict_for(self, need: int, keep=None) -> bool:
        """Evicts idle graphs, least recently used first, until ``need`` more bytes fit the
        budget. Returns whether they fit."""
        if self.memory_budget is None:
            return True
        busy = {id(g) for _, g in self.inflight}
        freed = False
        for key in list(self.graphs.keys()):
            if self.memory_in_use + need <= self.memory_budget:
                break
            g = self.graphs[key]
            if id(g) in busy or g is keep:
                continue
            del self.graphs[key]
            self.memory_in_use -= g.nbytes
            g.release()
            self.graphs_evicted += 1
            freed = True
        if freed:
            gc.collect()
            torch.cuda.empty_cache()
        return self.memory_in_use + need <= self.memory_budget

    def _graph_for(self, B: int, P: int, si: int, copy: Optional[int] = None):
        """An idle graph of the shape (the first idle copy, or copy ``copy``), made if there
        is none and it fits the budget; None otherwise."""
        busy = {id(g) for _, g in self.inflight}
        copies = [k for k in self.graphs if k[:2] == (B, P)]
        for key in copies:
            if (copy is None or key[2] == copy) and id(self.graphs[key]) not in busy:
                self.graphs.move_to_end(key)
                return self.graphs[key]
        if copy is None:
            used = {k[2] for k in copies}
            copy = next(c for c in range(len(copies) + 1) if c not in used)
        elif (B, P, copy) in self.graphs:
            return None
        key = (B, P, copy)
        need = self.graph_bytes_estimate(B, P)
        if not self._evict_for(need):
            return None
        # the graph's footprint: device memory reserved by making it, with the allocator's
        # free cache emptied before and after (the warm-up run's temporaries are released)
        torch.cuda.synchronize(self.device)
        gc.collect()
        torch.cuda.empty_cache()
        before = torch.cuda.memory_reserved(self.device)
        g = _GraphBatch(self, B, P, self.streams[si])
        torch.cuda.synchronize(self.device)
        torch.cuda.empty_cache()
        g.nbytes = max(torch.cuda.memory_reserved(self.device) - before, 1)
        self.graphs[key] = g
        self.memory_in_use += g.nbytes
        self.graphs_made += 1
        # the measured size may exceed the estimate: make room after the fact
        self._evict_for(0, keep=g)
        return g

    # -------------------------------------------------------------- requests
    def new_request(self, n_events: int, rid: int, arrival: float = 0.0) -> Request:
        """A request whose ``indices`` / ``times`` are views of ``n_events`` slots of the
        server's arena, for the caller to fill (time-ordered) and then ``submit``: its events
        are not copied again before the batch gather. The arena space is the server's again
        once the request completes (its ``indices`` and ``times`` are then set to None)."""
        blk = self.arena.alloc(int(n_events))
        if blk is None:
            raise MemoryError(f"the request arena ({self.arena.capacity} events) has no room "
                              f"for {n_events} events")
        r = Request(rid, _NO_EVENTS[0], _NO_EVENTS[1], arrival)
        r.n_events, r.block, r._data = int(n_events), blk, self.arena
        return r

    def submit(self, r: Request) -> None:
        if r.block is None:  # events outside the arena: copy them in
            idx, tm = r._data
            n = r.n_events = len(idx)
            blk = self.arena.alloc(n)
            if blk is None:
                raise MemoryError(f"the request arena ({self.arena.capacity} events) has no "
                                  f"room for {n} events")
            self.arena.idx[blk[0]:blk[1]] = idx
            self.arena.t[blk[0]:blk[1]] = tm
            r.block, r._data = blk, self.arena
        c = r.cls = self.events_class(r)
        q = self.queues.get(c)
        if q is None:
            q = self.queues[c] = collections.deque()
        q.append(r)

    @property
    def queue(self) -> List[Request]:
        """The waiting requests (all classes)."""
        return [r for q in self.queues.values() for r in q]

    def _waiting(self) -> bool:
        return any(self.queues.values())

    def _take(self) -> List[Request]:
        """Up to max_batch requests: the class of the oldest waiting request and the shorter
        classes, oldest first (all of them, in class order, when they fit one batch: the
        batch is ordered by length at launch anyway)."""
        heads = [(q[0].arrival, q[0].rid, c) for c, q in self.queues.items() if q]
        if not heads:
            return []
        c_top = min(heads)[2]
        eligible = [q for c, q in self.queues.items() if q and c <= c_top]
        batch: List[Request] = []
        if sum(len(q) for q in eligible) <= self.max_batch:
            for q in eligible:
                batch.extend(q)
                q.clear()
            return batch
        heap = [(q[0].arrival, q[0].rid, c) for c, q in self.queues.items() if q and c <= c_top]
        heapq.heapify(heap)
        while heap and len(batch) < self.max_batch:
            _, _, c = heapq.heappop(heap)
            q = self.queues[c]
            batch.append(q.popleft())
            if q:
                heapq.heappush(heap, (q[0].arrival, q[0].rid, c))
        return batch

    def _give_back(self, reqs: Sequence[Request]) -> None:
        """Returns requests to the front of their queues (in their order)."""
        for r in reversed(list(reqs)):
            self.queues.setdefault(r.cls, collections.deque()).appendleft(r)

    # -------------------------------------------------------------- batching
    def _launch(self, batch: List[Request], si: int) -> Optional[_GraphBatch]:
        batch.sort(key=_n_events, reverse=True)  # longest first (stable)
        Pb = _bucket(batch[0].n_events, self.min_events_bucket)
        Bb = self.batch_bucket(len(batch))
        g = self._graph_for(Bb, Pb, si)
        while g is None and Bb > self.min_batch_bucket:
            # the shape does not fit the memory budget: a smaller batch
            Bb = max(self.min_batch_bucket, Bb // self.batch_growth)
            g = self._graph_for(Bb, Pb, si)
        if g is None:
            if not self.inflight and not self.graphs:
                raise ValueError(f"the memory budget ({self.memory_budget} bytes) cannot hold "
                                 f"the smallest batch graph ({self.graph_bytes_estimate(Bb, Pb)} "
                                 f"bytes estimated)")
            self._give_back(batch)
            return None
        if len(batch) > Bb:
            self._give_back(batch[Bb:])
            batch = batch[:Bb]
        nb = len(batch)
        starts = np.fromiter((r.block[0] for r in batch), dtype=np.int64, count=nb)
        lens = np.fromiter(map(_n_events, batch), dtype=np.int32, count=nb)
        off = g.h_off.numpy()
        _lib.check(_lib.load().bats_pack_event_lists(
            nb, starts.ctypes.data, lens.ctypes.data, self.arena.idx_ptr, self.arena.t_ptr,
            g.h_off.data_ptr(), g.h_idx.data_ptr(), g.h_t.data_ptr(), g.B * g.P))
        total = int(off[nb])
        off[nb + 1:] = total  # unused slots: empty lists
        with torch.cuda.stream(self.streams[si]):
            g._round_trip(total)
        g.requests = batch
        return g

    def _retire(self, g: _GraphBatch) -> None:
        nb = len(g.requests)
        out = (g.h_n.numpy()[:nb].copy(), g.h_st.numpy()[:nb].copy(),
               g.h_pred.numpy()[:nb].tolist(), time.perf_counter())
        for s, r in enumerate(g.requests):
            r._out = out
            r._slot = s
            r.block[2] = True  # its arena space is free (returned in order below)
            r.block = r._data = None
        self.arena.release()
        self.completed.extend(g.requests)
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
        for si in range(len(self.streams)):
            if si in busy or not self._waiting():
                continue
            batch = self._take()
            if not batch:
                break
            g = self._launch(batch, si)
            if g is None:
                break  # waits for memory held by batches in flight
            self.inflight.append((si, g))

    def drain(self) -> None:
        while self._waiting() or self.inflight:
            self.poll()

    def warmup(self, shapes: Sequence[Tuple[int, int]], copies: int = 1) -> None:
        """Capture the graphs of the given (batch, events) buckets, each in ``copies``
        copies."""
        for B, P in shapes:
            for c in range(copies):
                self._graph_for(B, P, c % len(self.streams), copy=c)
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
    while k < len(requests) or server._waiting() or server.inflight:
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
                  seed: int = 0, server: Optional[InferenceServer] = None) -> List[Request]:
    """Synthetic requests: P ~ U{p_min..p_max} events, input index ~ U{0..n_inputs-1} with
    replacement, times ~ U[0, 1], sorted by time. With ``server`` they are written straight
    into its arena (``new_request``); the values are the same either way."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        P = int(rng.integers(p_min, p_max + 1))
        idx = rng.integers(0, n_inputs, size=P).astype(np.int32)
        tm = rng.uniform(0.0, 1.0, size=P).astype(np.float32)
        order = np.argsort(tm, kind="stable")
        if server is None:
            out.append(Request(rid=i, indices=idx[order], times=tm[order], arrival=0.0))
        else:
            r = server.new_request(P, rid=i)
            r.indices[:] = idx[order]
            r.times[:] = tm[order]
            out.append(r)
    return out
