#!/usr/bin/env python3
"""Training-step throughput of the 784-800-10 multi-spike LIF network on one MI355X.

Workload (one step = one batch through the hot path): latency-coded inputs (x ~ U[0,1],
input active iff x > 0.1, t = 1 - x, one spike per input), 784 -> 800 LIF (theta = delta_theta
= 1.0, tau_s = 0.130, 30-spike buffers) -> 10 LIF (theta = delta_theta = 1.3, 30-spike
buffers), batch 256, spike-count class loss (targets 15 / 3), BATS backward, Adam, weight
update; max_simulation = 1.0. Weights are drawn from N(0, 1) ("active" variant: the literal
N(0, 1/sqrt(fan_in)) initialisation leaves the hidden layer silent, see --variant literal).

The timed step is the whole step (forward of both layers, loss, backward of both layers,
Adam with an on-device step counter, weight update); inputs, labels and weights are resident
in HBM. It is timed both as stream launches and as replays of one captured HIP graph, K steps
each after W warm-up steps from the initial weights; `value` is the faster of the two (named in
config.launch), from the wall clock around exactly K steps. Every step is also bracketed by HIP
events on its stream (`step_ms_hip_events`: median, p10, p90, first and last step): training
raises the hidden layer's activity, so later steps cost more (`spikes_per_sample` gives the
activity of the first and of the last timed step). Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

METRIC = "train samples/s (784-800-10, 30 spk/neuron) vs plain form; p50/p99 serve latency"
KERNELS = ("lif_solve_seq_kernel", "lif_solve_wave_kernel", "sort_events_kernel",
           "wgrad_tile_kernel", "row_prep_kernel", "small_bwd_kernel")
HBM_PEAK_GBS = 8000.0

# network of the workload
N_IN, N_HID, N_OUT = 784, 800, 10
TAU_S = 0.130
THETA_H, THETA_O = 1.0, 1.3
K_HID, K_OUT = 30, 30
T_MAX = 1.0
TARGET_TRUE, TARGET_FALSE = 15, 3
LR = 0.003


def make_inputs(seed, batch, variant):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, size=(batch, N_IN))
    t = np.where(x > 0.1, 1.0 - x, np.inf).astype(np.float32)[:, :, None]
    n = np.isfinite(t[:, :, 0]).astype(np.int32)
    labels = rng.integers(0, N_OUT, size=batch).astype(np.int64)
    if variant == "active":
        w1 = rng.normal(0.0, 1.0, size=(N_HID, N_IN))
        w2 = rng.normal(0.0, 1.0, size=(N_OUT, N_HID))
    else:
        w1 = rng.normal(0.0, 1.0 / np.sqrt(N_IN), size=(N_HID, N_IN))
        w2 = rng.normal(0.0, 1.0 / np.sqrt(N_HID), size=(N_OUT, N_HID))
    return t, n, labels, w1.astype(np.float32), w2.astype(np.float32)


# ---------------------------------------------------------------- GPU path
def build_network(w1, w2):
    from bats_amd import Network
    from bats_amd.Layers import InputLayer, LIFLayer
    net = Network()
    inp = InputLayer(n_neurons=N_IN, name="input")
    net.add_layer(inp, input=True)
    hid = LIFLayer(previous_layer=inp, n_neurons=N_HID, tau_s=TAU_S, theta=THETA_H,
                   delta_theta=THETA_H, weight_initializer=lambda a, b: w1, max_n_spike=K_HID,
                   name="hidden")
    net.add_layer(hid)
    out = LIFLayer(previous_layer=hid, n_neurons=N_OUT, tau_s=TAU_S, theta=THETA_O,
                   delta_theta=THETA_O, weight_initializer=lambda a, b: w2, max_n_spike=K_OUT,
                   name="output")
    net.add_layer(out)
    return net, hid, out


def gpu_bench(args):
    import torch
    from bats_amd import ops
    from bats_amd.Losses import SpikeCountClassLoss
    from bats_amd.Optimizers import AdamOptimizer

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    t, n, labels, w1, w2 = make_inputs(args.seed, args.batch, args.variant)
    net, hid, out = build_network(w1, w2)
    t_d = torch.as_tensor(t, device=dev)
    n_d = torch.as_tensor(n, device=dev)
    lab_d = torch.as_tensor(labels, device=dev)
    loss_fct = SpikeCountClassLoss(target_true=TARGET_TRUE, target_false=TARGET_FALSE)
    # stream launches keep Adam's step counter on the host (as the reference does); the
    # captured graph needs it on the device
    opt_stream = AdamOptimizer(learning_rate=LR)
    opt_graph = AdamOptimizer(learning_rate=LR, capturable=True)
    opt = opt_stream

    def step():
        net.reset()
        net.forward(t_d, n_d, max_simulation=T_MAX, training=True)
        out_t, out_n = net.output_spike_trains
        loss, errors = loss_fct.compute_loss_and_errors(out_t, out_n, lab_d)
        grads = net.backward(errors)
        # the layers return the batch mean as [1, J, I]; dropping the unit batch axis is the
        # reference caller's mean over axis 0
        avg = [None if g is None else g[0] for g in grads]
        deltas = opt.step(avg)
        net.apply_deltas(deltas)
        return loss

    # Training changes the weights and with them the spike activity (and the cost) of later
    # steps, so every phase below starts from the initial weights and a fresh Adam state: the
    # kernel timing pass, the stream launches and the graph replays all run the same W + K
    # steps from the configured network.
    w_init = [hid.weights.clone(), out.weights.clone()]

    def restart(o):
        hid.weights.copy_(w_init[0])
        out.weights.copy_(w_init[1])
        if o._m is not None:
            for m in o._m + o._v:
                if m is not None:
                    m.zero_()
        o._t = 0
        if o._t_dev is not None:
            o._t_dev.zero_()
        torch.cuda.synchronize()

    # eager warm-up (lazy library init, Adam state allocation)
    for _ in range(3):
        step()
    restart(opt)
    net.reset()
    net.forward(t_d, n_d, max_simulation=T_MAX, training=True)
    hid_spikes = float(hid.spike_trains[1].sum().item()) / args.batch
    out_spikes = float(out.spike_trains[1].sum().item()) / args.batch

    # per-kernel durations: library-side HIP events on the launch stream, separate pass
    ops.timing(enable=True, reset=True)
    for _ in range(args.eager_steps):
        step()
    torch.cuda.synchronize()
    ops.timing(enable=False)
    kern = {}
    for name in KERNELS:
        ms, cnt = ops.timing_query(name)
        kern[name] = (ms / cnt if cnt else None, cnt)
    ops.timing(reset=True)

    # stream launches: W warm-up steps, then exactly K timed steps; each step is also bracketed
    # by HIP events on the launch stream (the library launches on torch's current stream)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    restart(opt)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record()
        step()
        ev[k][1].record()
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    eager_steps = [a.elapsed_time(b) for a, b in ev]
    hid_spikes_last = float(hid.spike_trains[1].sum().item()) / args.batch

    # graph capture of the whole step
    opt = opt_graph
    s = torch.cuda.Stream(device=dev)
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    restart(opt)
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record()
        graph.replay()
        ev[k][1].record()
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    graph_steps = [a.elapsed_time(b) for a, b in ev]
    return dict(graph_ms=graph_ms, eager_ms=eager_ms, kernels=kern, hid_spikes=hid_spikes,
                out_spikes=out_spikes, hid_spikes_last=hid_spikes_last,
                eager_steps=eager_steps, graph_steps=graph_steps)


def literal_bench(args):
    """The same step on the literal initialisation N(0, 1/sqrt(fan_in)) (same seed and inputs),
    stream launches, W warm-up + K timed steps: the hidden layer is silent there, so this times
    the path with empty solver walks and an all-zero backward."""
    import torch
    from bats_amd.Losses import SpikeCountClassLoss
    from bats_amd.Optimizers import AdamOptimizer
    dev = torch.device("cuda", 0)
    t, n, labels, w1, w2 = make_inputs(args.seed, args.batch, "literal")
    net, hid, out = build_network(w1, w2)
    t_d, n_d = torch.as_tensor(t, device=dev), torch.as_tensor(n, device=dev)
    lab_d = torch.as_tensor(labels, device=dev)
    loss_fct = SpikeCountClassLoss(target_true=TARGET_TRUE, target_false=TARGET_FALSE)
    opt = AdamOptimizer(learning_rate=LR)

    def step():
        net.reset()
        net.forward(t_d, n_d, max_simulation=T_MAX, training=True)
        out_t, out_n = net.output_spike_trains
        _, errors = loss_fct.compute_loss_and_errors(out_t, out_n, lab_d)
        grads = net.backward(errors)
        net.apply_deltas(opt.step([None if g is None else g[0] for g in grads]))

    step()
    spikes = [float(hid.spike_trains[1].sum().item()) / args.batch,
              float(out.spike_trains[1].sum().item()) / args.batch]
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    return {"workload": "same step, W~N(0,1/sqrt(fan_in)) literal initialisation",
            "ms_per_step": round(ms, 4), "samples_per_s": round(args.batch / (ms * 1e-3), 1),
            "spikes_per_sample": {"hidden": round(spikes[0], 1), "output": round(spikes[1], 2)}}


def _step_stats(ms_list):
    a = np.asarray(ms_list, dtype=np.float64)
    return {"median": round(float(np.median(a)), 4), "p10": round(float(np.percentile(a, 10)), 4),
            "p90": round(float(np.percentile(a, 90)), 4), "first": round(float(a[0]), 4),
            "last": round(float(a[-1]), 4)}


# ---------------------------------------------------------------- plain form (baseline)
def plain_bench(args):
    """The plain FP32 straight-translation form of the same step, eager (its structure has a
    device-to-host read per layer, so it cannot be captured), on the same inputs/weights."""
    import torch
    from bats_amd import Network
    from bats_amd.Layers import InputLayer
    from bats_amd.plain import PlainAdam, PlainLIFLayer, PlainSpikeCountClassLoss

    dev = torch.device("cuda", 0)
    t, n, labels, w1, w2 = make_inputs(args.seed, args.batch, args.variant)
    net = Network()
    inp = InputLayer(n_neurons=N_IN, name="input")
    net.add_layer(inp, input=True)
    hid = PlainLIFLayer(inp, tau_s=TAU_S, theta=THETA_H, delta_theta=THETA_H, n_neurons=N_HID,
                        weight_initializer=lambda a, b: w1, max_n_spike=K_HID, name="hidden")
    net.add_layer(hid)
    out = PlainLIFLayer(hid, tau_s=TAU_S, theta=THETA_O, delta_theta=THETA_O, n_neurons=N_OUT,
                        weight_initializer=lambda a, b: w2, max_n_spike=K_OUT, name="output")
    net.add_layer(out)
    t_d = torch.as_tensor(t, device=dev)
    n_d = torch.as_tensor(n, device=dev)
    lab_d = torch.as_tensor(labels, device=dev)
    loss_fct = PlainSpikeCountClassLoss(TARGET_TRUE, TARGET_FALSE)
    opt = PlainAdam(LR)

    def step():
        net.reset()
        net.forward(t_d, n_d, max_simulation=T_MAX, training=True)
        out_t, out_n = net.output_spike_trains
        loss, errors = loss_fct.compute_loss_and_errors(out_t, out_n, lab_d)
        grads = net.backward(errors)
        avg = [None if g is None else g.mean(dim=0) for g in grads]
        deltas = opt.step(avg)
        net.apply_deltas(deltas)

    net.forward(t_d, n_d, max_simulation=T_MAX)
    counts = (hid.spike_trains[1].cpu().numpy(), out.spike_trains[1].cpu().numpy())
    # the same W warm-up + K timed steps from the initial weights as the optimised path
    K = args.plain_steps if args.plain_steps else args.steps
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / K
    return dict(ms=ms, counts=counts, steps=K)


def serve_bench(args):
    """Inference serving of the same network: one request = one event list of
    P ~ U{50..2000} events, input index ~ U{0..783} with replacement (several spikes per input
    neuron), times ~ U[0, 1]. Measures the saturated throughput (all requests queued at once),
    then p50/p99 latency (arrival -> result on the host) under Poisson arrivals at 25%, 50% and
    90% of it, and the output spike counts of the first batch against the plain form's."""
    import torch
    from bats_amd.serving import InferenceServer, LayerSpec, make_requests, run_poisson

    _, _, _, w1, w2 = make_inputs(args.seed, 8, args.variant)
    layers = [LayerSpec(w1, TAU_S, THETA_H, THETA_H, K_HID),
              LayerSpec(w2, TAU_S, THETA_O, THETA_O, K_OUT)]
    # length classes cut at 256, 512 and 1024 events; padded shapes: 2048 events (the request
    # mix's upper end), batches of 8, 32, 128, 512 and max_batch, two copies of each (one per
    # stream); the captured graphs are held within the memory budget
    srv = InferenceServer(layers, N_IN, max_batch=args.serve_batch, max_simulation=T_MAX,
                          min_events_bucket=2048, batch_growth=4,
                          length_classes=(256, 512, 1024),
                          memory_budget=int(args.serve_memory_gib * (1 << 30)),
                          arena_events=1 << 25)
    seed = [args.seed + 1]

    def requests(count):
        # written straight into the server's request arena (as a front end receiving them
        # into it would), before any clock starts
        seed[0] += 1
        return make_requests(count, n_inputs=N_IN, p_min=50, p_max=2000, seed=seed[0],
                             server=srv)

    shapes, Bb = [], 8
    while True:
        shapes.append((min(Bb, args.serve_batch), 2048))
        if Bb >= args.serve_batch:
            break
        Bb *= 4
    srv.warmup(shapes, copies=2)
    agreement = plain_count_agreement(args, layers, requests(256), srv)
    # saturated throughput
    reqs = requests(args.serve_requests)
    p_sat = [len(r.indices) for r in reqs]
    srv.completed = []
    gc.collect()
    t0 = time.perf_counter()
    for r in reqs:
        r.arrival = t0
        srv.submit(r)
    srv.drain()
    torch.cuda.synchronize()
    sat = len(srv.completed) / (time.perf_counter() - t0)
    # untimed warm-up under arrivals: the first overlap of batches on the two streams carries
    # one-time runtime costs (milliseconds) that are not part of the steady state
    run_poisson(srv, requests(4000), 0.5 * sat, seed=args.seed + 100)
    loads = []
    for frac in (0.25, 0.5, 0.9):
        rate = frac * sat
        n_req = int(min(args.serve_requests, max(200, rate * args.serve_seconds)))
        res = run_poisson(srv, requests(n_req), rate, seed=args.seed)
        loads.append({"load": frac, "offered_per_s": round(rate, 1),
                      "served_per_s": round(res["samples_per_s"], 1),
                      "p50_ms": round(res["p50_ms"], 3), "p99_ms": round(res["p99_ms"], 3),
                      "requests": res["n"]})
    # saturated throughput against the HBM roofline: algorithmic bytes of one request are its
    # event list (index + time), the API-visible spike trains of both layers (times and
    # counts) and its prediction; the weights count once per batch of max_batch requests
    p_mean = float(np.mean(p_sat))
    per_req = (8.0 * p_mean + 4.0 * (N_HID * K_HID + N_HID) + 4.0 * (N_OUT * K_OUT + N_OUT)
               + 8.0)
    weights = 4.0 * (N_IN * N_HID + N_HID * N_OUT)
    achieved = (per_req * sat + weights * sat / args.serve_batch) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": 8000.0, "unit": "GB/s",
            "frac": round(achieved / 8000.0, 5),
            "bytes_per_request": round(per_req, 1), "mean_events_per_request": round(p_mean, 1)}
    return {"saturated_samples_per_s": round(sat, 1), "max_batch": args.serve_batch,
            "requests": "P ~ U{50..2000} events, idx ~ U{0..783} with replacement, t ~ U[0,1]",
            "batching": "length classes cut at 256/512/1024 events (oldest request's class "
                        "plus shorter ones), longest first; batch buckets 8/32/128/512/max_batch; "
                        "two streams",
            "memory_budget_gib": args.serve_memory_gib,
            "graphs_resident_gib": round(srv.memory_in_use / (1 << 30), 2),
            "graphs_made": srv.graphs_made, "graphs_evicted": srv.graphs_evicted,
            "roofline": roof, "count_agreement_with_plain": agreement, "loads": loads}


def plain_count_agreement(args, layers, reqs, srv):
    """Serves `reqs` and runs the plain form on the same requests (per-neuron buffers built
    from the event lists); returns the fraction of equal output spike counts."""
    import torch
    from bats_amd import Network, ops
    from bats_amd.Layers import InputLayer
    from bats_amd.plain import PlainLIFLayer

    # the plain form's padded lists first: a served request's events leave the arena
    P = max(len(r.indices) for r in reqs)
    idx = np.full((len(reqs), P), -1, dtype=np.int32)
    tm = np.full((len(reqs), P), np.inf, dtype=np.float32)
    for k, r in enumerate(reqs):
        idx[k, :len(r.indices)] = r.indices
        tm[k, :len(r.times)] = r.times
    rids = [r.rid for r in reqs]
    srv.completed = []
    for r in reqs:
        srv.submit(r)
    srv.drain()
    torch.cuda.synchronize()
    served = {r.rid: r.counts for r in srv.completed}
    per_t, per_n = ops.events_to_per_neuron(idx, tm, N_IN)
    net = Network()
    inp = InputLayer(n_neurons=N_IN, name="input")
    net.add_layer(inp, input=True)
    prev = inp
    for L, name in zip(layers, ("hidden", "output")):
        lay = PlainLIFLayer(prev, tau_s=L.tau_s, theta=L.theta, delta_theta=L.delta_theta,
                            n_neurons=L.weights.shape[0], weight_initializer=lambda a, b, w=L.weights: w,
                            max_n_spike=L.max_n_spike, name=name)
        net.add_layer(lay)
        prev = lay
    net.forward(per_t, per_n, max_simulation=T_MAX)
    plain_n = prev.spike_trains[1].cpu().numpy()
    same = [np.array_equal(served[rid], plain_n[k]) for k, rid in enumerate(rids)]
    return float(np.mean(same))


def first_forward_counts(args):
    """Spike counts of the optimised path's first forward on the bench inputs."""
    import torch
    t, n, labels, w1, w2 = make_inputs(args.seed, args.batch, args.variant)
    net, hid, out = build_network(w1, w2)
    net.forward(torch.as_tensor(t, device="cuda"), torch.as_tensor(n, device="cuda"),
                max_simulation=T_MAX)
    return hid.spike_trains[1].cpu().numpy(), out.spike_trains[1].cpu().numpy()


# ---------------------------------------------------------------- CPU baseline (oracle)
def _oracle_steps(job):
    """One CPU worker: the FP64 NumPy oracle's training step on its shard with BLAS on one
    thread, one warm-up step then `reps` timed steps; returns (samples per step, per-step
    seconds)."""
    seed, batch, variant, lo, hi, reps = job
    from threadpoolctl import threadpool_limits
    from oracle import lif as olif
    from oracle import losses as oloss
    from oracle import optim as ooptim

    t, n, labels, w1, w2 = make_inputs(seed, batch, variant)
    idx = np.arange(lo, hi) % batch
    t, labels = t[idx], labels[idx]
    times = []
    with threadpool_limits(limits=1):
        for _ in range(1 + reps):
            t0 = time.perf_counter()
            h = olif.lif_forward(t, w1, TAU_S, THETA_H, THETA_H, K_HID, tmax=T_MAX)
            o = olif.lif_forward(h.t, w2, TAU_S, THETA_O, THETA_O, K_OUT, tmax=T_MAX)
            tg = oloss.count_targets(o.n, labels, TARGET_TRUE, TARGET_FALSE)
            _, err = oloss.count_loss_and_errors(o.n, tg, K_OUT)
            _, g2, pe = olif.lif_backward(o, err, pre_trainable=True)
            _, g1, _ = olif.lif_backward(h, pe, pre_trainable=False)
            opt = ooptim.Adam(LR)
            opt.step([olif.batch_mean(g1), olif.batch_mean(g2)])
            times.append(time.perf_counter() - t0)
    return hi - lo, times[1:]


def cpu_share():
    """(os.cpu_count(), cores in this process's affinity set, the cgroup CPU quota in cores or
    None): the host cores the CPU baseline may use."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    for path in ("/sys/fs/cgroup/cpu.max",):
        try:
            with open(path) as f:
                q, period = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(period)
        except (OSError, ValueError):
            pass
    if quota is None:
        try:
            with open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us") as f:
                q = float(f.read())
            with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as f:
                period = float(f.read())
            if q > 0:
                quota = q / period
        except (OSError, ValueError):
            pass
    return os.cpu_count(), aff, quota


def cpu_baseline(args):
    """The oracle's training step timed as BASELINE.md's CPU baseline asks: one process per
    usable host core (the affinity set, capped by a cgroup CPU quota when there is one), each
    on its own shard of the batch with BLAS/OpenMP on one thread, one warm-up step then the
    median of `args.cpu_reps` steps; and the same on one core alone. Runs before the GPU is
    initialised (the workers are forked)."""
    import multiprocessing as mp
    n_cpu, aff, quota = cpu_share()
    procs = aff if quota is None else max(1, min(aff, int(quota)))
    reps = args.cpu_reps
    one_n, one_t = _oracle_steps((args.seed, args.batch, args.variant, 0, args.cpu_sample, reps))
    one = one_n / float(np.median(one_t))
    per = args.cpu_shard
    jobs = [(args.seed, args.batch, args.variant, k * per, k * per + per, reps)
            for k in range(procs)]
    with mp.get_context("fork").Pool(procs) as pool:
        t0 = time.perf_counter()
        res = pool.map(_oracle_steps, jobs)
        wall = time.perf_counter() - t0
    all_rate = sum(nn / float(np.median(tt)) for nn, tt in res)
    return dict(one=one, one_samples=one_n, all=all_rate, all_seconds=wall, cores=procs,
                per=per, cpu_count=n_cpu, affinity=aff, quota=quota, reps=reps)


# ---------------------------------------------------------------- the other configurations
def _time_steps(step, warmup, steps):
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def config_legs(args):
    """Training-step time of the other network configurations (stream launches, eager; each
    starts from its seeded weights): the single-spike TTFS network, the deep dense network
    and the convolutional network. Thresholds and weight scales are stated in the output;
    `spikes` is the mean number of spikes per sample of each LIF layer in the first step."""
    import torch
    from bats_amd import Network
    from bats_amd.Layers import (ConvInputLayer, ConvLIFLayer, InputLayer, LIFLayer,
                                 PoolingLayer)
    from bats_amd.Losses import SpikeCountClassLoss, TTFSSoftmaxCrossEntropy
    from bats_amd.Optimizers import AdamOptimizer

    dev = torch.device("cuda", 0)
    out = {}

    def dense_net(sizes, Ks, thetas, scale):
        net = Network()
        inp = InputLayer(n_neurons=sizes[0], name="input")
        net.add_layer(inp, input=True)
        prev = inp
        for k, (n_post, K, th) in enumerate(zip(sizes[1:], Ks, thetas)):
            w = (scale * rng.normal(0.0, 1.0, size=(n_post, sizes[k]))).astype(np.float32)
            lay = LIFLayer(previous_layer=prev, n_neurons=n_post, tau_s=TAU_S, theta=th,
                           delta_theta=th, weight_initializer=lambda a, b, w=w: w,
                           max_n_spike=K, name=f"l{k}")
            net.add_layer(lay)
            prev = lay
        return net

    wanted = [w.strip() for w in args.configs.split(",") if w.strip()]

    def run(name, net, batch, inputs, loss_fct, desc, steps=10, warmup=3):
        if not any(name.startswith(w) for w in wanted):
            return
        t_d, n_d = inputs
        lab = torch.as_tensor(rng.integers(0, 10, size=batch), device=dev)
        opt = AdamOptimizer(learning_rate=LR)

        def step():
            net.reset()
            net.forward(t_d, n_d, max_simulation=T_MAX, training=True)
            ot, on = net.output_spike_trains
            _, err = loss_fct.compute_loss_and_errors(ot, on, lab)
            grads = net.backward(err)
            net.apply_deltas(opt.step([None if g is None else g[0] for g in grads]))

        step()
        spikes = [round(float(l.spike_trains[1].sum().item()) / batch, 1)
                  for l in net.layers[1:] if isinstance(l, (LIFLayer, ConvLIFLayer))]
        ms = _time_steps(step, warmup, steps)
        out[name] = {"workload": desc, "batch": batch, "ms_per_step": round(ms, 4),
                     "samples_per_s": round(batch / (ms * 1e-3), 1), "steps": steps,
                     "spikes": spikes}

    def latency(batch):
        x = rng.uniform(0.0, 1.0, size=(batch, N_IN))
        t = np.where(x > 0.1, 1.0 - x, np.inf).astype(np.float32)[:, :, None]
        return (torch.as_tensor(t, device=dev),
                torch.as_tensor(np.isfinite(t[:, :, 0]).astype(np.int32), device=dev))

    # every leg draws its weights and inputs from its own generator, so that a leg's workload
    # does not depend on which other legs run
    # single-spike TTFS network (784-800-10, K = 1, batch 50, softmax on first spike times)
    rng = np.random.default_rng(args.seed + 1)
    run("ttfs_784-800-10_b50", dense_net([N_IN, N_HID, N_OUT], [1, 1], [1.0, 1.3], 1.0), 50,
        latency(50), TTFSSoftmaxCrossEntropy(tau=0.005),
        "784-800-10, K=1/1, theta 1.0/1.3, W~N(0,1), TTFS loss tau 0.005, Adam")
    # deep dense network (784-800-800-800-10, K = 30, batch 256)
    rng = np.random.default_rng(args.seed + 2)
    run("deep_784-800x3-10_b256",
        dense_net([N_IN, 800, 800, 800, N_OUT], [30, 30, 30, 30], [1.0, 1.0, 1.0, 1.3],
                  1.0), 256, latency(256), SpikeCountClassLoss(target_true=15, target_false=3),
        "784-800-800-800-10, K=30, theta 1.0 (1.3 output), W~N(0,1), count loss 15/3, Adam")
    # convolutional network (batch 128): conv5x5x15 (K 30) -> pool -> conv5x5x40 (K 3) ->
    # pool -> 300 (K 10) -> 10 (K 30), thresholds of experiments/mnist/train_conv.py
    rng = np.random.default_rng(args.seed + 3)
    net = Network()
    inp = ConvInputLayer(neurons_shape=np.array([28, 28, 1]), name="input")
    net.add_layer(inp, input=True)
    # N(0, 1/sqrt(fan_in)) with gains 1, 8, 8, 8: the plain scale leaves conv2 and every
    # later layer silent at these thresholds; with the gains every layer fires
    w1 = rng.normal(0.0, 1.0 / 5.0, size=(15, 5, 5, 1)).astype(np.float32)
    w2 = (8.0 * rng.normal(0.0, 1.0 / np.sqrt(5 * 5 * 15), size=(40, 5, 5, 15))).astype(np.float32)
    w3 = (8.0 * rng.normal(0.0, 1.0 / np.sqrt(640), size=(300, 640))).astype(np.float32)
    w4 = (8.0 * rng.normal(0.0, 1.0 / np.sqrt(300), size=(10, 300))).astype(np.float32)
    c1 = ConvLIFLayer(inp, filters_shape=np.array([5, 5, 15]), tau_s=TAU_S, theta=0.04,
                      delta_theta=0.04, weight_initializer=lambda *a: w1, max_n_spike=30,
                      name="conv1")
    net.add_layer(c1)
    p1 = PoolingLayer(c1, name="pool1")
    net.add_layer(p1)
    c2 = ConvLIFLayer(p1, filters_shape=np.array([5, 5, 40]), tau_s=TAU_S, theta=0.8,
                      delta_theta=0.8, weight_initializer=lambda *a: w2, max_n_spike=3,
                      name="conv2")
    net.add_layer(c2)
    p2 = PoolingLayer(c2, name="pool2")
    net.add_layer(p2)
    d3 = LIFLayer(p2, n_neurons=300, tau_s=TAU_S, theta=0.6, delta_theta=0.6,
                  weight_initializer=lambda *a: w3, max_n_spike=10, name="fc")
    net.add_layer(d3)
    d4 = LIFLayer(d3, n_neurons=10, tau_s=TAU_S, theta=0.3, delta_theta=0.3,
                  weight_initializer=lambda *a: w4, max_n_spike=30, name="output")
    net.add_layer(d4)
    run("conv_28x28-c15-p-c40-p-300-10_b128", net, 128, latency(128),
        SpikeCountClassLoss(target_true=15, target_false=3),
        "conv5x5x15 K30 -> pool -> conv5x5x40 K3 -> pool -> 300 K10 -> 10 K30, thresholds "
        "0.04/0.8/0.6/0.3, W~N(0,1/sqrt(fan_in)) x gains 1/8/8/8 (active variant), count loss "
        "15/3, Adam")
    return out


# Kernel sources whose text fixes lif_solve_seq_kernel's memory traffic: the measured figure in
# profiles/traffic.json is used only while their hash equals the one stored with it.
TRAFFIC_SOURCES = ("bats_amd/csrc/lif_forward.hip", "bats_amd/csrc/solver.hpp",
                   "bats_amd/csrc/common.hpp")


def traffic_source_hash():
    import hashlib
    root = os.path.dirname(os.path.abspath(__file__))
    h = hashlib.sha256()
    for rel in TRAFFIC_SOURCES:
        with open(os.path.join(root, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(kernel):
    """(HBM bytes per launch, note) of `kernel` at this workload, from the calibrated
    FETCH_SIZE / WRITE_SIZE counters recorded in profiles/traffic.json; the bytes are None when
    the record is absent or was measured on other kernel sources than the current ones."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)[kernel]
    except (OSError, KeyError, ValueError):
        return None, "no traffic record"
    cur = traffic_source_hash()
    if rec.get("source_hash") != cur:
        return None, (f"stale: measured on sources {rec.get('source_hash')}, current {cur}")
    return rec["bytes"], f"profiles/traffic.json, sources {cur}"


def main():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--eager-steps", type=int, default=20)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--variant", choices=("active", "literal"), default="active")
    ap.add_argument("--cpu-sample", type=int, default=32,
                    help="samples of the one-core CPU baseline")
    ap.add_argument("--cpu-shard", type=int, default=16,
                    help="samples per process of the all-core CPU baseline")
    ap.add_argument("--cpu-reps", type=int, default=5,
                    help="timed CPU steps per process (after one warm-up step)")
    ap.add_argument("--skip-literal", action="store_true",
                    help="skip the literal-initialisation variant")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=0,
                    help="timed steps of the plain form (default: --steps)")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--serve-batch", type=int, default=1024)
    ap.add_argument("--serve-memory-gib", type=float, default=48.0,
                    help="device memory budget of the server's captured batch graphs")
    ap.add_argument("--serve-requests", type=int, default=20000)
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--configs", default="ttfs,deep,conv",
                    help="training-step legs to run (comma list of ttfs, deep, conv)")
    ap.add_argument("--skip-configs", action="store_true",
                    help="skip the TTFS, deep and conv training-step legs")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (the batch would "
              "shard over GPUs with one gradient all-reduce per step); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    c = None if args.skip_cpu else cpu_baseline(args)  # before the GPU is initialised
    r = gpu_bench(args)
    B = args.batch
    # whole-job rate of the faster launch mode (both time exactly K steps)
    launch = "graph" if r["graph_ms"] <= r["eager_ms"] else "stream"
    ms = min(r["graph_ms"], r["eager_ms"])
    value = B / (ms * 1e-3)

    # roofline of the dominant kernel: the hidden layer's lane-per-neuron solver (chained to
    # the output layer: it also builds that layer's event lists, which are internal data and
    # add no algorithmic bytes)
    solve_ms, solve_cnt = r["kernels"]["lif_solve_seq_kernel"]
    alg_bytes = 4 * N_IN * N_HID + 4 * (B * N_HID * K_HID + B * N_HID)
    traffic, traffic_note = measured_traffic("lif_solve_seq_kernel")
    roof = None
    if solve_ms:
        achieved = alg_bytes / (solve_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": traffic if B == 256 else None, "traffic_source": traffic_note,
                "kernel": "lif_solve_seq_kernel (784->800 layer, chained to the output layer)",
                "kernel_avg_us": round(solve_ms * 1e3, 2), "algorithmic_bytes": alg_bytes}

    plain = None
    if not args.skip_plain:
        pr = plain_bench(args)
        opt_counts = first_forward_counts(args)
        plain_value = B / (pr["ms"] * 1e-3)
        plain = {"value": round(plain_value, 1), "unit": "samples/s",
                 "ms_per_step": round(pr["ms"], 3), "steps": pr["steps"],
                 "speedup": round(value / plain_value, 2),
                 "hidden_count_agreement": float(np.mean(opt_counts[0] == pr["counts"][0])),
                 "output_count_agreement": float(np.mean(opt_counts[1] == pr["counts"][1]))}

    literal = None if (args.skip_literal or args.variant == "literal") else literal_bench(args)
    serve = None if args.skip_serve else serve_bench(args)
    legs = None if args.skip_configs else config_legs(args)

    cpu = None
    if c is not None:
        cpu = {"value": round(c["all"], 3), "unit": "samples/s", "cores": c["cores"],
               "kind": "port",
               "sample": f"full training step of the FP64 NumPy oracle (BLAS/OpenMP on 1 thread "
                         f"per process), {c['cores']} forked processes at once, each on its own "
                         f"shard of {c['per']} samples of the batch, median of {c['reps']} steps "
                         f"after 1 warm-up per process ({c['all_seconds']:.1f} s wall); 1 core: "
                         f"{c['one']:.3f} samples/s on {c['one_samples']} samples (median of "
                         f"{c['reps']} after 1 warm-up)",
               "value_1_core": round(c["one"], 3), "os_cpu_count": c["cpu_count"],
               "affinity_cores": c["affinity"],
               "cgroup_cpu_quota": None if c["quota"] is None else round(c["quota"], 2)}

    step_bytes = 176e6 * B / 256.0
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "784-800-10 LIF train step, K=30/30, count loss 15/3, Adam, "
                               f"t_max=1.0, W~{'N(0,1) active' if args.variant == 'active' else 'N(0,1/sqrt(fan_in)) literal'} variant",
                   "global_batch": B, "seed": args.seed, "launch": launch},
        "roofline": roof, "cpu_baseline": cpu, "plain_form": plain, "serve": serve,
        "other_configs": legs,
        "stream_ms_per_step": round(r["eager_ms"], 4),
        "graph_ms_per_step": round(r["graph_ms"], 4),
        "step_ms_hip_events": {"stream": _step_stats(r["eager_steps"]),
                               "graph": _step_stats(r["graph_steps"])},
        "literal_variant": literal,
        "step_hbm_fraction": round(step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "spikes_per_sample": {"hidden": round(r["hid_spikes"], 1), "output": round(r["out_spikes"], 2),
                              "hidden_last_timed_step": round(r["hid_spikes_last"], 1)},
        "kernel_avg_us": {k: (None if v[0] is None else round(v[0] * 1e3, 2))
                          for k, v in r["kernels"].items()},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
