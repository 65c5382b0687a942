"""Event list -> per-neuron spike buffers.  Test infrastructure only."""
from __future__ import annotations

import numpy as np


def events_to_per_neuron(indices, times, n_neurons, max_per_neuron=None):
    """Restates bats/CudaKernels/CuFiles/Inference/convert_to_spike_per_neuron.cu:9-24 and
    the buffer depth of bats/CudaKernels/Wrappers/Inference/convert_to_spike_per_neuron.py:11-18
    (largest per-neuron multiplicity over the batch, when ``max_per_neuron`` is None).

    indices [B,P] int (a negative index ends a sample's list), times [B,P] float32.
    Returns (times_per_neuron [B,N,K] f32 +inf padded, counts [B,N] int32).
    """
    idx = np.asarray(indices)
    tim = np.asarray(times, dtype=np.float32)
    B, P = idx.shape
    counts = np.zeros((B, n_neurons), dtype=np.int32)
    ends = []
    for b in range(B):
        neg = np.nonzero(idx[b] < 0)[0]
        end = int(neg[0]) if neg.size else P
        ends.append(end)
        if end:
            np.add.at(counts[b], idx[b, :end], 1)
    K = int(counts.max()) if max_per_neuron is None else int(max_per_neuron)
    out = np.full((B, n_neurons, max(K, 0)), np.inf, dtype=np.float32)
    fill = np.zeros((B, n_neurons), dtype=np.int64)
    for b in range(B):
        for p in range(ends[b]):
            i = idx[b, p]
            if fill[b, i] < K:
                out[b, i, fill[b, i]] = tim[b, p]
            fill[b, i] += 1
    return out, counts
