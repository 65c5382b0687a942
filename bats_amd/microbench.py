"""Per-kernel microbenchmark of the dense forward / backward at the 784-800-10 training shape,
for profiling one kernel family at a time:

    python -m bats_amd.microbench --what forward1 --reps 20

It runs only the requested operator (after one warm-up) on the same synthetic inputs as
bench.py, plus one elementwise copy of a known size used to calibrate HBM traffic counters.
The conv_* operators are the two conv layers of bench.py's convolutional leg (batch 128).
"""
from __future__ import annotations

import argparse
import time

import numpy as np
import torch

from . import ops
from ._constants import lif_constants


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--what", choices=("forward1", "forward1_chain", "forward2", "backward1",
                                       "backward2",
                                       "forward_hh", "backward_hh", "conv_forward1",
                                       "conv_forward2", "conv_backward1", "conv_backward2",
                                       "copy"),
                    default="forward1")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--solver-mode", type=int, default=0,
                    help="solver form of the timed dense forwards (bats_lif_forward_solver)")
    args = ap.parse_args()
    sm = args.solver_mode
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(0)
    B = args.batch
    x = rng.uniform(size=(B, 784))
    t = torch.as_tensor(np.where(x > 0.1, 1.0 - x, np.inf).astype(np.float32)[:, :, None], device=dev)
    w1 = torch.as_tensor(rng.normal(0, 1.0, size=(800, 784)).astype(np.float32), device=dev)
    w2 = torch.as_tensor(rng.normal(0, 1.0, size=(10, 800)).astype(np.float32), device=dev)
    c1 = lif_constants(0.130, 1.0, 1.0)
    c2 = lif_constants(0.130, 1.3, 1.3)
    h = ops.lif_forward(t, w1, *c1, 1.0, 30)
    o = ops.lif_forward(h[1], w2, *c2, 1.0, 30)
    err2 = torch.randn(o[1].shape, device=dev)
    err1 = torch.randn(h[1].shape, device=dev)
    if args.what in ("forward_hh", "backward_hh"):
        # 800 -> 800 hidden layer of the 784-800-800-800-10 network, fed by the first
        # hidden layer's multi-spike trains
        w3 = torch.as_tensor(rng.normal(0, 1.0, size=(800, 800)).astype(np.float32), device=dev)
        hh = ops.lif_forward(h[1], w3, *c1, 1.0, 30)
        err3 = torch.randn(hh[1].shape, device=dev)
    if args.what.startswith("conv_"):
        # the two conv layers of the 28x28-c15-p-c40-p-300-10 network at batch 128, with the
        # weights and inputs that bench.py's conv leg draws (same generator and order)
        crng = np.random.default_rng(3)
        cw1 = torch.as_tensor(crng.normal(0.0, 1.0 / 5.0, size=(15, 5, 5, 1)).astype(np.float32),
                              device=dev)
        cw2 = torch.as_tensor((8.0 * crng.normal(0.0, 1.0 / np.sqrt(375), size=(40, 5, 5, 15)))
                              .astype(np.float32), device=dev)
        crng.normal(size=(300, 640))
        crng.normal(size=(10, 300))
        cx = crng.uniform(0.0, 1.0, size=(128, 784))
        ct = torch.as_tensor(np.where(cx > 0.1, 1.0 - cx, np.inf).astype(np.float32)[:, :, None],
                             device=dev)
        k1 = lif_constants(0.130, 0.04, 0.04)
        k2 = lif_constants(0.130, 0.8, 0.8)
        o1 = ops.conv_lif_forward(ct, (28, 28, 1), cw1, *k1, 1.0, 30)
        p1n, p1t, _ = ops.pool_forward(o1[0], o1[1], (24, 24, 15))
        o2 = ops.conv_lif_forward(p1t, (12, 12, 15), cw2, *k2, 1.0, 3)
        ce1 = torch.randn(o1[1].shape, device=dev)
        ce2 = torch.randn(o2[1].shape, device=dev)
    src = torch.empty(args.copy_mib * 2 ** 18, dtype=torch.float32, device=dev).uniform_()

    def run():
        if args.what == "forward1":
            ops.lif_forward(t, w1, *c1, 1.0, 30, solver_mode=sm)
        elif args.what == "forward1_chain":
            # the hidden layer chained to the output layer (its solver also builds the output
            # layer's event lists), as the training step runs it
            ops.lif_forward_chain(t, w1, *c1, 1.0, 30, (c2[0], c2[1], c2[2], 1.0))
        elif args.what == "forward2":
            ops.lif_forward(h[1], w2, *c2, 1.0, 30, solver_mode=sm)
        elif args.what == "backward1":
            ops.lif_backward(*h[:5], err1.clone(), t, h[5], h[6], w1, *c1)
        elif args.what == "backward2":
            ops.lif_backward(*o[:5], err2.clone(), h[1], o[5], o[6], w2, *c2)
        elif args.what == "forward_hh":
            ops.lif_forward(h[1], w3, *c1, 1.0, 30, solver_mode=sm)
        elif args.what == "backward_hh":
            ops.lif_backward(*hh[:5], err3.clone(), h[1], hh[5], hh[6], w3, *c1)
        elif args.what == "conv_forward1":
            ops.conv_lif_forward(ct, (28, 28, 1), cw1, *k1, 1.0, 30)
        elif args.what == "conv_forward2":
            ops.conv_lif_forward(p1t, (12, 12, 15), cw2, *k2, 1.0, 3)
        elif args.what == "conv_backward1":
            ops.conv_lif_backward(*o1[:5], ce1.clone(), ct, (28, 28, 1), o1[5], o1[6], cw1, *k1,
                                  pre_trainable=False)
        elif args.what == "conv_backward2":
            ops.conv_lif_backward(*o2[:5], ce2.clone(), p1t, (12, 12, 15), o2[5], o2[6], cw2,
                                  *k2, pre_trainable=True)
        else:
            torch.add(src, 0.0)

    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.reps):
        run()
    torch.cuda.synchronize()
    extra = ""
    if args.what.startswith("conv_"):
        extra = (f", conv1 / pool1 / conv2 spikes per sample {o1[0].sum().item() / 128:.1f} / "
                 f"{p1n.sum().item() / 128:.1f} / {o2[0].sum().item() / 128:.1f}")
    print(f"{args.what}: {(time.perf_counter() - t0) * 1e3 / args.reps:.3f} ms/call "
          f"(hidden spikes/sample {h[0].sum().item() / B:.1f}{extra})", flush=True)


if __name__ == "__main__":
    main()
