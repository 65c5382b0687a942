"""Per-kernel microbenchmark of the dense forward / backward at the 784-800-10 training shape,
for profiling one kernel family at a time:

    python -m bats_amd.microbench --what forward1 --reps 20

It runs only the requested operator (after one warm-up) on the same synthetic inputs as
bench.py, plus one elementwise copy of a known size used to calibrate HBM traffic counters.
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
    ap.add_argument("--what", choices=("forward1", "forward2", "backward1", "backward2",
                                       "forward_hh", "backward_hh", "copy"),
                    default="forward1")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--copy-mib", type=int, default=256)
    args = ap.parse_args()
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
    src = torch.empty(args.copy_mib * 2 ** 18, dtype=torch.float32, device=dev).uniform_()

    def run():
        if args.what == "forward1":
            ops.lif_forward(t, w1, *c1, 1.0, 30)
        elif args.what == "forward2":
            ops.lif_forward(h[1], w2, *c2, 1.0, 30)
        elif args.what == "backward1":
            ops.lif_backward(*h[:5], err1.clone(), t, h[5], h[6], w1, *c1)
        elif args.what == "backward2":
            ops.lif_backward(*o[:5], err2.clone(), h[1], o[5], o[6], w2, *c2)
        elif args.what == "forward_hh":
            ops.lif_forward(h[1], w3, *c1, 1.0, 30)
        elif args.what == "backward_hh":
            ops.lif_backward(*hh[:5], err3.clone(), h[1], hh[5], hh[6], w3, *c1)
        else:
            torch.add(src, 0.0)

    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.reps):
        run()
    torch.cuda.synchronize()
    print(f"{args.what}: {(time.perf_counter() - t0) * 1e3 / args.reps:.3f} ms/call "
          f"(hidden spikes/sample {h[0].sum().item() / B:.1f})", flush=True)


if __name__ == "__main__":
    main()
