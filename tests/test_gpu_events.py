"""Event list -> per-neuron buffers on the GPU (dense [B,P] lists and packed lists with row
offsets) against the oracle and the reference test's hand arrays. Bit-exact: the c-th slot of
a neuron holds its c-th event in list order, counts include the events dropped beyond K."""
import numpy as np
import pytest
import torch

from oracle import events as oev

pytestmark = pytest.mark.gpu
F32 = np.float32


def _dense_gpu(idx, tim, N, K):
    from bats_amd import ops
    t, n = ops.events_to_per_neuron(torch.as_tensor(idx, dtype=torch.int32, device="cuda"),
                                    torch.as_tensor(tim, dtype=torch.float32, device="cuda"),
                                    N, K)
    torch.cuda.synchronize()
    return t.cpu().numpy(), n.cpu().numpy()


def _packed(rows):
    off = np.zeros(len(rows) + 1, dtype=np.int32)
    off[1:] = np.cumsum([len(r[0]) for r in rows])
    idx = np.concatenate([r[0] for r in rows] + [np.zeros(0, np.int32)]).astype(np.int32)
    tim = np.concatenate([r[1] for r in rows] + [np.zeros(0, F32)]).astype(F32)
    return off, idx, tim


def _packed_gpu(off, idx, tim, N, K):
    from bats_amd import ops
    dev = torch.device("cuda")
    if idx.size == 0:  # a zero-length device tensor has no data pointer to pass
        idx, tim = np.full(1, -1, np.int32), np.full(1, np.inf, F32)
    t, n, m = ops.events_csr_to_per_neuron(torch.as_tensor(off, device=dev),
                                           torch.as_tensor(idx, device=dev),
                                           torch.as_tensor(tim, device=dev), N, K)
    torch.cuda.synchronize()
    return t.cpu().numpy(), n.cpu().numpy(), m.cpu().numpy()


def _random_rows(rng, B, N, p_max, grid=None, stop_frac=0.0):
    rows = []
    for b in range(B):
        P = int(rng.integers(0, p_max + 1)) if b % 7 else 0  # some empty rows
        ii = rng.integers(0, N, size=P).astype(np.int32)
        tt = rng.uniform(0.0, 1.0, size=P)
        if grid:
            tt = np.round(tt / grid) * grid  # ties, also between events of one neuron
        order = np.argsort(tt, kind="stable")
        ii, tt = ii[order], tt[order].astype(F32)
        if P and rng.uniform() < stop_frac:
            ii[int(rng.integers(0, P))] = -1  # a negative index ends the list there
        rows.append((ii, tt))
    return rows


def _dense_from_rows(rows, P):
    B = len(rows)
    idx = np.full((B, P), -1, dtype=np.int32)
    tim = np.full((B, P), np.inf, dtype=F32)
    for b, (ii, tt) in enumerate(rows):
        idx[b, :len(ii)] = ii
        tim[b, :len(tt)] = tt
    return idx, tim


def test_reference_hand_arrays():
    # arrays of tests/CudaKernels/Inference/convert_to_spike_per_neuron_tests.py:11-32
    idx = np.array([[0, 0, 2, 2, 0, 1], [2, 2, 0, 1, -1, -1], [1, -1, -1, -1, -1, -1]])
    tim = np.array([[0.1, 0.2, 0.3, 0.4, 0.5, 0.6], [0.25, 0.3, 0.45, 0.57, np.inf, np.inf],
                    [0.33, np.inf, np.inf, np.inf, np.inf, np.inf]], dtype=F32)
    true_t = np.array([[[0.1, 0.2, 0.5], [0.6, np.inf, np.inf], [0.3, 0.4, np.inf]],
                       [[0.45, np.inf, np.inf], [0.57, np.inf, np.inf], [0.25, 0.3, np.inf]],
                       [[np.inf] * 3, [0.33, np.inf, np.inf], [np.inf] * 3]], dtype=F32)
    true_n = np.array([[3, 1, 2], [1, 1, 2], [0, 1, 0]])
    t, n = _dense_gpu(idx, tim, 3, 3)
    assert np.array_equal(n, true_n)
    assert np.array_equal(t, true_t)
    rows = [(r[r >= 0].astype(np.int32), tr[r >= 0]) for r, tr in zip(idx, tim)]
    t2, n2, m2 = _packed_gpu(*_packed(rows), 3, 3)
    assert np.array_equal(t2, true_t) and np.array_equal(n2, true_n)
    assert np.array_equal(m2, true_n.max(axis=1))


@pytest.mark.parametrize("K_mode", ["exact", "drop"])
def test_random_lists_match_oracle(K_mode):
    rng = np.random.default_rng(11)
    N = 784
    rows = _random_rows(rng, 37, N, 3000, grid=1e-3, stop_frac=0.3)
    P = max(len(r[0]) for r in rows)
    idx, tim = _dense_from_rows(rows, P)
    _, n_all = oev.events_to_per_neuron(idx, tim, N)
    K = int(n_all.max()) if K_mode == "exact" else max(1, int(n_all.max()) // 2)
    print(f"[events] B=37 P={P} N={N} K={K} ({K_mode})", flush=True)
    ref_t, ref_n = oev.events_to_per_neuron(idx, tim, N, max_per_neuron=K)
    t, n = _dense_gpu(idx, tim, N, K)
    assert np.array_equal(n, ref_n)
    assert np.array_equal(t, ref_t)
    t2, n2, m2 = _packed_gpu(*_packed(rows), N, K)
    assert np.array_equal(n2, ref_n)
    assert np.array_equal(t2, ref_t)
    assert np.array_equal(m2, ref_n.max(axis=1))


def test_packed_all_empty_and_wide_layer():
    # every row empty
    rows = [(np.zeros(0, np.int32), np.zeros(0, F32)) for _ in range(5)]
    t, n, m = _packed_gpu(*_packed(rows), 10, 4)
    assert np.all(np.isinf(t)) and not n.any() and not m.any()
    # a layer wider than the LDS counters takes the sequential form of the dense entry point
    rng = np.random.default_rng(5)
    N = 6000
    rows = _random_rows(rng, 6, N, 4000, grid=1e-3)
    P = max(len(r[0]) for r in rows)
    idx, tim = _dense_from_rows(rows, P)
    ref_t, ref_n = oev.events_to_per_neuron(idx, tim, N, max_per_neuron=3)
    t, n = _dense_gpu(idx, tim, N, 3)
    assert np.array_equal(n, ref_n) and np.array_equal(t, ref_t)


def test_long_runs_of_one_neuron():
    # whole 64-event chunks on one neuron, the K boundary falling inside a chunk
    N, K = 50, 150
    rows = []
    for b, run in enumerate((200, 64, 130, 1)):
        ii = np.concatenate([np.full(run, 5), np.arange(N), np.full(70, 7)]).astype(np.int32)
        tt = np.sort(np.random.default_rng(b).uniform(0, 1, size=ii.size)).astype(F32)
        rows.append((ii, tt))
    P = max(len(r[0]) for r in rows)
    idx, tim = _dense_from_rows(rows, P)
    ref_t, ref_n = oev.events_to_per_neuron(idx, tim, N, max_per_neuron=K)
    t, n = _dense_gpu(idx, tim, N, K)
    assert np.array_equal(n, ref_n) and np.array_equal(t, ref_t)
    t2, n2, m2 = _packed_gpu(*_packed(rows), N, K)
    assert np.array_equal(n2, ref_n) and np.array_equal(t2, ref_t)
    assert np.array_equal(m2, ref_n.max(axis=1))
