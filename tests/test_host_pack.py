"""The serving host helper bats_pack_event_lists (no GPU work): gathering event lists out of
an arena gives the concatenation of the lists with their cumulative offsets, for batches that
split over several host threads, empty lists and one list; a batch beyond the buffer fails."""
import ctypes

import numpy as np

from bats_amd import _lib


def _pack(arena_idx, arena_t, starts, lens, capacity):
    lib = _lib.load()
    n = len(starts)
    starts = np.ascontiguousarray(starts, dtype=np.int64)
    lens = np.ascontiguousarray(lens, dtype=np.int32)
    off = np.zeros(n + 1, np.int32)
    out_i = np.full(capacity, -7, np.int32)
    out_t = np.full(capacity, -7.0, np.float32)
    rc = lib.bats_pack_event_lists(n, starts.ctypes.data, lens.ctypes.data,
                                   arena_idx.ctypes.data, arena_t.ctypes.data, off.ctypes.data,
                                   out_i.ctypes.data, out_t.ctypes.data, capacity)
    return rc, off, out_i, out_t


def test_gather_equals_concatenation():
    rng = np.random.default_rng(0)
    arena_i = rng.integers(0, 784, size=3_000_000).astype(np.int32)
    arena_t = rng.uniform(size=3_000_000).astype(np.float32)
    for n, hi in ((1, 2000), (7, 3), (900, 2000), (2000, 2000)):
        lens = rng.integers(0, hi + 1, size=n)
        starts = rng.integers(0, len(arena_i) - hi, size=n)
        cap = int(lens.sum()) + 5
        rc, off, out_i, out_t = _pack(arena_i, arena_t, starts, lens, cap)
        assert rc == 0
        want_i = np.concatenate([arena_i[s:s + l] for s, l in zip(starts, lens)])
        want_t = np.concatenate([arena_t[s:s + l] for s, l in zip(starts, lens)])
        assert np.array_equal(off, np.concatenate([[0], np.cumsum(lens)]))
        assert np.array_equal(out_i[:len(want_i)], want_i)
        assert np.array_equal(out_t[:len(want_t)], want_t)
        assert np.all(out_i[len(want_i):] == -7)  # nothing written past the batch


def test_empty_batch_and_capacity():
    a_i = np.arange(100, dtype=np.int32)
    a_t = np.arange(100, dtype=np.float32)
    rc, off, _, _ = _pack(a_i, a_t, [], [], 4)
    assert rc == 0 and off[0] == 0
    rc, _, _, _ = _pack(a_i, a_t, [0, 10], [30, 30], 40)
    assert rc != 0
    assert b"larger than its buffer" in _lib.load().bats_last_error()
