"""The dense forward from packed event lists (bats_lif_forward_packed, the inference server's
first layer) against the per-neuron path (event lists -> per-neuron buffers -> per-sample sort
-> solver) on the same requests, and against the FP64 oracle.

With distinct event times both paths give the solver the same event order, so spike counts,
times, a, x and z are bit-identical (+inf tails included), in both solver forms (one wave per
(sample, neuron) below 32768 pairs, one lane per pair above). Edge cases: empty lists, a list
ended by a negative index, by an index outside [0, I) and by a non-finite time, lists longer
than P_max (the rest dropped), and 200 events on one input neuron. Times tied across inputs
keep their list order in the packed path: where that is the per-neuron order the results are
again bit-identical; otherwise counts are identical and times within 1e-5 relative.
"""
import numpy as np
import pytest
import torch

from oracle import events as oev
from oracle import lif as olif
from tests.helpers import assert_forward_match, spike_slots

pytestmark = pytest.mark.gpu
F32 = np.float32
I, J, K, TAU_S, THETA, T_MAX = 784, 800, 30, 0.130, 1.0, 1.0


def _requests(rng, B, p_hi, grid=None):
    lists = []
    for _ in range(B):
        P = int(rng.integers(0, p_hi + 1))
        idx = rng.integers(0, I, size=P).astype(np.int32)
        tm = rng.uniform(0.0, 1.0, size=P)
        if grid:
            tm = np.round(tm * grid) / grid
        tm = tm.astype(F32)
        order = np.lexsort((idx, tm)) if grid else np.argsort(tm, kind="stable")
        lists.append((idx[order], tm[order]))
    return lists


def _packed(lists):
    off = np.zeros(len(lists) + 1, np.int32)
    off[1:] = np.cumsum([len(i) for i, _ in lists])
    idx = np.concatenate([i for i, _ in lists] + [np.zeros(0, np.int32)])
    tm = np.concatenate([t for _, t in lists] + [np.zeros(0, F32)])
    return off, idx, tm


def _padded(lists):
    P = max(1, max(len(i) for i, _ in lists))
    idx = np.full((len(lists), P), -1, np.int32)
    tm = np.full((len(lists), P), np.inf, F32)
    for k, (i, t) in enumerate(lists):
        idx[k, :len(i)] = i
        tm[k, :len(t)] = t
    return idx, tm


def _both(lists, W, P_max=None, cut=None):
    from bats_amd import ops
    from bats_amd._constants import lif_constants
    consts = lif_constants(TAU_S, THETA, THETA)
    Wd = torch.as_tensor(W, device="cuda")
    off, idx, tm = _packed(lists)
    P_max = max(1, max(len(i) for i, _ in lists)) if P_max is None else P_max
    got = ops.lif_forward_packed(off, idx, tm, P_max, Wd, *consts, T_MAX, K)
    # the per-neuron path on the lists as the packed path reads them
    eff = lists if cut is None else cut
    pidx, ptm = _padded(eff)
    mult = max(1, max((np.bincount(i, minlength=I).max() if len(i) else 0) for i, _ in eff))
    per_t, _ = ops.events_to_per_neuron(pidx, ptm, I, int(mult))
    ref = ops.lif_forward(per_t, Wd, *consts, T_MAX, K)
    torch.cuda.synchronize()
    return [g.cpu().numpy() for g in got], [r.cpu().numpy() for r in ref[:5]]


def _bit_equal(got, ref, what):
    n_g, n_r = got[0], ref[0]
    assert np.array_equal(n_g, n_r), f"{what}: counts differ"
    m = spike_slots(n_r, K)
    for name, g, r in zip("taxz", got[1:], ref[1:]):
        assert np.array_equal(g[m], r[m]), f"{what}: {name} differs"
    assert np.all(np.isinf(got[1][~m])), f"{what}: +inf tails"


@pytest.mark.parametrize("B", [2, 64])
def test_packed_equals_per_neuron_path(B):
    rng = np.random.default_rng(B)
    W = (0.5 * rng.normal(0, 1.0, size=(J, I))).astype(F32)
    lists = _requests(rng, B, 1500)
    lists[0] = (np.zeros(0, np.int32), np.zeros(0, F32))  # an empty list
    got, ref = _both(lists, W)
    print(f"[packed] B={B}: spikes {int(got[0].sum())}", flush=True)
    assert got[0].sum() > 0
    _bit_equal(got, ref, f"B={B}")
    # the oracle on two samples
    for b in (1, B - 1):
        per, _ = oev.events_to_per_neuron(lists[b][0][None, :], lists[b][1][None, :], I)
        o = olif.lif_forward(per, W, TAU_S, THETA, THETA, K, tmax=T_MAX)
        n_g, t_g = got[0][b:b + 1], got[1][b:b + 1]
        assert np.array_equal(n_g, o.n)
        m = spike_slots(o.n, K)
        assert np.all(np.abs(t_g[m] - o.t[m]) <= 1e-5 * np.abs(o.t[m]))


def test_packed_list_ends_and_limits():
    rng = np.random.default_rng(7)
    W = (0.5 * rng.normal(0, 1.0, size=(J, I))).astype(F32)
    lists = _requests(rng, 6, 900)
    # 200 events on one input neuron
    t200 = np.sort(rng.uniform(0, 1, size=200)).astype(F32)
    lists[1] = (np.full(200, 3, np.int32), t200)
    cut = list(lists)
    i2, t2 = lists[2]
    i2 = i2.copy(); i2[len(i2) // 2] = -1             # a negative index ends the list
    lists[2] = (i2, t2)
    cut[2] = (i2[:len(i2) // 2], t2[:len(t2) // 2])
    i3, t3 = lists[3]
    i3 = i3.copy(); i3[len(i3) // 3] = I + 5          # so does an index past I
    lists[3] = (i3, t3)
    cut[3] = (i3[:len(i3) // 3], t3[:len(t3) // 3])
    i4, t4 = lists[4]
    t4 = t4.copy(); t4[len(t4) // 4:] = np.inf        # and a non-finite time
    lists[4] = (i4, t4)
    cut[4] = (i4[:len(i4) // 4], t4[:len(t4) // 4])
    P_max = 700                                        # longer lists lose their tail
    cut = [(i[:P_max], t[:P_max]) for i, t in cut]
    got, ref = _both(lists, W, P_max=P_max, cut=cut)
    _bit_equal(got, ref, "list ends and P_max")


def test_packed_tied_times():
    rng = np.random.default_rng(11)
    W = (0.5 * rng.normal(0, 1.0, size=(J, I))).astype(F32)
    # ties ordered by input neuron (the per-neuron path's order): bit-identical
    lists = _requests(rng, 8, 1200, grid=64)
    assert sum(len(t) - len(np.unique(t)) for _, t in lists) > 100
    got, ref = _both(lists, W)
    _bit_equal(got, ref, "ties in neuron order")
    # ties in another order: same counts, times within 1e-5
    shuffled = []
    for i, t in lists:
        order = np.lexsort((-i, t))
        shuffled.append((i[order], t[order]))
    got2, ref2 = _both(shuffled, W)
    assert np.array_equal(got2[0], ref2[0])
    m = spike_slots(ref2[0], K)
    assert np.all(np.abs(got2[1][m] - ref2[1][m]) <= 1e-5 * np.abs(ref2[1][m]))
