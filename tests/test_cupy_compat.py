"""The CuPy-name adapter (bats_amd.cupy_compat) on the CPU: the calls the reference's scripts
make (`.get()`, `cp.mean(g, axis=0)`, `cp.random.seed` / `uniform`) and the array helpers,
against NumPy's results on the same values."""
import numpy as np
import torch

import bats_amd.cupy_compat as cp


def test_get_copies_to_numpy():
    t = torch.arange(6, dtype=torch.float32).reshape(2, 3)
    g = t.get()
    assert isinstance(g, np.ndarray) and g.dtype == np.float32
    assert np.array_equal(g, np.arange(6, dtype=np.float32).reshape(2, 3))
    g[0, 0] = 99.0  # a copy, as cupy.ndarray.get returns
    assert float(t[0, 0]) == 0.0
    out = np.empty((2, 3), np.float32)
    assert t.get(out=out) is out and np.array_equal(out, t.numpy())


def test_mean_and_reductions_match_numpy():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(4, 5, 3)).astype(np.float32)
    t = cp.asarray(a)
    assert np.allclose(cp.mean(t, axis=0).get(), a.mean(axis=0), rtol=1e-6)
    assert np.allclose(float(cp.mean(t)), a.mean(), rtol=1e-6)
    assert cp.mean(t, axis=0).dtype == torch.float32
    ints = cp.asarray(np.array([[1, 2], [3, 5]], np.int32))
    assert cp.mean(ints, axis=1).dtype == torch.float64
    assert np.allclose(cp.mean(ints, axis=1).get(), [1.5, 4.0])
    assert np.allclose(cp.sum(t, axis=(0, 2)).get(), a.sum(axis=(0, 2)), rtol=1e-5)
    assert np.allclose(cp.max(t, axis=1).get(), a.max(axis=1))
    assert np.array_equal(cp.argmax(t, axis=2).get(), a.argmax(axis=2))
    assert np.allclose(cp.mean(t, axis=0, keepdims=True).get(), a.mean(axis=0, keepdims=True))


def test_random_seed_and_uniform():
    cp.random.seed(7)
    a = cp.random.uniform(-1.0, 1.0, size=(50, 40), dtype=cp.float32)
    cp.random.seed(7)
    b = cp.random.uniform(-1.0, 1.0, size=(50, 40), dtype=cp.float32)
    assert a.dtype == torch.float32 and tuple(a.shape) == (50, 40)
    assert torch.equal(a, b)
    assert float(a.min()) >= -1.0 and float(a.max()) < 1.0
    assert abs(float(a.mean())) < 0.1
    n = cp.random.normal(0.0, 2.0, size=1000)
    assert n.dtype == torch.float64 and 1.7 < float(n.std()) < 2.3


def test_array_helpers():
    z = cp.zeros((2, 3), dtype=cp.int32)
    assert z.dtype == torch.int32 and not z.any()
    f = cp.full(3, np.inf, dtype=cp.float32)
    assert bool(cp.isinf(f).all())
    w = cp.where(cp.asarray(np.array([True, False])), 1.0, cp.asarray(np.array([5.0, 6.0])))
    assert np.array_equal(cp.asnumpy(w), [1.0, 6.0])
    assert np.array_equal(cp.asnumpy(cp.arange(4)), np.arange(4))
    assert np.array_equal(cp.asnumpy(np.array([1, 2])), [1, 2])
