"""CPU pins of the FP64 oracle: every function is checked against something that does not
share its code -- the brute-force membrane integrator, closed forms derived in a different
variable, the hand-written arrays of the reference's tests, brute-force loops, and float64
central finite differences of the forward.
"""
import numpy as np
import pytest

from oracle import events as oev
from oracle import integrator
from oracle import lif as olif
from oracle import losses as oloss
from oracle import optim as oopt

F32 = np.float32
F64 = np.float64


# ------------------------------------------------------------------ forward solver
def _per_neuron_from_events(idx, times, n):
    ii = np.where(np.isfinite(times), idx, -1).astype(np.int32)
    per, _ = oev.events_to_per_neuron(ii, times.astype(F32), n)
    return per


def _integrator_check(per, W, tau_s, theta, dth, K, tmax, tol=1e-9):
    st = olif.lif_forward(per, W, tau_s, theta, dth, K, tmax=tmax, store=F64)
    ts32, tau32, c32, dtt32 = olif.lif_constants(tau_s, theta, dth)
    B, I, Kp = per.shape
    J = W.shape[0]
    checked = 0
    for b in range(B):
        ii, mm = np.nonzero(np.isfinite(per[b]))
        tin = per[b][ii, mm].astype(F64)
        for j in range(J):
            ref = integrator.simulate_neuron(tin, W[j, ii].astype(F64), F64(ts32), F64(tau32),
                                             F64(c32) * F64(tau32), F64(dtt32) * F64(tau32),
                                             F64(F32(tmax)), K)
            got = st.t[b, j, :st.n[b, j]]
            assert len(ref) == st.n[b, j], (b, j, len(ref), st.n[b, j])
            if len(ref):
                assert np.max(np.abs(ref - got)) < tol
            checked += len(ref)
    return checked


def test_solver_matches_integrator_on_reference_seed42_inputs():
    # inputs of tests/CudaKernels/Inference/compute_spike_times_tests.py:44-55
    np.random.seed(42)
    idx = np.random.randint(10, size=(1, 100))
    tt = np.sort(np.random.uniform(0.0, 0.5, size=(1, 100)), axis=1)
    tt[0, -28:] = np.inf
    W = np.random.normal(0, 1.0, size=(5, 10)).astype(F32)
    per = _per_neuron_from_events(idx, tt, 10)
    assert _integrator_check(per, W, 0.1, 0.1, 0.2, 128, 0.5) > 20


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_solver_matches_integrator_random(seed):
    rng = np.random.default_rng(seed)
    B, I, J, K = 2, 12, 6, 4
    n = rng.integers(0, K + 1, size=(B, I))
    t = np.sort(rng.uniform(0, 0.8, size=(B, I, K)), axis=2).astype(F32)
    t[np.arange(K)[None, None, :] >= n[:, :, None]] = np.inf
    W = rng.normal(1.0, 2.0, size=(J, I)).astype(F32)
    # small buffers exercise the buffer-full stop, t_max the time limit
    assert _integrator_check(t, W, 0.05, 0.3, 0.3, 3, 0.7) > 0


def test_single_input_closed_form():
    # One input of weight w at t0: u = w tau (y - y^2), y = exp(-(t - t0)/tau) (tau = 2 tau_s),
    # so u = theta at the larger root y = (1 + sqrt(1 - 4 theta/(w tau)))/2, and there is no
    # spike when w tau / 4 < theta.
    tau_s = 0.1
    ts32, tau32, c32, dtt32 = olif.lif_constants(tau_s, 0.02, 0.02)
    tau = F64(tau32)
    theta = F64(c32) * tau
    t0 = F32(0.125)
    for w in (0.5, 1.0, 3.0, 10.0):
        per = np.full((1, 1, 1), t0, dtype=F32)
        st = olif.lif_forward(per, np.array([[w]], dtype=F32), tau_s, 0.02, 0.02, 1,
                              tmax=np.inf, store=F64)
        disc = 1.0 - 4.0 * theta / (w * tau)
        if disc < 0:
            assert st.n[0, 0] == 0
        else:
            y = (1.0 + np.sqrt(disc)) / 2.0
            assert st.n[0, 0] == 1
            assert abs(st.t[0, 0, 0] - (F64(t0) - tau * np.log(y))) < 1e-12


def test_single_input_multi_spike_sequence():
    # After k spikes at y_1..y_k the reset adds -delta_theta * sum_l y / y_l to u, so the
    # next spike is the larger root of w tau y^2 - (w tau - D_k) y + theta = 0,
    # D_k = delta_theta * sum_l 1/y_l.
    tau_s = 0.1
    theta_in, dth_in = 0.05, 0.02
    ts32, tau32, c32, dtt32 = olif.lif_constants(tau_s, theta_in, dth_in)
    tau = F64(tau32)
    theta = F64(c32) * tau
    dth = F64(dtt32) * tau
    w = 2.0
    st = olif.lif_forward(np.zeros((1, 1, 1), dtype=F32), np.array([[w]], dtype=F32), tau_s,
                          theta_in, dth_in, 30, tmax=np.inf, store=F64)
    expect = []
    D = 0.0
    for _ in range(30):
        wt = w * tau
        disc = (wt - D) ** 2 - 4.0 * wt * theta
        if disc < 0:
            break
        y = ((wt - D) + np.sqrt(disc)) / (2.0 * wt)
        if expect and -tau * np.log(y) <= expect[-1]:
            break
        expect.append(-tau * np.log(y))
        D += dth / y
    assert len(expect) >= 3
    assert st.n[0, 0] == len(expect)
    np.testing.assert_allclose(st.t[0, 0, :len(expect)], expect, rtol=0, atol=1e-11)


def test_sort_invariants():
    rng = np.random.default_rng(3)
    t, n = _multi(rng, 4, 30, 5)
    st, neu, slot, counts = olif.sort_spikes(t)
    flat = t.reshape(4, -1)
    for b in range(4):
        c = counts[b]
        assert c == np.isfinite(flat[b]).sum()
        assert np.all(np.diff(st[b, :c]) >= 0)
        assert sorted(slot[b, :c].tolist()) == sorted(np.nonzero(np.isfinite(flat[b]))[0].tolist())
        assert np.array_equal(neu[b, :c], slot[b, :c] // 5)
        assert np.array_equal(st[b, :c], flat[b, slot[b, :c]])
        assert np.all(np.isinf(st[b, c:]))


def _multi(rng, B, I, K, tmax=1.0):
    n = rng.integers(0, K + 1, size=(B, I))
    t = np.sort(rng.uniform(0, tmax, size=(B, I, K)), axis=2).astype(F32)
    t[np.arange(K)[None, None, :] >= n[:, :, None]] = np.inf
    return t, n.astype(np.int32)


def test_pre_exps_closed_form():
    rng = np.random.default_rng(4)
    t, _ = _multi(rng, 3, 7, 4)
    es, e = olif.pre_exps(t, F32(0.13), F32(0.26))
    fin = np.isfinite(t)
    np.testing.assert_allclose(es[fin], np.exp(t[fin].astype(F64) / F64(F32(0.13))), rtol=1e-7)
    np.testing.assert_allclose(e[fin], np.exp(t[fin].astype(F64) / F64(F32(0.26))), rtol=1e-7)
    assert np.all(np.isinf(es[~fin])) and np.all(np.isinf(e[~fin]))


# ------------------------------------------------------------------ backward pieces
def test_recurrent_errors_reference_hand_arrays():
    # hand-written arrays and O(N^2) formula of
    # tests/CudaKernels/Backpropagation/propagate_recurrent_errors_tests.py:9-41
    dth, tau = 1.0, 0.2
    dtt = dth / tau
    spikes = np.array([[[0.1, 0.2, 0.3, 0.4], [0.3, 0.35, 0.4, np.inf], [np.inf] * 4,
                        [0.4, 1.0, np.inf, np.inf]],
                       [[0.2, 0.3, 0.35, 0.5], [0.2, 0.4, 0.7, np.inf], [np.inf] * 4,
                        [0.7, np.inf, np.inf, np.inf]]])
    x = spikes
    z = np.exp(spikes / tau)
    errors = np.array([[[1.0, 2.0, -3.0, 4.0], [3.0, 3.5, -4.0, 0.0], [0.0] * 4, [0.4, -1.0, 0.0, 0.0]],
                       [[-0.2, -0.3, -0.35, 0.5], [-1.2, 4.0, -0.7, 0.0], [0.0] * 4, [4.2, 0.0, 0.0, 0.0]]])
    ref = errors.copy()
    for sx, sz, se in zip(x, z, ref):
        for nx, nz, ne in zip(sx, sz, se):
            for i in reversed(range(len(ne))):
                if ne[i] == 0.0:
                    continue
                for j in range(i + 1, len(ne)):
                    if ne[j] == 0.0:
                        break
                    ne[i] += ne[j] * dtt / nx[j] * nz[i]
    got = olif.recurrent_errors(x.astype(F32), z.astype(F32), errors.astype(F32), F32(dtt))
    np.testing.assert_allclose(got, ref, rtol=1e-5)


def _root_t(a, b, c, tau):
    return tau * np.log(2 * a / (b + np.sqrt(b * b - 4 * a * c)))


def test_factors_are_root_derivatives():
    rng = np.random.default_rng(5)
    c, tau = 3.0, 0.26
    a = rng.uniform(0.5, 2.0, 50)
    b = 2.0 * np.sqrt(a * c) * rng.uniform(1.05, 3.0, 50)  # positive discriminant
    x = np.sqrt(b * b - 4 * a * c)
    z = 2 * a / (b + x)
    t = tau * np.log(z)
    f1, f2 = olif.factors(t, a, x, z, c, tau, store=F64)
    h = 1e-6
    da = (_root_t(a + h * a, b, c, tau) - _root_t(a - h * a, b, c, tau)) / (2 * h * a)
    db = (_root_t(a, b + h * b, c, tau) - _root_t(a, b - h * b, c, tau)) / (2 * h * b)
    np.testing.assert_allclose(f1, da, rtol=1e-6)
    np.testing.assert_allclose(f2, -db, rtol=1e-6)
    # closed form of tests/CudaKernels/Backpropagation/compute_factors_tests.py:9-12
    np.testing.assert_allclose(f1, tau / a * (1 + c / x * z), rtol=1e-12)


def _brute_weights_gradient(f1, f2, post_t, pre_t, es, e, err):
    # per-(sample, post spike, pre neuron) loops, as
    # tests/CudaKernels/Backpropagation/compute_weights_gradient_tests.py:9-24
    B, J, K = f1.shape
    _, I, Kp = pre_t.shape
    out = np.zeros((B, J, I))
    for b in range(B):
        for j in range(J):
            for k in range(K):
                if np.isinf(post_t[b, j, k]):
                    break
                for i in range(I):
                    m = pre_t[b, i] < post_t[b, j, k]
                    out[b, j, i] += err[b, j, k] * (f1[b, j, k] * es[b, i][m].astype(F64).sum()
                                                    - f2[b, j, k] * e[b, i][m].astype(F64).sum())
    return out


def _brute_pre_errors(f1, f2, post_t, pre_t, es, e, W, err, tau_s, tau):
    # loops of tests/CudaKernels/Backpropagation/propagate_errors_to_pre_spikes_tests.py:9-31
    B, J, K = f1.shape
    _, I, Kp = pre_t.shape
    out = np.zeros((B, I, Kp))
    for b in range(B):
        for i in range(I):
            for m in range(Kp):
                tp = pre_t[b, i, m]
                if np.isinf(tp):
                    break
                for j in range(J):
                    for k in range(K):
                        tq = post_t[b, j, k]
                        if np.isinf(tq):
                            break
                        if tq <= tp:
                            continue
                        out[b, i, m] += err[b, j, k] * W[j, i] * (
                            f1[b, j, k] * es[b, i, m] / tau_s - f2[b, j, k] * e[b, i, m] / tau)
    return out


def test_weights_gradient_and_pre_errors_brute_force():
    rng = np.random.default_rng(6)
    B, I, J, K, Kp = 3, 9, 5, 6, 4
    tau_s, tau = 0.1, 0.2
    pre_t, _ = _multi(rng, B, I, Kp)
    post_t, _ = _multi(rng, B, J, K)
    es, e = olif.pre_exps(pre_t, tau_s, tau, store=F64)
    f1 = rng.uniform(0, 1, (B, J, K))
    f2 = rng.uniform(0, 1, (B, J, K))
    err = rng.uniform(-1, 1, (B, J, K))
    W = rng.normal(size=(J, I))
    got = olif.weights_gradient(f1, f2, post_t, pre_t, es, e, err, store=F64)
    ref = _brute_weights_gradient(f1, f2, post_t, pre_t, es, e, err)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12)
    got_p = olif.pre_spike_errors(f1, f2, post_t, pre_t, es, e, W, err, tau_s, tau, store=F64)
    ref_p = _brute_pre_errors(f1, f2, post_t, pre_t, es, e, W, err, tau_s, tau)
    np.testing.assert_allclose(got_p, ref_p, rtol=1e-10, atol=1e-12)


def _fd_setup(seed):
    rng = np.random.default_rng(seed)
    B, I, J, K = 2, 16, 4, 5
    t, _ = _multi(rng, B, I, 2, tmax=0.5)
    W = rng.normal(1.0, 1.0, size=(J, I))
    return t.astype(F64), W, rng


def _loss(t_in, W, e_out, params):
    st = olif.lif_forward(t_in, W, *params, store=F64)
    return np.nansum(np.where(np.isfinite(st.t), st.t * e_out, 0.0)), st


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_central_differences(seed):
    # L = sum_spikes e * t; dL/dW from the oracle's BATS backward (reset chain + factors +
    # weight gradient) against float64 central differences, h = 1e-6 (|w| + 1)
    t_in, W, rng = _fd_setup(seed)
    params = (0.1, 0.3, 0.15, 8, 0.6)  # tau_s, theta, delta_theta, K, t_max
    st0 = olif.lif_forward(t_in, W, *params[:4], tmax=params[4], store=F64)
    assert st0.n.sum() >= 4
    e_out = rng.normal(size=st0.t.shape)
    _, grad, pre_err = olif.lif_backward(st0, e_out.copy(), pre_trainable=True, store=F64)
    g = grad.sum(axis=0)
    fparams = (params[0], params[1], params[2], params[3])

    def L(Wp, tp):
        st = olif.lif_forward(tp, Wp, *fparams, tmax=params[4], store=F64)
        if not np.array_equal(st.n, st0.n):
            return None
        return np.where(np.isfinite(st.t), st.t * e_out, 0.0).sum()

    checked = 0
    for j in range(W.shape[0]):
        for i in range(W.shape[1]):
            h = 1e-6 * (abs(W[j, i]) + 1.0)
            Wp, Wm = W.copy(), W.copy()
            Wp[j, i] += h
            Wm[j, i] -= h
            lp, lm = L(Wp, t_in), L(Wm, t_in)
            if lp is None or lm is None:
                continue
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[j, i]) <= 1e-6 * max(1.0, abs(fd)), (j, i, fd, g[j, i])
            checked += 1
    assert checked > 20
    # pre-spike errors = dL/dt_pre
    checked = 0
    fin = np.argwhere(np.isfinite(t_in))
    for (b, i, m) in fin[:40]:
        h = 1e-6
        tp, tm = t_in.copy(), t_in.copy()
        tp[b, i, m] += h
        tm[b, i, m] -= h
        if (m + 1 < t_in.shape[2] and tp[b, i, m] >= t_in[b, i, m + 1]) or (m > 0 and tm[b, i, m] <= t_in[b, i, m - 1]):
            continue
        lp, lm = L(W, tp), L(W, tm)
        if lp is None or lm is None:
            continue
        fd = (lp - lm) / (2 * h)
        assert abs(fd - pre_err[b, i, m]) <= 1e-4 * max(1.0, abs(fd)), (b, i, m, fd, pre_err[b, i, m])
        checked += 1
    assert checked > 10


# ------------------------------------------------------------------ losses, optimizers, events
def test_ttfs_loss_reference_hand_arrays():
    # arrays of tests/Losses/TTFSSoftmaxCrossEntropy_tests.py:18-28 with their closed form
    tau = 0.2
    spk = np.array([[[0.1, 0.2, 0.3, np.inf], [0.2, np.inf, np.inf, np.inf], [np.inf] * 4,
                     [0.3, 0.4, 0.5, 0.6]],
                    [[0.4, 0.5, np.inf, np.inf], [1.0, np.inf, np.inf, np.inf],
                     [0.9, 1.1, np.inf, np.inf], [0.5, np.inf, np.inf, np.inf]]]).astype(F32)
    labels = np.array([3, 0])
    exps = np.exp(-spk[..., 0].astype(F64) / tau)
    soft = exps / exps.sum(axis=1, keepdims=True)
    true_loss = -np.log(soft[np.arange(2), labels])
    neg = -soft
    neg[np.arange(2), labels] += 1
    true_err = np.zeros(spk.shape)
    true_err[..., 0] = neg / tau
    loss, err = oloss.ttfs_loss_and_errors(spk, labels, tau)
    np.testing.assert_allclose(loss, true_loss, rtol=1e-6)
    np.testing.assert_allclose(err, true_err, rtol=1e-5, atol=1e-6)
    assert np.array_equal(oloss.ttfs_predict(spk), np.argmin(spk, axis=1).flatten())


def test_ttfs_loss_silent_outputs_give_zero_errors():
    spk = np.full((2, 3, 1), np.inf, dtype=F32)
    loss, err = oloss.ttfs_loss_and_errors(spk, np.array([0, 2]), 0.005)
    assert np.all(err == 0.0)
    assert np.all(np.isnan(loss))


def test_count_loss_closed_form():
    n = np.array([[3, 0, 15, 1], [0, 0, 0, 0]], dtype=np.int32)
    labels = np.array([2, 1])
    tg = oloss.count_targets(n, labels, 15, 3)
    assert np.array_equal(tg, [[3, 3, 15, 3], [3, 15, 3, 3]])
    loss, err = oloss.count_loss_and_errors(n, tg, 5)
    np.testing.assert_array_equal(loss, [0.5 * (0 + 9 + 0 + 4), 0.5 * (9 + 225 + 9 + 9)])
    assert err.shape == (2, 4, 5)
    np.testing.assert_array_equal(err[:, :, 3], tg - n)
    assert np.array_equal(oloss.count_predict(n), [2, 0])


def test_adam_closed_form_and_convergence():
    rng = np.random.default_rng(7)
    g = rng.normal(size=(5, 4)).astype(F32)
    opt = oopt.Adam(0.05)
    d = opt.step([None, g])
    assert d[0] is None
    # first step: m_hat = g, v_hat = g^2  ->  delta = -lr * g / (|g| + eps)
    np.testing.assert_allclose(d[1], -0.05 * g / (np.abs(g) + 1e-8), rtol=1e-5)
    # tests/Optimizers/AdamOptimizer.py:21-33: 1000 steps on 0.5*|w|^2 drive w to 0
    w = [None, rng.normal(size=(5, 10)).astype(F32), rng.normal(size=(2, 5)).astype(F32)]
    opt = oopt.Adam(0.05)
    for _ in range(1000):
        for k, dd in enumerate(opt.step(w)):
            if dd is not None:
                w[k] = (w[k] + dd).astype(F32)
    for ww in w[1:]:
        assert np.allclose(ww, 0.0, atol=1e-6)


def test_events_to_per_neuron_reference_hand_arrays():
    # arrays of tests/CudaKernels/Inference/convert_to_spike_per_neuron_tests.py:11-32
    idx = np.array([[0, 0, 2, 2, 0, 1], [2, 2, 0, 1, -1, -1], [1, -1, -1, -1, -1, -1]])
    tim = np.array([[0.1, 0.2, 0.3, 0.4, 0.5, 0.6], [0.25, 0.3, 0.45, 0.57, np.inf, np.inf],
                    [0.33, np.inf, np.inf, np.inf, np.inf, np.inf]])
    true_t = np.array([[[0.1, 0.2, 0.5], [0.6, np.inf, np.inf], [0.3, 0.4, np.inf]],
                       [[0.45, np.inf, np.inf], [0.57, np.inf, np.inf], [0.25, 0.3, np.inf]],
                       [[np.inf] * 3, [0.33, np.inf, np.inf], [np.inf] * 3]])
    true_n = np.array([[3, 1, 2], [1, 1, 2], [0, 1, 0]])
    t, n = oev.events_to_per_neuron(idx, tim, 3)
    assert np.array_equal(n, true_n)
    np.testing.assert_allclose(t, true_t)


def test_nadam_closed_form_and_convergence():
    """NAdam (bats/Optimizers/NAdamOptimizer.py:36-58). With a constant gradient g the bias
    corrections make m_hat = g and v_hat = g^2 at every step, so the update reduces to
    delta_t = -lr * g / (|g| + eps) * (beta_1 + (1 - beta_1) / (1 - beta_1^t)):
    t = 1 gives -lr (1 + beta_1) g / (|g| + eps), t = 2 gives the factor beta_1 + 1/(1 + beta_1)."""
    rng = np.random.default_rng(17)
    g = rng.normal(size=(6, 5)).astype(F32)
    lr, b1 = 0.05, 0.9
    opt = oopt.NAdam(lr)
    base = -lr * g.astype(F64) / (np.abs(g.astype(F64)) + 1e-8)
    d1 = opt.step([None, g])
    assert d1[0] is None
    np.testing.assert_allclose(d1[1], base * (1.0 + b1), rtol=1e-5)
    d2 = opt.step([None, g])
    np.testing.assert_allclose(d2[1], base * (b1 + 1.0 / (1.0 + b1)), rtol=1e-5)
    d3 = opt.step([None, g])
    np.testing.assert_allclose(d3[1], base * (b1 + (1 - b1) / (1 - b1 ** 3)), rtol=1e-5)
    # the convergence test of tests/Optimizers/AdamOptimizer.py:8-33 with this optimizer
    w = [None, rng.normal(size=(5, 10)).astype(F32), rng.normal(size=(2, 5)).astype(F32)]
    opt = oopt.NAdam(0.05)
    for _ in range(1000):
        for k, dd in enumerate(opt.step(w)):
            if dd is not None:
                w[k] = (w[k] + dd).astype(F32)
    for ww in w[1:]:
        assert np.allclose(ww, 0.0, atol=1e-6)


def test_gd_closed_form_and_convergence():
    """GradientDescentOptimizer (bats/Optimizers/GradientDescentOptimizer.py:8-9): delta =
    -lr * g, None passes through; tests/Optimizers/GradientDescentOptimizer.py:8-33 drives
    0.5*|w|^2 to zero (w_t = 0.95^t w_0 exactly in real arithmetic)."""
    rng = np.random.default_rng(18)
    g = rng.normal(size=(4, 3)).astype(F32)
    d = oopt.gd_step([None, g, None], 0.05)
    assert d[0] is None and d[2] is None
    assert d[1].dtype == F32
    np.testing.assert_array_equal(d[1], (F64(F32(0.05)) * -g.astype(F64)).astype(F32))
    w0 = [None, rng.normal(size=(5, 10)).astype(F32), rng.normal(size=(2, 5)).astype(F32)]
    w = [None if x is None else x.copy() for x in w0]
    for step in range(1000):
        for k, dd in enumerate(oopt.gd_step(w, 0.05)):
            if dd is not None:
                w[k] = (w[k] + dd).astype(F32)
        if step == 9:
            for a, b in zip(w[1:], w0[1:]):
                np.testing.assert_allclose(a, b * (1 - F64(F32(0.05))) ** 10, rtol=1e-5)
    for ww in w[1:]:
        assert np.allclose(ww, 0.0)


def test_spike_time_loss_hand_arrays():
    """SpikeTimeLoss (bats/Losses/SpikeTimeLoss.py:38-47) on hand arrays: an infinite target
    or first spike counts as 1.0 in the loss, while the errors subtract the sanitised targets
    from the raw times (an absent spike keeps an infinite error, SpikeTimeLoss.py:45)."""
    inf = np.inf
    t = np.array([[[0.1, inf], [inf, inf], [0.3, 0.5]],
                  [[0.7, 0.8], [0.2, inf], [inf, inf]]], dtype=F32)
    tg = np.array([[[0.2, inf], [0.4, inf], [inf, inf]],
                   [[0.5, 0.9], [inf, inf], [0.6, 0.1]]], dtype=F32)
    loss, err = oloss.spike_time_loss_and_errors(t, tg)
    want_loss = [0.5 * ((0.2 - 0.1) ** 2 + (0.4 - 1.0) ** 2 + (1.0 - 0.3) ** 2),
                 0.5 * ((0.5 - 0.7) ** 2 + (1.0 - 0.2) ** 2 + (0.6 - 1.0) ** 2)]
    np.testing.assert_allclose(loss, want_loss, rtol=1e-6)
    want_err = np.array([[[-0.1, inf], [inf, inf], [-0.7, -0.5]],
                         [[0.2, -0.1], [-0.8, inf], [inf, inf]]])
    fin = np.isfinite(want_err)
    assert np.array_equal(np.isinf(err), ~fin) and np.all(err[~fin] > 0)
    np.testing.assert_allclose(err[fin], want_err[fin], rtol=1e-6, atol=1e-7)
    # the caller's targets are not modified by the oracle (it works on a copy)
    assert np.isinf(tg[0, 0, 1])


def test_lif_layer_descends_to_target_times():
    """The reference's end-to-end layer test (tests/Layers/LIFLayer_tests.py:91-132) on the
    oracle: 7 inputs -> 5 LIF neurons, all weights 1, errors = first spike time - target,
    gradient descent at lr 3e3 on the batch-mean gradient for 200 epochs; the first spike
    times reach the targets within 5e-4. A wrong sign or a dropped term in the backward does
    not descend."""
    tau_s, theta, K, inf = 10e-3, 0.01, 8, np.inf
    t_in = np.array([[[0.0, inf], [3e-3, inf], [6e-3, inf], [8e-3, inf], [9e-3, inf],
                      [11e-3, inf], [17e-3, inf]]], dtype=F32)
    targets = np.array([[4e-3, 6e-3, 8e-3, 12e-3, 16e-3]], dtype=F32)
    W = np.full((5, 7), 1.0, dtype=F32)
    for _ in range(200):
        st = olif.lif_forward(t_in, W, tau_s, theta, theta, K, tmax=0.02)
        err = np.zeros(st.t.shape, dtype=F32)
        err[:, :, 0] = st.t[:, :, 0] - targets
        _, g, pre = olif.lif_backward(st, err, pre_trainable=False)
        assert pre is None
        W = (W.astype(F64) + (-F64(F32(3e3)) * olif.batch_mean(g).astype(F64))).astype(F32)
    st = olif.lif_forward(t_in, W, tau_s, theta, theta, K, tmax=0.02)
    assert np.allclose(st.t[:, :, 0], targets, atol=5e-4), st.t[:, :, 0]
