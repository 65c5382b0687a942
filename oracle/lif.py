"""FP64 restatement of the dense LIF forward (exact spike-time solver) and the
BATS backward, vectorised over (sample, neuron) lanes.  Test infrastructure only.

Every function cites the reference file:line it restates (paths relative to
the reference repository root).
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
F64 = np.float64


def lif_constants(tau_s, theta, delta_theta):
    """Per-layer float32 constants, as bats/Layers/LIFLayer.py:16-19.

    tau = f32(2*tau_s); c = f32(theta / tau) and delta_theta_tau = f32(delta_theta / tau),
    where the divisions are carried out in float64 on the float32 tau (the legacy NumPy
    scalar promotion of ``python float / numpy.float32``) and then rounded to float32.
    Returns (tau_s, tau, c, delta_theta_tau) as numpy.float32.
    """
    tau_s32 = F32(tau_s)
    tau32 = F32(2 * F64(tau_s))
    c32 = F32(F64(theta) / F64(tau32))
    dtt32 = F32(F64(delta_theta) / F64(tau32))
    return tau_s32, tau32, c32, dtt32


def pre_exps(times, tau_s, tau, store=F32):
    """exp(t/tau_s), exp(t/tau) of every spike slot (+inf stays +inf).

    Restates bats/CudaKernels/CuFiles/Inference/compute_pre_exps.cu:16-21.
    """
    t = np.asarray(times, dtype=F64)
    with np.errstate(over="ignore", invalid="ignore"):
        es = np.exp(t / F64(tau_s))
        e = np.exp(t / F64(tau))
    return es.astype(store), e.astype(store)


def sort_spikes(times):
    """Per-sample time order of all finite spike slots.

    Restates bats/CudaKernels/Wrappers/Inference/get_sorted_spikes_indices.py:8-16
    plus the gathers of bats/Layers/LIFLayer.py:80-84. The reference's argpartition
    leaves the order of equal times unspecified; here ties are broken by flat slot
    index (stable sort), which is also the order the HIP sort produces.

    times: [B, N, K] float32 (+inf padded).
    Returns (sorted_times [B,P] f32 +inf padded, sorted_neuron [B,P] int32 (0 in pads),
             sorted_slot [B,P] int64 (-1 in pads), counts [B]), P = max finite count.
    """
    times = np.asarray(times)
    B, N, K = times.shape
    flat = times.reshape(B, N * K)
    finite = np.isfinite(flat)
    counts = finite.sum(axis=1)
    P = int(counts.max()) if B > 0 else 0
    keyed = np.where(finite, flat, np.inf)
    order = np.argsort(keyed, axis=1, kind="stable")[:, :P]
    st = np.take_along_axis(keyed, order, axis=1).astype(F32)
    pad = np.arange(P)[None, :] >= counts[:, None]
    st[pad] = np.inf
    slot = np.where(pad, -1, order)
    neuron = np.where(pad, 0, order // K).astype(np.int32)
    return st, neuron, slot, counts


def lif_solve(st, sidx, ses, se, W, c, dtt, tau, tmax, max_n_post, store=F32):
    """Exact event-driven multi-spike LIF solver.

    Restates bats/CudaKernels/CuFiles/Inference/compute_spike_times.cu:34-138: for each
    (sample, post neuron), walk the time-sorted input events; after each event add
    w*exp(t/tau_s) to a and w*exp(t/tau) to b, then repeatedly take the smaller root
    z = 2a/(b + sqrt(b^2 - 4ac)) of c z^2 - b z + a = 0, t = tau*ln z, and accept it iff
    last < t <= min(next event, t_max) (:53-83), resetting b -= (delta_theta/tau) z.
    A full buffer (max_n_post spikes) ends that neuron's event loop (:80-82, :133-136).

    All locals are float64; the outputs are rounded to ``store``.
    st, ses, se: [B,P] sorted times / exps (+inf padded); sidx [B,P] int pre-neuron ids;
    W [J,I]. Returns n [B,J] int32, a, x, t, z [B,J,K] (t and z +inf padded, a and x NaN
    padded -- the reference leaves them uninitialised).
    """
    st = np.asarray(st)
    B, P = st.shape
    J = W.shape[0]
    K = int(max_n_post)
    W64 = np.asarray(W, dtype=F64)
    c = F64(c)
    dtt = F64(dtt)
    tau = F64(tau)
    tmax = F64(tmax)

    L = B * J
    a = np.zeros(L)
    b = np.zeros(L)
    n = np.zeros(L, dtype=np.int64)
    alive = np.ones(L, dtype=bool)
    out_a = np.full((L, K), np.nan, dtype=store)
    out_x = np.full((L, K), np.nan, dtype=store)
    out_t = np.full((L, K), np.inf, dtype=store)
    out_z = np.full((L, K), np.inf, dtype=store)

    st64 = st.astype(F64)
    ses64 = np.asarray(ses, dtype=F64)
    se64 = np.asarray(se, dtype=F64)
    with np.errstate(all="ignore"):
        for q in range(P):
            tq = st64[:, q]
            fin = np.isfinite(tq)
            if not fin.any():
                break
            lane_ok = alive & np.repeat(fin, J)
            if not lane_ok.any():
                continue
            w = W64[:, sidx[:, q]].T.reshape(L)  # [B,J] -> lanes
            es = np.repeat(ses64[:, q], J)
            e = np.repeat(se64[:, q], J)
            a = np.where(lane_ok, a + w * es, a)
            b = np.where(lane_ok, b + w * e, b)
            nxt = np.repeat(st64[:, q + 1] if q + 1 < P else np.full(B, np.inf), J)
            # pending lanes evaluated with index subsets
            lanes = np.nonzero(lane_ok)[0]
            last = np.repeat(tq, J)[lanes]
            while lanes.size:
                al = a[lanes]
                bl = b[lanes]
                disc = bl * bl - 4.0 * al * c
                keep = ~(disc < 0)
                x = np.sqrt(disc)
                den = x + bl
                keep &= ~(den == 0.0)
                z = 2.0 * al / den
                keep &= ~(z < 0)
                t = tau * np.log(z)
                keep &= ~((t <= last) | (t > tmax) | (t > nxt[lanes]))
                lanes_acc = lanes[keep]
                if lanes_acc.size == 0:
                    break
                k = n[lanes_acc]
                out_a[lanes_acc, k] = al[keep]
                out_x[lanes_acc, k] = x[keep]
                out_t[lanes_acc, k] = t[keep]
                out_z[lanes_acc, k] = z[keep]
                b[lanes_acc] = bl[keep] - dtt * z[keep]
                n[lanes_acc] = k + 1
                full = n[lanes_acc] >= K
                alive[lanes_acc[full]] = False
                cont = ~full
                lanes = lanes_acc[cont]
                last = t[keep][cont]
    shp = (B, J, K)
    return (n.reshape(B, J).astype(np.int32), out_a.reshape(shp), out_x.reshape(shp),
            out_t.reshape(shp), out_z.reshape(shp))


class LayerState:
    """What a forward pass keeps for the backward (bats/Layers/LIFLayer.py:26-35)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def lif_forward(pre_times, W, tau_s, theta, delta_theta, max_n_post, tmax=np.inf,
                store=F32):
    """One dense LIF layer forward, as bats/Layers/LIFLayer.py:63-91 (K1 + sort + gathers
    + K2 + K3, including the all-silent branch :71-78 which the solver reproduces by
    itself: no event, no spike).

    pre_times: [B,I,K_pre] (+inf padded). Returns a LayerState with n, a, x, t, z and the
    pre exps, plus the constants used.
    """
    tau_s32, tau32, c32, dtt32 = lif_constants(tau_s, theta, delta_theta)
    pre_times = np.asarray(pre_times, dtype=store)
    es, e = pre_exps(pre_times, tau_s32, tau32, store=store)
    st, sneu, sslot, counts = sort_spikes(pre_times)
    B = pre_times.shape[0]
    flat_es = es.reshape(B, -1)
    flat_e = e.reshape(B, -1)
    safe = np.where(sslot < 0, 0, sslot)
    ses = np.take_along_axis(flat_es, safe, axis=1)
    se = np.take_along_axis(flat_e, safe, axis=1)
    tmax32 = F32(tmax)
    n, a, x, t, z = lif_solve(st, sneu, ses, se, np.asarray(W, dtype=store), c32, dtt32,
                              tau32, tmax32, max_n_post, store=store)
    return LayerState(n=n, a=a, x=x, t=t, z=z, pre_times=pre_times, pre_exp_s=es,
                      pre_exp=e, tau_s=tau_s32, tau=tau32, c=c32, dtt=dtt32,
                      W=np.asarray(W, dtype=store))


def recurrent_errors(x, z, errors, dtt, store=F32):
    """Reverse O(N) error scan through the reset chain; returns the new errors.

    Restates bats/CudaKernels/CuFiles/Backpropagation/propagate_recurrent_errors.cu:17-22
    (which works in place on ``errors``): from the last slot to the first, skipping
    slots whose z is +inf, e_k <- e_k + C z_k; C <- C + e_k (delta_theta/tau) / x_k.
    e_k is an array element, so it is rounded to ``store`` before C reads it back.
    """
    e = np.array(errors, dtype=store, copy=True)
    B, J, K = e.shape
    C = np.zeros((B, J))
    z64 = np.asarray(z, dtype=F64)
    x64 = np.asarray(x, dtype=F64)
    dtt = F64(dtt)
    with np.errstate(all="ignore"):
        for k in range(K - 1, -1, -1):
            valid = ~np.isinf(z64[..., k])
            ek = e[..., k].astype(F64) + C * z64[..., k]
            ek_s = ek.astype(store)
            e[..., k] = np.where(valid, ek_s, e[..., k])
            C = np.where(valid, C + ek_s.astype(F64) * dtt / x64[..., k], C)
    return e


def factors(t, a, x, z, c, tau, store=F32):
    """f1 = tau/a*(1 + z*c/x) = dt/da and f2 = tau/x = -dt/db; 0 where t is +-inf.

    Restates bats/CudaKernels/CuFiles/Backpropagation/compute_factors.cu:20-30, in its
    operation order.
    """
    t64 = np.asarray(t, dtype=F64)
    a64 = np.asarray(a, dtype=F64)
    x64 = np.asarray(x, dtype=F64)
    z64 = np.asarray(z, dtype=F64)
    with np.errstate(all="ignore"):
        f1 = F64(tau) / a64 * (1.0 + z64 * F64(c) / x64)
        f2 = F64(tau) / x64
    inf = np.isinf(t64)
    f1 = np.where(inf, 0.0, f1).astype(store)
    f2 = np.where(inf, 0.0, f2).astype(store)
    return f1, f2


def _post_list(t_b, e_b, f1_b, f2_b):
    """Finite post spikes of one sample: (j, t, e*f1, e*f2) arrays."""
    jj, kk = np.nonzero(np.isfinite(t_b))
    tt = t_b[jj, kk].astype(F64)
    p1 = e_b[jj, kk].astype(F64) * f1_b[jj, kk].astype(F64)
    p2 = e_b[jj, kk].astype(F64) * f2_b[jj, kk].astype(F64)
    return jj, kk, tt, p1, p2


def _stop_at_first_inf(t):
    """Slots at or after the first +inf of each (sample, neuron) row become +inf, as the
    reference's ``break`` on the first infinite time (compute_weights_gradient.cu:37-38,
    propagate_errors_to_pre_spikes.cu:45-46)."""
    t = np.array(t, copy=True)
    bad = np.cumsum(np.isinf(t), axis=-1) > 0
    t[bad] = np.inf
    return t


def weights_gradient(f1, f2, post_t, pre_t, pre_es, pre_e, errors, store=F32):
    """Per-sample weight gradient [B,J,I].

    Restates bats/CudaKernels/CuFiles/Backpropagation/compute_weights_gradient.cu:33-45:
    grad[b,j,i] = sum_k e_k (f1_k SA_k - f2_k SB_k) with SA_k, SB_k the sums of
    exp(t/tau_s), exp(t/tau) over the pre spikes of neuron i strictly before t_k. The sum
    is regrouped per sample as dense float64 products (the order of a float64 sum is the
    only difference from the reference's two-pointer merge).
    """
    post_t = _stop_at_first_inf(post_t)
    B, J, K = post_t.shape
    _, I, Kp = pre_t.shape
    out = np.zeros((B, J, I), dtype=F64)
    for bb in range(B):
        jj, kk, tt, p1, p2 = _post_list(post_t[bb], errors[bb], f1[bb], f2[bb])
        if jj.size == 0:
            continue
        ii, mm = np.nonzero(np.isfinite(pre_t[bb]))
        pt = pre_t[bb][ii, mm].astype(F64)
        pes = pre_es[bb][ii, mm].astype(F64)
        pe = pre_e[bb][ii, mm].astype(F64)
        mask = (pt[None, :] < tt[:, None]).astype(F64)  # [Q,P]
        onehot = np.zeros((ii.size, I))
        onehot[np.arange(ii.size), ii] = 1.0
        SA = mask @ (onehot * pes[:, None])  # [Q,I]
        SB = mask @ (onehot * pe[:, None])
        rows = p1[:, None] * SA - p2[:, None] * SB
        np.add.at(out[bb], jj, rows)
    return out.astype(store)


def pre_spike_errors(f1, f2, post_t, pre_t, pre_es, pre_e, W, errors, tau_s, tau,
                     store=F32):
    """Errors of the pre-synaptic spikes [B,I,K_pre].

    Restates bats/CudaKernels/CuFiles/Backpropagation/propagate_errors_to_pre_spikes.cu:41-53
    and the reduction over post neurons in
    bats/CudaKernels/Wrappers/Backpropagation/propagate_errors_to_pre_spikes.py:26:
    pre_err[b,i,m] = sum_j W[j,i] * (C1 * es_m / tau_s - C2 * e_m / tau), with
    C1, C2 = sum of e*f1, e*f2 over the post spikes of j strictly after t_m. The per-(m,j)
    terms are summed in float64 and rounded once (the reference materialises them as a
    float32 [B,I,K_pre,J] array first).
    """
    post_t = _stop_at_first_inf(post_t)
    B, J, K = post_t.shape
    _, I, Kp = pre_t.shape
    out = np.zeros((B, I, Kp), dtype=F64)
    W64 = np.asarray(W, dtype=F64)
    tau_s = F64(tau_s)
    tau = F64(tau)
    for bb in range(B):
        jj, kk, tt, p1, p2 = _post_list(post_t[bb], errors[bb], f1[bb], f2[bb])
        ii, mm = np.nonzero(np.isfinite(pre_t[bb]))
        if jj.size == 0 or ii.size == 0:
            continue
        pt = pre_t[bb][ii, mm].astype(F64)
        pes = pre_es[bb][ii, mm].astype(F64)
        pe = pre_e[bb][ii, mm].astype(F64)
        mask = (tt[:, None] > pt[None, :]).astype(F64)  # [Q,P]
        onehot_j = np.zeros((J, jj.size))
        onehot_j[jj, np.arange(jj.size)] = 1.0
        C1 = onehot_j @ (mask * p1[:, None])  # [J,P]
        C2 = onehot_j @ (mask * p2[:, None])
        terms = W64[:, ii] * (C1 * pes[None, :] / tau_s - C2 * pe[None, :] / tau)
        out[bb, ii, mm] = terms.sum(axis=0)
    return out.astype(store)


def lif_backward(state, errors, pre_trainable=True, store=F32):
    """One dense LIF layer backward, as bats/Layers/LIFLayer.py:93-110.

    Returns (errors_after_reset_chain, per_sample_grad [B,J,I], pre_errors or None).
    """
    e = recurrent_errors(state.x, state.z, errors, state.dtt, store=store)
    f1, f2 = factors(state.t, state.a, state.x, state.z, state.c, state.tau, store=store)
    grad = weights_gradient(f1, f2, state.t, state.pre_times, state.pre_exp_s,
                            state.pre_exp, e, store=store)
    pre_err = None
    if pre_trainable:
        pre_err = pre_spike_errors(f1, f2, state.t, state.pre_times, state.pre_exp_s,
                                   state.pre_exp, state.W, e, state.tau_s, state.tau,
                                   store=store)
    return e, grad, pre_err


def batch_mean(grad, store=F32):
    """The caller's batch mean of per-sample gradients (experiments/mnist/train_spike_count.py:167)."""
    return np.asarray(grad, dtype=F64).mean(axis=0).astype(store)
