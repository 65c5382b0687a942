"""FP64 restatement of the reference optimizers.  Test infrastructure only."""
from __future__ import annotations

import numpy as np

F32 = np.float32
F64 = np.float64


class Adam:
    """bats/Optimizers/AdamOptimizer.py:22-55 with float32 arrays (m, v, m_hat, v_hat,
    delta) and float64 arithmetic; beta^t is evaluated in float64 on the float32 betas."""

    def __init__(self, learning_rate, beta_1=0.9, beta_2=0.999, epsilon=1e-8):
        self.lr = F32(learning_rate)
        self.b1 = F32(beta_1)
        self.omb1 = F32(1.0 - beta_1)
        self.b2 = F32(beta_2)
        self.omb2 = F32(1.0 - beta_2)
        self.eps = F32(epsilon)
        self.m = None
        self.v = None
        self.t = 0

    def step(self, grads):
        self.t += 1
        if self.m is None:
            self.m = [None if g is None else np.zeros(g.shape, F32) for g in grads]
            self.v = [None if g is None else np.zeros(g.shape, F32) for g in grads]
        s1 = 1.0 - F64(self.b1) ** self.t
        s2 = 1.0 - F64(self.b2) ** self.t
        out = []
        for i, g in enumerate(grads):
            if g is None:
                out.append(None)
                continue
            g64 = np.asarray(g, dtype=F64)
            m = (F64(self.b1) * self.m[i].astype(F64) + F64(self.omb1) * g64).astype(F32)
            v = (F64(self.b2) * self.v[i].astype(F64) + F64(self.omb2) * g64 * g64).astype(F32)
            self.m[i], self.v[i] = m, v
            mh = (m.astype(F64) / s1).astype(F32)
            vh = (v.astype(F64) / s2).astype(F32)
            d = -(F64(self.lr) * mh.astype(F64) / (np.sqrt(vh.astype(F64)) + F64(self.eps)))
            out.append(d.astype(F32))
        return out


class NAdam(Adam):
    """bats/Optimizers/NAdamOptimizer.py:36-58."""

    def step(self, grads):
        self.t += 1
        if self.m is None:
            self.m = [None if g is None else np.zeros(g.shape, F32) for g in grads]
            self.v = [None if g is None else np.zeros(g.shape, F32) for g in grads]
        s1 = 1.0 - F64(self.b1) ** self.t
        s2 = 1.0 - F64(self.b2) ** self.t
        out = []
        for i, g in enumerate(grads):
            if g is None:
                out.append(None)
                continue
            g64 = np.asarray(g, dtype=F64)
            m = (F64(self.b1) * self.m[i].astype(F64) + F64(self.omb1) * g64).astype(F32)
            v = (F64(self.b2) * self.v[i].astype(F64) + F64(self.omb2) * g64 * g64).astype(F32)
            self.m[i], self.v[i] = m, v
            mh = (m.astype(F64) / s1).astype(F32).astype(F64)
            vh = (v.astype(F64) / s2).astype(F32).astype(F64)
            d = -(F64(self.lr) / (np.sqrt(vh) + F64(self.eps))) * (
                F64(self.b1) * mh + F64(self.omb1) * g64 / s1)
            out.append(d.astype(F32))
        return out


def gd_step(grads, learning_rate):
    """bats/Optimizers/GradientDescentOptimizer.py:8-9."""
    lr = F64(F32(learning_rate))
    return [None if g is None else (-lr * np.asarray(g, dtype=F64)).astype(F32) for g in grads]
