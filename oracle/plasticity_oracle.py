"""FP64 oracle of the reward-modulated plasticity rule and the DN-rate readout -- TEST
INFRASTRUCTURE ONLY (the checker; the product never imports it).

Restates flyzero/learning.py:116-131 in its batched form flyzero/batch.py:482-495 (with the
per-fly averaging of batch.py:158-172), with this build's numeric contract: float64 throughout,
weights in synapse units, eligibility stored as float32, sums over a fly's slots in ascending
slot order, the simulation's weight = rint(w) synapses.  The DN readout restates
flyzero/batch.py:107-111 in float32.
"""

from __future__ import annotations

import numpy as np


class RewardOracle:
    def __init__(self, pl_pre, pl_post, w0_syn, B, cap=3.0, w_abs_max=np.inf, fly_of_slot=None):
        self.pre = np.asarray(pl_pre, np.int64)
        self.post = np.asarray(pl_post, np.int64)
        self.upost, self.post_k = np.unique(self.post, return_inverse=True)
        w0 = np.asarray(w0_syn, np.float64)
        self.sign = np.sign(w0)
        self.wmax = np.minimum(cap * np.maximum(np.abs(w0), 1.0), w_abs_max)
        self.B = B
        self.fos = np.arange(B) if fly_of_slot is None else np.asarray(fly_of_slot)
        self.elig = np.zeros((B, len(w0)), np.float32)
        self.w = np.tile(w0, (B, 1))
        self.post_mean = np.zeros((B, len(self.upost)))
        self.r_mean = np.zeros(B)

    def count(self):
        return np.rint(self.w).astype(np.int32)

    def pw(self, w_syn=0.275):
        return self.count().astype(np.float32) * np.float32(w_syn)

    def step(self, counts, reward, learn, ae, ap, ar, eta):
        """``eta`` in synapses per unit, one value or one per slot (the per-slot rate of
        flyzero/school.py:289-304); ``counts`` (B, n) int."""
        counts = np.asarray(counts)
        reward = np.broadcast_to(np.asarray(reward, np.float64), (self.B,))
        learn = np.broadcast_to(np.asarray(learn, bool), (self.B,))
        eta = np.broadcast_to(np.asarray(eta, np.float64), (self.B,))
        adv = reward - self.r_mean
        for b in range(self.B):
            c_pre = counts[b, self.pre].astype(np.float64)
            c_post = counts[b, self.post].astype(np.float64)
            dev = c_post - self.post_mean[b, self.post_k]
            e = ae * self.elig[b].astype(np.float64) + c_pre * dev
            self.elig[b] = e.astype(np.float32)
        if learn.any():
            for f in np.unique(self.fos):
                slots = np.flatnonzero(self.fos == f)            # ascending
                act = [b for b in slots if learn[b]]
                if not act:
                    continue
                acc = np.zeros(self.w.shape[1])
                for b in act:
                    acc = acc + (eta[b] * adv[b]) * self.elig[b].astype(np.float64)
                d = acc / float(len(act))
                for b in slots:
                    w = self.w[b] + d
                    m = np.clip(self.sign * w, 0.0, self.wmax)
                    self.w[b] = self.sign * m
        for b in range(self.B):
            c_post = counts[b, self.upost].astype(np.float64)
            self.post_mean[b] = self.post_mean[b] + ap * (c_post - self.post_mean[b])
        self.r_mean = self.r_mean + ar * adv
        return adv


def dn_rates_step(rates, counts, groups, window_ms, tau_ms=80.0):
    """float32 restatement of the DN readout: mean count of each group * 1000/window, smoothed."""
    alpha = np.float32(1.0 - np.exp(-window_ms / tau_ms))
    hz = np.float32(1000.0 / window_ms)
    out = rates.astype(np.float32).copy()
    for g, members in enumerate(groups):
        if len(members):
            s = counts[:, members].sum(1).astype(np.float32)
            inst = (s / np.float32(len(members))) * hz
        else:
            inst = np.zeros(len(rates), np.float32)
        out[:, g] = out[:, g] + alpha * (inst - out[:, g])
    return out
