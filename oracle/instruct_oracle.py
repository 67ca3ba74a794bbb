"""CPU restatement of the instructed learning rules (test infrastructure only).

Restates flyzero/instruct.py:124-148 as batched in flyzero/batch.py:226-451 (BatchInstruct,
BatchInstructDeep, BatchInstructDeep2) including the deep_update kernel of
flyzero/batch.py:27-52, in float32 numpy with the summation orders the device uses: each fly's
slots in ascending order, each L1 / L2 neuron's synapses in ascending plastic index.  Every
float32 operation is a single IEEE operation on float32 operands, so the results are the
device's bits.  Nothing in the product imports this module.
"""

from __future__ import annotations

import numpy as np

f32 = np.float32


def traces(counts, upre, upost, trace, rate, a_pre, pre_gain, a_rate, hz):
    """flyzero/batch.py:289-293: trace += a_pre * (c_pre * hz * tau_pre_s - trace),
    rate += a_rate * (c_post * hz - rate); in place, float32."""
    a_pre, pre_gain, a_rate, hz = f32(a_pre), f32(pre_gain), f32(a_rate), f32(hz)
    c = counts[:, upre].astype(f32)
    trace[...] = trace + a_pre * (c * hz * pre_gain - trace)
    c = counts[:, upost].astype(f32)
    rate[...] = rate + a_rate * (c * hz - rate)


def fly_slots(fly_of_slot):
    fos = np.asarray(fly_of_slot)
    return [np.flatnonzero(fos == f) for f in range(int(fos.max()) + 1)]


def clip_dale(w_plus_d, sign, wmax):
    """sign * clip(sign * w, 0, wmax) (flyzero/batch.py:166-172, 47-49)."""
    return sign * np.minimum(np.maximum(sign * w_plus_d, f32(0.0)), wmax)


def delta_rule(pw, sel, pre_k, group, post_k, mask, trace, rate, tgt, eta_b, learn, fly_of_slot, sign, wmax):
    """The DN synapses' update (flyzero/batch.py:320-322 with apply, :158-172): per fly, the sum
    over its learning slots in ascending order of (1 / n) * (eta_b * trace * err * mask), then
    every slot of the fly gets w = sign * clip(sign * (w + d), 0, wmax).  ``sel`` None = all
    plastic synapses.  In place on ``pw`` [B, n_pl]."""
    sel = np.arange(pw.shape[1]) if sel is None else np.asarray(sel)
    m = np.ones(len(sel), f32) if mask is None else np.asarray(mask).astype(f32)
    for slots in fly_slots(fly_of_slot):
        nl = int(np.sum(learn[slots]))
        inv_n = f32(1.0) / f32(max(nl, 1))
        acc = np.zeros(len(sel), f32)
        for b in slots:
            if not learn[b]:
                continue
            err = tgt[b, group] - rate[b, post_k]
            dw = f32(eta_b[b]) * trace[b, pre_k] * err * m
            acc = acc + inv_n * dw
        for b in slots:
            pw[b, sel] = clip_dale(pw[b, sel] + acc, sign[sel], wmax[sel])


def retrograde(pw, lists, tgt, rate, w_syn):
    """delta[b][l] = sum over the neuron's synapses (syn, group, post_k) in the given order of
    (pw[b][syn] / w_syn) * (tgt[b][group] - rate[b][post_k]) (flyzero/batch.py:395-399)."""
    B = pw.shape[0]
    out = np.zeros((B, len(lists)), f32)
    w_syn = f32(w_syn)
    for l, entries in enumerate(lists):
        acc = np.zeros(B, f32)
        for syn, g, k in entries:
            acc = acc + (pw[:, syn] / w_syn) * (tgt[:, g] - rate[:, k])
        out[:, l] = acc
    return out


def deep_update(pw, sel, pre_k, post_l, trace, delta, coef, fly_of_slot, sign, wmax):
    """flyzero/batch.py:27-52: per fly, acc = sum over its slots with coef != 0, ascending, of
    coef * trace * delta; slots of the fly get w = sign * clip(sign * (w + acc), 0, wmax) where
    acc != 0.  In place."""
    sel = np.asarray(sel)
    for slots in fly_slots(fly_of_slot):
        acc = np.zeros(len(sel), f32)
        for b in slots:
            c = f32(coef[b])
            if c != 0:
                acc = acc + c * trace[b, pre_k] * delta[b, post_l]
        nz = acc != 0
        for b in slots:
            pw[b, sel] = np.where(nz, clip_dale(pw[b, sel] + acc, sign[sel], wmax[sel]), pw[b, sel])


def intrinsic_bias(ib, bias_ext, dn_slot, dn_group, dn_rate_k, tgt, rate, learn, fly_of_slot, eta_bias, ratio, limit):
    """flyzero/batch.py:302-310: per fly and DN the error averaged over the fly's learning slots
    (ascending, each term scaled by 1 / n); then every slot: ib' = clip(ib + eta_bias * mean *
    ratio_b, -limit, limit), bias_ext += ib' - ib.  In place."""
    eta_bias, limit = f32(eta_bias), f32(limit)
    for slots in fly_slots(fly_of_slot):
        nl = int(np.sum(learn[slots]))
        inv_n = f32(1.0) / f32(max(nl, 1))
        mean = np.zeros(len(dn_group), f32)
        for b in slots:
            if learn[b]:
                mean = mean + inv_n * (tgt[b, dn_group] - rate[b, dn_rate_k])
        for b in slots:
            step = eta_bias * mean * f32(ratio[b])
            new = np.minimum(np.maximum(ib[b] + step, -limit), limit)
            bias_ext[b, dn_slot] = bias_ext[b, dn_slot] + (new - ib[b])
            ib[b] = new


def deep_coef(learn, eta_b, eta_layer, eta, deep_scale, fly_of_slot):
    """coef_b = eta_b * (eta_layer / eta) * scale_b / learning slots of its fly, 0 where not
    learning, computed in float64 and rounded once (flyzero/batch.py:394-397)."""
    fos = np.asarray(fly_of_slot)
    rate = np.asarray(eta_b, np.float64) * (eta_layer / eta) * np.asarray(deep_scale, np.float64)
    n_act = np.bincount(fos, weights=np.asarray(learn, np.float64), minlength=int(fos.max()) + 1)
    return np.where(learn, rate / np.maximum(n_act[fos], 1), 0.0).astype(f32)
