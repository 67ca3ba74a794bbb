"""Vectorised numpy restatement of the LIF step -- TEST INFRASTRUCTURE ONLY.

A second, independently written oracle (per-step array operations instead of the C oracle's
per-neuron loop) used to cross-check ``lif_oracle.c`` on small networks.  Same semantics:
flyzero/brain.py:36-57 (numba order), ext slots and hash RNG of flyzero/gpu.py:29-31,55-69,
integer arrivals scaled once (see lif_oracle.h).
"""

from __future__ import annotations

import numpy as np

M32 = np.uint64(0xFFFFFFFF)


def mix(x):
    """lowbias32 on uint64 arrays holding 32-bit values (flyzero/gpu.py:29-31)."""
    x = np.asarray(x, np.uint64) & M32
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & M32
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & M32
    x ^= x >> np.uint64(16)
    return x


def run(net, ext_idx, v, g, ref, ring, t, step_no, steps, p_ext, bias_ext, pw, seed, hits=None, alive=None):
    """Advance in place; ``net`` is an ``oracle.OracleNet``, ``ext_idx`` the [n] slot map, ``v``
    the membrane potential relative to rest.  Returns counts [B, n]."""
    c = net.c
    B, n = v.shape
    nd = c["n_delay"]
    f32 = np.float32
    counts = np.zeros((B, n), np.int32)
    has_ext = p_ext is not None and p_ext.shape[1] > 0
    for s in range(steps):
        slot = (t + s) % nd
        step = np.uint64((step_no + s) & 0xFFFFFFFF)
        for b in range(B):
            if alive is not None and not alive[b]:
                continue
            acc = ring[b, slot].copy()
            ring[b, slot] = 0
            g[b] = g[b] * c["decay_syn"] + acc.astype(f32) * net.in_scale
            bias = np.zeros(n, f32)
            hit = np.zeros(n, bool)
            if has_ext:
                j = ext_idx
                m = j >= 0
                bias[m] = bias_ext[b, j[m]]
                if hits is not None:
                    jj = j[m]
                    words = hits[s, b, jj >> 5].astype(np.uint64)
                    hit[m] = ((words >> (jj & 31).astype(np.uint64)) & np.uint64(1)).astype(bool)
                else:
                    pj = np.zeros(n, f32)
                    pj[m] = p_ext[b, j[m]]
                    tid = (np.uint64(b) * np.uint64(n) + np.arange(n, dtype=np.uint64)) & M32
                    inner = mix((step * np.uint64(0x9E3779B9)) & M32 ^ mix(np.uint64((seed + b) & 0xFFFFFFFF)))
                    h = mix(tid ^ inner)
                    u = (h >> np.uint64(8)).astype(f32) * f32(1.0 / 16777216.0)
                    hit = m & (pj > 0) & (u < pj)
            refr = ref[b] > 0
            leak = ((g[b] + bias) - v[b]) * c["decay_m"]
            v[b] = np.where(refr, v[b], v[b] + leak).astype(f32)
            ref[b] = np.where(refr, ref[b] - 1, ref[b]).astype(np.int8)
            v[b] = np.where(hit, v[b] + c["kick"], v[b]).astype(f32)
            spk = np.flatnonzero((v[b] >= c["u_th"]) & (ref[b] <= 0))
            v[b, spk] = c["u_reset"]
            ref[b, spk] = c["n_ref"]
            counts[b, spk] += 1
            # push: integer adds into the slot just consumed
            for i in spk:
                k0, k1 = net.indptr[i], net.indptr[i + 1]
                np.add.at(ring[b, slot], net.col[k0:k1], net.w[k0:k1])
                if len(net.plastic_pos):
                    q0, q1 = net.pl_indptr[i], net.pl_indptr[i + 1]
                    q = np.rint((pw[b, q0:q1] * net.pw_qscale).astype(f32)).astype(np.int32)
                    np.add.at(ring[b, slot], net.pl_post[q0:q1], q)
    return counts
