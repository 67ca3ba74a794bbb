"""Model parameters, derived constants and the integer weight format.

``LIFParams`` keeps the reference's fields and defaults (flyzero/brain.py:65-76) so existing
callers construct it unchanged.  Derived constants follow flyzero/brain.py:92-96.

Weights.  The reference stores ``data = f32(count) * w_syn`` and sums float32 mV terms.  This
engine stores the signed synapse count of every connection as an integer and accumulates
integers in an int32 delay ring, scaling once on arrival by ``in_scale = w_syn * 2**-frac_bits``:

* whole synapse counts (FlyWire, the synthetic FlyWire-shaped brain): ``frac_bits = 0``,
  stored as int16 when every |count| < 32768;
* any other real-valued matrix: int32 fixed point with ``frac_bits`` fractional bits, the most
  (<= 24) that keep the largest possible per-step arrival (column sum of |count|) below 2**27.

Plastic weights stay in mV (the reference's ``pw`` contract, flyzero/gpu.py:134-136) and are
quantised on use as ``rintf(pw * pw_qscale)`` with ``pw_qscale = f32(2**frac_bits / w_syn)``.
Learning rules that edit them continuously (the instructed and deep rules) would lose every
edit under half a synapse with ``frac_bits = 0``, so a whole-synapse connectome with plastic
synapses gets ``PLASTIC_FRAC_BITS`` fractional bits (``plastic_frac_bits``): plastic weights
resolve w_syn / 256 = 1.07e-3 mV, and the fixed synapses' arrivals are unchanged bit for bit
(``(float)(c * 2**f) * (w_syn * 2**-f)`` is ``(float)c * w_syn`` exactly).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class LIFParams:
    v_rest: float = -52.0   # mV
    v_reset: float = -52.0  # mV
    v_th: float = -45.0     # mV
    tau_m: float = 20.0     # ms, membrane time constant
    tau_syn: float = 5.0    # ms, synaptic time constant
    t_ref: float = 2.2      # ms, refractory period
    delay: float = 1.8      # ms, synaptic delay
    w_syn: float = 0.275    # mV per synapse
    f_poisson: float = 250.0  # Poisson input weight = f_poisson * w_syn
    dt: float = 0.1         # ms


def derived(p: LIFParams) -> dict:
    """flyzero/brain.py:92-96 (same in flyzero/gpu.py:164-171)."""
    return {
        "n_delay": max(1, int(round(p.delay / p.dt))),
        "n_ref": int(round(p.t_ref / p.dt)),
        "decay_m": np.float32(p.dt / p.tau_m),
        "decay_syn": np.float32(np.exp(-p.dt / p.tau_syn)),
        "kick": np.float32(p.f_poisson * p.w_syn),
    }


def frac_bits_for(counts: np.ndarray, col: np.ndarray, n: int) -> int:
    """Fractional bits of the fixed-point weight format (see module docstring)."""
    c = np.asarray(counts, np.float32)
    if c.size == 0 or bool(np.all(c == np.rint(c))):
        return 0
    colsum = np.bincount(np.asarray(col), weights=np.abs(c.astype(np.float64)), minlength=n)
    cmax = max(float(colsum.max()), 1e-30)
    return int(min(24, max(0, math.floor(math.log2(2.0 ** 27 / cmax)))))


PLASTIC_FRAC_BITS = 8


def plastic_frac_bits(counts: np.ndarray, col: np.ndarray, n: int, tile: int | None = None) -> int:
    """Fractional bits for a whole-synapse connectome with plastic synapses: the most, up to
    ``PLASTIC_FRAC_BITS``, that keep the largest possible per-step arrival (column sum of
    |count|) below 2**27 and, for the tile engine, every fixed value beside its target's offset in
    one packed int32 (|count| * 2**f < 2**(31 - log2 tile))."""
    c = np.abs(np.asarray(counts, np.float64))
    if c.size == 0:
        return PLASTIC_FRAC_BITS
    colmax = float(np.bincount(np.asarray(col), weights=c, minlength=n).max())
    lbits = max(0, math.ceil(math.log2(tile))) if tile else 0
    f = PLASTIC_FRAC_BITS
    while f > 0 and (colmax * 2.0 ** f >= 2.0 ** 27 or (tile and float(c.max()) * 2.0 ** f >= 2.0 ** (31 - lbits))):
        f -= 1
    return f


def quantize(counts: np.ndarray, frac_bits: int) -> np.ndarray:
    """Signed synapse counts (float32) -> fixed-point int64 ``rint(count * 2**frac_bits)``."""
    return np.rint(np.asarray(counts, np.float32).astype(np.float64) * (2.0 ** frac_bits)).astype(np.int64)


def scales(w_syn: float, frac_bits: int) -> tuple[np.float32, np.float32]:
    """(in_scale, pw_qscale) as float32."""
    ws = np.float64(np.float32(w_syn))
    return np.float32(ws * 2.0 ** -frac_bits), np.float32(2.0 ** frac_bits / ws)
