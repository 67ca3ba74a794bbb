"""CPU oracle for the whole-brain LIF step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
module, and only as the checker.  The shipped library (``flyhip``) never imports it.

``OracleBrain`` builds its own network arrays from the same inputs the product receives (a
``scipy.sparse`` matrix of signed synapse counts, ``LIFParams``, plastic CSR positions) and runs
the C restatement in ``lif_oracle.c`` (``liblif_oracle.so``, built by ``oracle/Makefile``).
Reference semantics it restates:

* constants ``n_delay``, ``n_ref``, ``decay_m``, ``decay_syn``, ``kick``: flyzero/brain.py:92-96
* CSR layout and ``data = count * w_syn``: flyzero/brain.py:84-89
* fixed / plastic split of the CSR by position: flyzero/gpu.py:150-163
* per-step order of operations: flyzero/brain.py:36-57 (see lif_oracle.c)
* ext slots, per-step probability ``p = Hz * dt / 1000``: flyzero/gpu.py:174-226
* counter-hash RNG: flyzero/gpu.py:29-31,66-67

Integer synaptic sums are this build's contract (the reference sums float32 mV): a weight of
``c`` synapses is the fixed-point integer ``rint(c * 2**frac_bits)`` and the ring value is
scaled once on arrival by ``w_syn * 2**-frac_bits``.
"""

from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblif_oracle.so")


@dataclass
class OracleParams:
    """Same fields and defaults as flyzero/brain.py:65-76 (``LIFParams``)."""
    v_rest: float = -52.0
    v_reset: float = -52.0
    v_th: float = -45.0
    tau_m: float = 20.0
    tau_syn: float = 5.0
    t_ref: float = 2.2
    delay: float = 1.8
    w_syn: float = 0.275
    f_poisson: float = 250.0
    dt: float = 0.1


def constants(p) -> dict:
    """Derived constants, restated from flyzero/brain.py:92-96."""
    return {
        "n_delay": max(1, int(round(p.delay / p.dt))),
        "n_ref": int(round(p.t_ref / p.dt)),
        "decay_m": np.float32(p.dt / p.tau_m),
        "decay_syn": np.float32(np.exp(-p.dt / p.tau_syn)),
        "kick": np.float32(p.f_poisson * p.w_syn),
        "v_rest": np.float32(p.v_rest),
        "v_th": np.float32(p.v_th),
        "v_reset": np.float32(p.v_reset),
        "u_th": np.float32(np.float32(p.v_th) - np.float32(p.v_rest)),
        "u_reset": np.float32(np.float32(p.v_reset) - np.float32(p.v_rest)),
        "w_syn": np.float32(p.w_syn),
    }


def choose_frac_bits(counts_f32: np.ndarray, col: np.ndarray, n: int) -> int:
    """0 for whole synapse counts; otherwise the most fractional bits (<= 24) that keep the
    largest possible per-step arrival (column sum of |count|) under 2**27."""
    if counts_f32.size == 0 or np.all(counts_f32 == np.rint(counts_f32)):
        return 0
    colsum = np.bincount(col, weights=np.abs(counts_f32.astype(np.float64)), minlength=n)
    cmax = max(float(colsum.max()), 1e-30)
    return int(min(24, max(0, math.floor(math.log2(2.0 ** 27 / cmax)))))


class _Net(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("n_delay", ctypes.c_int), ("n_ref", ctypes.c_int),
        ("decay_syn", ctypes.c_float), ("decay_m", ctypes.c_float), ("u_th", ctypes.c_float),
        ("u_reset", ctypes.c_float), ("kick", ctypes.c_float),
        ("in_scale", ctypes.c_float), ("pw_qscale", ctypes.c_float),
        ("indptr", ctypes.c_void_p), ("col", ctypes.c_void_p), ("w", ctypes.c_void_p),
        ("n_pl", ctypes.c_int), ("pl_indptr", ctypes.c_void_p), ("pl_post", ctypes.c_void_p),
        ("n_ext", ctypes.c_int), ("ext_idx", ctypes.c_void_p),
    ]


class _State(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_int), ("u", ctypes.c_void_p), ("g", ctypes.c_void_p), ("ref", ctypes.c_void_p),
        ("ring", ctypes.c_void_p), ("t", ctypes.c_longlong), ("step_no", ctypes.c_uint),
        ("u64", ctypes.c_void_p), ("g64", ctypes.c_void_p),
    ]


class _Input(ctypes.Structure):
    _fields_ = [
        ("p_ext", ctypes.c_void_p), ("bias_ext", ctypes.c_void_p), ("pw", ctypes.c_void_p),
        ("hits", ctypes.c_void_p), ("alive", ctypes.c_void_p), ("seed", ctypes.c_uint),
        ("nthreads", ctypes.c_int),
    ]


class _Report(ctypes.Structure):
    _fields_ = [("max_dv", ctypes.c_double), ("max_dv_sub", ctypes.c_double), ("max_abs_v", ctypes.c_double),
                ("max_dg", ctypes.c_double), ("flips", ctypes.c_longlong),
                ("dv_step", ctypes.c_void_p), ("one_step", ctypes.c_int), ("step_max_dv", ctypes.c_double),
                ("step_max_dg", ctypes.c_double), ("step_flips", ctypes.c_longlong),
                ("step_dv_step", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: build it with `make -C {_HERE}`")
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_mix.restype = ctypes.c_uint
        _lib.orc_mix.argtypes = [ctypes.c_uint]
        _lib.orc_hash.restype = ctypes.c_uint
        _lib.orc_hash.argtypes = [ctypes.c_uint] * 4
        _lib.orc_lif_run.restype = ctypes.c_longlong
        _lib.orc_lif_run.argtypes = [ctypes.POINTER(_Net), ctypes.POINTER(_State), ctypes.POINTER(_Input),
                                     ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(_Report)]
    return _lib


def mix(x: int) -> int:
    return int(lib().orc_mix(x & 0xFFFFFFFF))


def _ptr(a):
    return None if a is None else a.ctypes.data


class OracleNet:
    """Network arrays for the oracle, built independently of the product's host code."""

    def __init__(self, weights, params=None, plastic_pos=None, frac_bits=None):
        self.p = p = params or OracleParams()
        self.c = c = constants(p)
        w = sp.csr_matrix(weights, dtype=np.float32)
        w.sort_indices()
        self.n = n = w.shape[0]
        self.indptr_full = w.indptr.astype(np.int64)
        self.indices_full = w.indices.astype(np.int32)
        cnt = w.data.astype(np.float32)
        self.frac_bits = s = choose_frac_bits(cnt, self.indices_full, n) if frac_bits is None else int(frac_bits)
        q = np.rint(cnt.astype(np.float64) * (2.0 ** s)).astype(np.int64)
        if q.size and np.abs(q).max() >= 2 ** 31:
            raise ValueError("weights do not fit int32 fixed point")
        pos = np.zeros(0, np.int64) if plastic_pos is None else np.sort(np.asarray(plastic_pos, np.int64))
        self.plastic_pos = pos
        rows = np.repeat(np.arange(n), np.diff(self.indptr_full))
        keep = np.ones(len(cnt), bool)
        keep[pos] = False
        self.indptr = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=n))]).astype(np.int64)
        self.col = np.ascontiguousarray(self.indices_full[keep])
        self.w = np.ascontiguousarray(q[keep].astype(np.int32))
        self.pl_pre = rows[pos]
        self.pl_post = np.ascontiguousarray(self.indices_full[pos])
        self.pl_indptr = np.searchsorted(self.pl_pre, np.arange(n + 1)).astype(np.int64)
        self.w0 = (cnt * c["w_syn"]).astype(np.float32)[pos]          # mV, like Brain.data[pos]
        self.in_scale = np.float32(np.float64(c["w_syn"]) * 2.0 ** -s)
        self.pw_qscale = np.float32(2.0 ** s / np.float64(c["w_syn"]))


class MembraneMV:
    """mV view (v = u + v_rest) of a membrane state held relative to rest; slice reads and
    writes, a write stores u = f32(f32(value) - f32(v_rest))."""

    def __init__(self, u: np.ndarray, rest: np.float32):
        self._u, self._rest = u, np.float32(rest)

    def __getitem__(self, k):
        return (self._u[k] + self._rest).astype(np.float32)

    def __setitem__(self, k, val):
        self._u[k] = np.asarray(val, np.float32) - self._rest

    def __array__(self, dtype=None, copy=None):
        a = (self._u + self._rest).astype(np.float32)
        return a if dtype is None else a.astype(dtype)

    @property
    def shape(self):
        return self._u.shape


class OracleBrain:
    """B flies on the CPU oracle.  State layout mirrors the GPU engine's: u (membrane potential
    relative to rest), g float32 [B, n]; ref int8 [B, n]; ring int32 [B, n_delay, n]."""

    def __init__(self, weights, params=None, batch=1, seed=0, plastic_pos=None, frac_bits=None):
        self.net = net = OracleNet(weights, params, plastic_pos, frac_bits)
        self.p, self.c, self.n, self.B = net.p, net.c, net.n, batch
        self.seed = int(seed) & 0xFFFFFFFF
        self.ext_idx = np.full(self.n, -1, np.int32)
        self.ext = np.zeros(0, np.int64)
        self.p_ext = np.zeros((batch, 0), np.float32)
        self.bias_ext = np.zeros((batch, 0), np.float32)
        self.pw = np.tile(net.w0, (batch, 1)).astype(np.float32)
        self.alive = None
        self.reset()

    def reset(self):
        n, B, nd = self.n, self.B, self.c["n_delay"]
        self.u = np.zeros((B, n), np.float32)
        self.g = np.zeros((B, n), np.float32)
        self.ref = np.zeros((B, n), np.int8)
        self.ring = np.zeros((B, nd, n), np.int32)
        self.t = 0
        self.step_no = 0
        self.u64 = None
        self.g64 = None

    @property
    def v(self) -> MembraneMV:
        return MembraneMV(self.u, self.c["v_rest"])

    def enable_shadow(self):
        self.u64 = self.u.astype(np.float64)
        self.g64 = self.g.astype(np.float64)

    def slots(self, idx) -> np.ndarray:
        idx = np.asarray(idx, np.int64).ravel()
        new = np.unique(idx[self.ext_idx[idx] < 0])
        if len(new):
            k = len(self.ext)
            self.ext = np.r_[self.ext, new]
            self.ext_idx[new] = np.arange(k, k + len(new), dtype=np.int32)
            pad = np.zeros((self.B, len(new)), np.float32)
            self.p_ext = np.concatenate([self.p_ext, pad], 1)
            self.bias_ext = np.concatenate([self.bias_ext, pad], 1)
        return self.ext_idx[idx]

    def set_input(self, idx, rate_hz):
        j = self.slots(idx)
        self.p_ext[:] = 0
        self.p_ext[:, j] = np.asarray(rate_hz, np.float32) * np.float32(self.p.dt / 1000.0)

    def set_bias(self, idx, mv):
        j = self.slots(idx)
        self.bias_ext[:] = 0
        self.bias_ext[:, j] = np.asarray(mv, np.float32)

    def run_steps(self, steps: int, hits=None, raster=False, nthreads=1, dv_per_step=False, one_step=False):
        """Advance ``steps`` steps; returns counts [B, n] (and raster words).  ``one_step``: also
        advance every step once in FP64 from the FP32 state and record the difference
        (``last_report['step_max_dv']`` etc.)."""
        net, c = self.net, self.c
        counts = np.zeros((self.B, self.n), np.int32)
        rw = (self.n + 31) // 32
        ras = np.zeros((steps, self.B, rw), np.uint32) if raster else None
        nn = _Net(self.n, c["n_delay"], c["n_ref"], c["decay_syn"], c["decay_m"], c["u_th"], c["u_reset"],
                  c["kick"], net.in_scale, net.pw_qscale, _ptr(net.indptr), _ptr(net.col),
                  _ptr(net.w), len(net.plastic_pos), _ptr(net.pl_indptr), _ptr(net.pl_post),
                  self.p_ext.shape[1], _ptr(self.ext_idx))
        st = _State(self.B, _ptr(self.u), _ptr(self.g), _ptr(self.ref), _ptr(self.ring), self.t,
                    self.step_no & 0xFFFFFFFF, _ptr(self.u64), _ptr(self.g64))
        p_ext = np.ascontiguousarray(self.p_ext, np.float32)
        bias = np.ascontiguousarray(self.bias_ext, np.float32)
        pw = np.ascontiguousarray(self.pw, np.float32) if len(net.plastic_pos) else None
        if hits is not None:
            hits = np.ascontiguousarray(hits, np.uint32)
            hw = (p_ext.shape[1] + 31) // 32
            assert hits.shape == (steps, self.B, hw), (hits.shape, (steps, self.B, hw))
        alive = None if self.alive is None else np.ascontiguousarray(self.alive, np.uint8)
        inp = _Input(_ptr(p_ext), _ptr(bias), _ptr(pw), _ptr(hits), _ptr(alive), self.seed, nthreads)
        dv = np.zeros(steps, np.float64) if dv_per_step else None
        sdv = np.zeros(steps, np.float64) if (dv_per_step and one_step) else None
        rep = _Report(0.0, 0.0, 0.0, 0.0, 0, _ptr(dv), int(bool(one_step)), 0.0, 0.0, 0, _ptr(sdv))
        n_sp = lib().orc_lif_run(ctypes.byref(nn), ctypes.byref(st), ctypes.byref(inp), steps,
                                 counts.ctypes.data, _ptr(ras), ctypes.byref(rep))
        if n_sp < 0:
            raise RuntimeError(f"oracle rejected the arguments ({n_sp})")
        self.t += steps
        self.step_no = (self.step_no + steps) & 0xFFFFFFFF
        self.last_report = {"max_dv": rep.max_dv, "max_dv_sub": rep.max_dv_sub, "max_abs_v": rep.max_abs_v,
                            "max_dg": rep.max_dg, "flips": rep.flips, "dv_step": dv,
                            "step_max_dv": rep.step_max_dv, "step_max_dg": rep.step_max_dg,
                            "step_flips": rep.step_flips, "step_dv_step": sdv}
        if raster:
            return counts, ras
        return counts

    def run(self, duration_ms: float, **kw):
        """``Brain.run`` / ``BatchBrain.run`` window: steps = max(1, round(ms / dt))."""
        return self.run_steps(max(1, int(round(duration_ms / self.p.dt))), **kw)


def raster_to_bool(ras: np.ndarray, n: int) -> np.ndarray:
    """[steps, B, words] uint32 -> [steps, B, n] bool."""
    bits = np.unpackbits(ras.view(np.uint8).reshape(ras.shape[0], ras.shape[1], -1), axis=-1,
                         bitorder="little")
    return bits[..., :n].astype(bool)
