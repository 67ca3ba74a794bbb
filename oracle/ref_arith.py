"""The reference's own LIF arithmetic (``ref_arith.c``) -- TEST INFRASTRUCTURE ONLY.

Used by tests/ to measure how far this build's numeric contract (integer synaptic sums scaled
once, membrane state held relative to rest) takes the spike trains from the arithmetic of
flyzero/brain.py:32-58 (numba backend; ``fused=True``: with the fused multiply-adds its
fastmath may emit) and of its numpy backend, brain.py:147-170 (``numpy=True``).  Kicks come from
precomputed hits, as for the oracle's raster tests.
"""

from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp

from .oracle import OracleParams, constants, lib


class _RefNet(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("n_delay", ctypes.c_int), ("n_ref", ctypes.c_int),
                ("decay_syn", ctypes.c_float), ("decay_m", ctypes.c_float), ("v_rest", ctypes.c_float),
                ("v_th", ctypes.c_float), ("v_reset", ctypes.c_float), ("kick", ctypes.c_float),
                ("indptr", ctypes.c_void_p), ("indices", ctypes.c_void_p), ("data", ctypes.c_void_p),
                ("n_in", ctypes.c_int), ("input_idx", ctypes.c_void_p), ("numpy", ctypes.c_int),
                ("fused", ctypes.c_int)]


class _RefState(ctypes.Structure):
    _fields_ = [("v", ctypes.c_void_p), ("g", ctypes.c_void_p), ("ref", ctypes.c_void_p),
                ("pending", ctypes.c_void_p), ("bias", ctypes.c_void_p), ("t", ctypes.c_longlong)]


class RefArithBrain:
    """One fly in the reference's arithmetic: v in mV (float32), float32 ``data = count * w_syn``
    added per spike into a float32 delay ring, int16 refractory counter (brain.py:84-110)."""

    def __init__(self, weights, params=None, numpy_backend=False, fused=False):
        self.p = p = params or OracleParams()
        self.c = c = constants(p)
        w = sp.csr_matrix(weights, dtype=np.float32)
        w.sort_indices()
        self.n = n = w.shape[0]
        self.indptr = w.indptr.astype(np.int64)
        self.indices = w.indices.astype(np.int32)
        self.data = (w.data * np.float32(p.w_syn)).astype(np.float32)   # brain.py:89, float32 product
        self.numpy_backend, self.fused = bool(numpy_backend), bool(fused)
        self.v = np.full(n, c["v_rest"], np.float32)
        self.g = np.zeros(n, np.float32)
        self.ref = np.zeros(n, np.int16)
        self.pending = np.zeros((c["n_delay"], n), np.float32)
        self.bias = np.zeros(n, np.float32)
        self.input_idx = np.zeros(0, np.int64)
        self.t = 0
        f = lib().orc_ref_run
        f.restype = ctypes.c_longlong
        f.argtypes = [ctypes.POINTER(_RefNet), ctypes.POINTER(_RefState), ctypes.c_int, ctypes.c_void_p,
                      ctypes.c_void_p, ctypes.c_void_p]

    def set_input(self, idx):
        """The input list (one hit bit per entry and step; entries may repeat a neuron)."""
        self.input_idx = np.ascontiguousarray(idx, np.int64)

    def set_bias(self, idx, mv):
        self.bias[:] = 0.0
        self.bias[np.asarray(idx, np.int64)] = mv

    def run_steps(self, steps: int, hits=None, raster=False):
        """hits: uint32 [steps, ceil(n_in/32)] or [steps, 1, ceil(n_in/32)] (the oracle's layout
        for one fly).  Returns counts [n] (and raster words [steps, ceil(n/32)])."""
        c = self.c
        hw = (len(self.input_idx) + 31) // 32
        if hits is not None:
            hits = np.ascontiguousarray(hits, np.uint32).reshape(steps, hw)
        counts = np.zeros(self.n, np.int32)
        ras = np.zeros((steps, (self.n + 31) // 32), np.uint32) if raster else None
        net = _RefNet(self.n, c["n_delay"], c["n_ref"], c["decay_syn"], c["decay_m"], c["v_rest"], c["v_th"],
                      c["v_reset"], c["kick"], self.indptr.ctypes.data, self.indices.ctypes.data,
                      self.data.ctypes.data, len(self.input_idx), self.input_idx.ctypes.data if len(self.input_idx)
                      else None, int(self.numpy_backend), int(self.fused))
        st = _RefState(self.v.ctypes.data, self.g.ctypes.data, self.ref.ctypes.data, self.pending.ctypes.data,
                       self.bias.ctypes.data, self.t)
        rc = lib().orc_ref_run(ctypes.byref(net), ctypes.byref(st), steps,
                               None if hits is None else hits.ctypes.data, counts.ctypes.data,
                               None if ras is None else ras.ctypes.data)
        if rc < 0:
            raise RuntimeError(f"orc_ref_run rejected the arguments ({rc})")
        self.t = st.t
        return (counts, ras) if raster else counts
