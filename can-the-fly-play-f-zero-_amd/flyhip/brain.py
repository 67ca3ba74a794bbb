"""``Brain``: one fly on the MI355X engine, drop-in for flyzero/brain.py:79-172.

Same constructor (``weights, params, seed, backend``), methods (``reset``, ``set_bias``,
``set_input``, ``_propagate``, ``run``) and attributes (``p, n, indptr, indices, data, rng,
n_delay, n_ref, decay_m, decay_syn, kick, v, g, ref, t, spike_count, input_idx, input_rate,
bias``).  ``run`` returns numpy int32 counts as the reference's CPU ``Brain`` does.

Deliberate differences:

* ``backend`` accepts only ``"auto"`` and ``"hip"``; there is no CPU fallback.
* Poisson input uses the counter-hash RNG of flyzero/gpu.py:66 keyed by (seed, step, neuron)
  instead of numba's reseeded MT19937 (brain.py:34,144); a neuron listed twice in
  ``set_input`` gets the sum of its rates (superposed Poisson trains) in one draw per step.
* ``g, ref`` are torch views of the GPU state (``ref`` is int8); ``v`` is an mV view over the
  rest-relative membrane state ``u = v - v_rest``; ``pending`` is a read-only mV copy of the
  integer ring.
* Edits of ``data`` (mV) reach the engine on the next ``run``: reading the ``data`` attribute
  marks it for a diff against the uploaded weights, and changed entries are re-quantised.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from .gpu import BatchBrain
from .params import LIFParams


class Brain:
    def __init__(self, weights: sp.csr_matrix, params: LIFParams | None = None, seed: int = 0,
                 backend: str = "auto", frac_bits: int | None = None, device=None):
        if backend not in ("auto", "hip"):
            raise ValueError(f"flyhip.Brain runs on the GPU only (backend {backend!r} is not available)")
        self.backend = "hip"
        self._bb = bb = BatchBrain(weights, params, batch=1, seed=seed, frac_bits=frac_bits, device=device)
        self.p = bb.p
        self.n = bb.n
        self.indptr = bb.indptr
        self.indices = bb.indices
        self._data = bb.data
        self._data_uploaded = bb.data.copy()
        self._data_touched = False
        self.rng = np.random.default_rng(seed)
        self.n_delay, self.n_ref = bb.n_delay, bb.n_ref
        self.decay_m, self.decay_syn, self.kick = bb.decay_m, bb.decay_syn, bb.kick
        self.input_idx = np.zeros(0, np.int64)
        self.input_rate = np.zeros(0, np.float32)
        self.bias = np.zeros(self.n, np.float32)
        self._in_sig = None
        self.spike_count = np.zeros(self.n, np.int32)

    # state views -------------------------------------------------------------
    @property
    def v(self):
        """mV view over the engine's rest-relative membrane state (slice reads and writes)."""
        from .state import MembraneView

        return MembraneView(self._bb.u[0], self.p.v_rest)

    @property
    def u(self) -> torch.Tensor:
        return self._bb.u[0]

    @property
    def g(self) -> torch.Tensor:
        return self._bb.g[0]

    @property
    def ref(self) -> torch.Tensor:
        return self._bb.ref[0]

    @property
    def pending(self) -> torch.Tensor:
        return self._bb.pending[:, 0]

    @property
    def t(self) -> int:
        return self._bb.t

    @t.setter
    def t(self, value: int):
        self._bb.t = value

    @property
    def data(self) -> np.ndarray:
        self._data_touched = True
        return self._data

    @data.setter
    def data(self, value):
        self._data[...] = value
        self._data_touched = True

    # API ---------------------------------------------------------------------
    def reset(self):
        self._bb.reset()
        self.spike_count = np.zeros(self.n, np.int32)

    def set_bias(self, idx: np.ndarray, mv):
        """Steady depolarising current (mV of steady-state shift) for the given neurons."""
        self.bias[:] = 0.0
        self.bias[np.asarray(idx, np.int64)] = mv

    def set_input(self, idx: np.ndarray, rate_hz):
        """Poisson drive (Hz) for the given neurons, replacing previous input."""
        self.input_idx = np.asarray(idx, np.int64)
        self.input_rate = np.broadcast_to(np.asarray(rate_hz, np.float32), self.input_idx.shape).copy()

    def _propagate(self, spiking: np.ndarray, out: np.ndarray):
        """``out += w[spiking].sum(0) * w_syn`` (flyzero/brain.py:123-131), summed as integers on
        the GPU and scaled once."""
        self._sync_weights()
        bb = self._bb
        lst = torch.as_tensor(np.asarray(spiking, np.int32), device=bb.device)
        acc = torch.zeros(self.n, dtype=torch.int32, device=bb.device)
        _lib.check(bb._lib.fz_propagate(bb._h, int(lst.numel()), _lib.ptr(lst), _lib.ptr(acc),
                                        _lib.stream_handle(bb.device)), "fz_propagate")
        out += (acc.to(torch.float32) * float(bb.in_scale)).cpu().numpy()

    def run(self, duration_ms: float) -> np.ndarray:
        """Advance the simulation; returns per-neuron spike counts for this window."""
        self._sync_weights()
        self._sync_inputs()
        counts = self._bb.run(duration_ms, copy=False)[0].cpu().numpy().copy()
        self.spike_count += counts
        return counts

    # host -> engine ----------------------------------------------------------
    def _sync_inputs(self):
        sig = (self.input_idx.tobytes(), self.input_rate.tobytes(), self.bias.tobytes())
        if sig == self._in_sig:
            return
        bb, p = self._bb, self.p
        idx = np.asarray(self.input_idx, np.int64)
        prob = (np.asarray(self.input_rate, np.float32) * np.float32(p.dt / 1000.0)).astype(np.float32)
        uidx, inv = np.unique(idx, return_inverse=True)
        psum = np.zeros(len(uidx), np.float32)
        np.add.at(psum, inv, prob)
        bidx = np.flatnonzero(self.bias)
        ext = np.union1d(uidx, bidx)
        if len(ext):
            bb.slots(ext)
        bb.p_ext.zero_()
        bb.bias_ext.zero_()
        if len(uidx):
            bb.p_ext[0, bb.slots(uidx)] = torch.from_numpy(psum).to(bb.device)
        if len(bidx):
            bb.bias_ext[0, bb.slots(bidx)] = torch.from_numpy(self.bias[bidx].astype(np.float32)).to(bb.device)
        self._in_sig = sig

    def _sync_weights(self):
        if not self._data_touched:
            return
        self._data_touched = False
        changed = np.flatnonzero(self._data != self._data_uploaded)
        if not len(changed):
            return
        bb = self._bb
        q = np.rint((self._data[changed] * bb.pw_qscale).astype(np.float32).astype(np.float64)).astype(np.int64)
        lim = 2 ** 15 if bb.w_bits == 16 else 2 ** 31
        if np.abs(q).max() >= lim:
            raise ValueError("edited weight does not fit the engine's weight format")
        dt = torch.int16 if bb.w_bits == 16 else torch.int32
        bb.d_w[torch.from_numpy(changed).to(bb.device)] = torch.from_numpy(q).to(device=bb.device, dtype=dt)
        self._data_uploaded[changed] = self._data[changed]
