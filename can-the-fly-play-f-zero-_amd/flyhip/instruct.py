"""Instructed learning on a ``BatchBrain``: drop-ins for ``BatchInstruct``,
``BatchInstructDeep`` and ``BatchInstructDeep2`` (flyzero/batch.py:226-451), the delta rule of
flyzero/instruct.py:124-148 for many slots at once.

Each motor DN gets an instructive signal, the rate it should have had given the teacher's
intent, and the synapses onto it follow ``w += eta * trace_pre * (target - rate)`` (Dale's law
and caps as everywhere).  The deep variants pass the DN errors back through the fly's own
synapses: an L1 neuron (a presynaptic partner of the DNs) gets ``delta = sum_j n_ij err_j``
and the synapses onto it learn ``eta_deep * trace * delta``; ``Deep2`` goes one layer further.

The work per frame is four HIP kernels on the device (``fz_instruct_traces``,
``fz_instruct_delta``, ``fz_instruct_update``, ``fz_instruct_bias``): float32 as written, each
fly's slots summed in ascending order, so a CPU restatement in the same order gives the same
bits.  Differences from the reference: its slot averaging is a float32 matrix product and its
retrograde sum an atomic scatter (orders unspecified); here both are fixed-order sums.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from .batch import fly_csr
from .connectome import GROUPS as _GROUPS

# the batched rules also instruct the giant fiber (flyzero/batch.py:20-22)
GROUPS = {**_GROUPS, "gf": ("DNp01", None)}
PLASTIC_TYPES = ("DNa02", "DNa01", "DNg02*", "DNp09", "MDN", "DNp01")


@dataclass
class InstructParams:
    """flyzero/instruct.py:29-38 (same names and defaults)."""
    eta: float = 2e-4          # mV per (spike trace x Hz of error)
    tau_pre_ms: float = 50.0   # presynaptic trace
    tau_rate_ms: float = 80.0  # motor neuron rate estimate
    high_hz: float = 80.0
    low_hz: float = 0.0
    rest_hz: float = 20.0      # steering neurons when going straight
    cap: float = 3.0


# the teacher's intent axes (flyzero/instruct.py:51-68): (button that pushes the axis up, button
# that pushes it down, or None)
INTENT_AXES = (("RIGHT", "LEFT"),   # steer: -1 left .. +1 right
               ("R", "L"),          # lean
               ("B", None),         # gas
               ("Y", None),         # brake
               ("A", None))         # boost


def intent(masks: np.ndarray, buttons: list, tau_frames: float = 15.0) -> np.ndarray:
    """The teacher's smoothed intention per frame, (frames, 5): steer, lean, gas, brake, boost
    (``INTENT_AXES``), each a causal exponential moving average in float32 with time constant
    ``tau_frames``, started at the first frame's value (flyzero/instruct.py:51-68)."""
    col = {name: k for k, name in enumerate(buttons)}
    pressed = np.asarray(masks, np.float32)
    axes = np.zeros((pressed.shape[0], len(INTENT_AXES)), np.float32)
    for k, (up, down) in enumerate(INTENT_AXES):
        axes[:, k] = pressed[:, col[up]]
        if down is not None:
            axes[:, k] -= pressed[:, col[down]]
    gain = np.float32(1.0 / tau_frames)
    smoothed = np.empty_like(axes)
    state = axes[0].copy() if len(axes) else np.zeros(len(INTENT_AXES), np.float32)
    for t, frame in enumerate(axes):
        state += gain * (frame - state)
        smoothed[t] = state
    return smoothed


def _opponent(x: float, p: InstructParams) -> tuple:
    """(left-side, right-side) target rates for a signed command x in [-1, 1]: the side the
    command points to rises linearly from rest to the high rate, the other falls from rest to
    the low rate (x = 0: both at rest)."""
    rise, fall = p.high_hz - p.rest_hz, p.rest_hz - p.low_hz
    neg, pos = max(-x, 0.0), max(x, 0.0)
    return p.rest_hz + rise * neg - fall * pos, p.rest_hz + rise * pos - fall * neg


def _level(x: float, p: InstructParams) -> float:
    """Target rate for an unsigned command x in [0, 1]: low .. high."""
    return p.low_hz + (p.high_hz - p.low_hz) * x


def targets_from_intent(x, p: InstructParams) -> dict:
    """Target rates of the 8 DN groups for an intent vector (the continuous form of
    flyzero/instruct.py:71-82): DNa02 follows steering on its own side, DNg02 (the wings) the
    opposite side's legs, DNa01 leaning, DNp09 gas and MDN brake."""
    steer, lean, gas, brake = (float(v) for v in x[:4])
    legs_l, legs_r = _opponent(steer, p)
    return {"a02L": legs_l, "a02R": legs_r, "g02L": legs_r, "g02R": legs_l,
            "a01L": _level(max(-lean, 0.0), p), "a01R": _level(max(lean, 0.0), p),
            "gas": _level(gas, p), "brake": _level(brake, p)}


def targets_from_buttons(b: dict, p: InstructParams) -> dict:
    """Target rates for pressed buttons (flyzero/instruct.py:85-99): the all-or-none form of
    ``targets_from_intent``.  One steering button held: that side's legs (DNa02) and the other
    side's wings (DNg02) at the high rate, the rest low; both or none: all four at rest."""
    left, right = bool(b.get("LEFT", False)), bool(b.get("RIGHT", False))
    hi, lo, rest = p.high_hz, p.low_hz, p.rest_hz
    legs_l, legs_r = (hi, lo) if left and not right else (lo, hi) if right and not left else (rest, rest)
    on = lambda key: hi if b.get(key) else lo   # noqa: E731
    return {"a02L": legs_l, "a02R": legs_r, "g02R": legs_l, "g02L": legs_r,
            "a01L": on("L"), "a01R": on("R"), "gas": on("B"), "brake": on("Y")}


def targets(x, p: InstructParams) -> list:
    """Target rates in ``GROUPS`` order for an intent vector, the giant fiber following boost
    (flyzero/batch.py:68-74)."""
    t = targets_from_intent(x, p)
    t["gf"] = p.low_hz + (p.high_hz - p.low_hz) * float(x[4])
    return [t[g] for g in GROUPS]


def _csr(conn):
    w = sp.csr_matrix(conn.weights, dtype=np.float32)
    w.sort_indices()
    return w, np.repeat(np.arange(w.shape[0]), np.diff(w.indptr))


def plastic_positions(conn, types=PLASTIC_TYPES) -> np.ndarray:
    """CSR positions of every synapse onto the given cell types (flyzero/batch.py:58-65)."""
    w, _ = _csr(conn)
    is_t = np.zeros(conn.n, bool)
    is_t[np.unique(np.concatenate([conn.find(t) for t in types]))] = True
    return np.flatnonzero(is_t[w.indices])


def deep_positions(conn, dn_pos: np.ndarray) -> np.ndarray:
    """CSR positions of every synapse onto the DNs' presynaptic partners, L1
    (flyzero/batch.py:325-333)."""
    w, rows = _csr(conn)
    is_l1 = np.zeros(conn.n, bool)
    is_l1[np.unique(rows[dn_pos])] = True
    return np.flatnonzero(is_l1[w.indices])


def deep2_positions(conn, dn_pos: np.ndarray) -> np.ndarray:
    """CSR positions of every synapse onto L2, the inputs of L1, leaving out L1 and the DNs
    (flyzero/batch.py:406-421)."""
    w, rows = _csr(conn)
    l1 = np.unique(rows[dn_pos])
    is_l1 = np.zeros(conn.n, bool)
    is_l1[l1] = True
    is_l2 = np.zeros(conn.n, bool)
    is_l2[np.unique(rows[is_l1[w.indices]])] = True
    is_l2[l1] = False
    is_l2[w.indices[dn_pos]] = False
    return np.flatnonzero(is_l2[w.indices])


class BatchInstruct:
    """Delta rule towards a teacher's target rates (flyzero/batch.py:226-322)."""

    def __init__(self, brain, conn, params: InstructParams | None = None, fly_of_slot=None,
                 eta_bias: float = 0.0, bias_limit: float = 8.0):
        """``eta_bias``: intrinsic plasticity, each instructed DN's own bias current (mV) follows
        ``eta_bias * (target - rate)`` within +-``bias_limit`` of where it started."""
        self.p = p = params or InstructParams()
        self.brain = brain
        self.B = brain.B
        self.dev = dev = brain.device
        self.eta_bias, self.bias_limit = float(eta_bias), float(bias_limit)
        self.names = list(GROUPS)
        group_of = np.full(conn.n, -1, np.int32)
        for k, g in enumerate(self.names):
            t, s = GROUPS[g]
            group_of[conn.find(t, s)] = k
        pre, post = np.asarray(brain.pl_pre), np.asarray(brain.pl_post)
        pg = group_of[post]
        self.mask_np = pg >= 0                            # only synapses onto instructed groups learn
        self.post_group_np = np.maximum(pg, 0)
        self.upre, self.pre_k_np = np.unique(pre, return_inverse=True)
        self.upost, self.post_k_np = np.unique(post, return_inverse=True)
        w0 = np.asarray(brain.w0, np.float32)
        self.sign_np = np.sign(w0).astype(np.float32)
        self.wmax_np = (p.cap * np.maximum(np.abs(w0), np.float32(brain.p.w_syn))).astype(np.float32)
        fos = np.arange(self.B) if fly_of_slot is None else np.asarray(fly_of_slot)
        self.fly_of_slot = fos
        self.n_flies, fptr, fslots = fly_csr(fos)
        T = self._t
        self.d_pre_k, self.d_post_k = T(self.pre_k_np, np.int32), T(self.post_k_np, np.int32)
        self.d_group, self.d_mask = T(self.post_group_np, np.int32), T(self.mask_np, np.uint8)
        self.d_upre, self.d_upost = T(self.upre, np.int32), T(self.upost, np.int32)
        self.sign, self.wmax = T(self.sign_np, np.float32), T(self.wmax_np, np.float32)
        self.d_fly_ptr, self.d_fly_slots = T(fptr, np.int32), T(fslots, np.int32)
        self.trace = torch.zeros((self.B, len(self.upre)), dtype=torch.float32, device=dev)
        self.rate = torch.zeros((self.B, len(self.upost)), dtype=torch.float32, device=dev)
        self.group_members = [np.flatnonzero(group_of[self.upost] == k) for k in range(len(self.names))]
        self.dn = np.flatnonzero(group_of >= 0)          # the instructed DNs themselves
        if not np.isin(self.dn, self.upost).all():
            raise ValueError("every instructed DN needs plastic input synapses")
        self.d_dn_group = T(group_of[self.dn], np.int32)
        self.d_dn_rate_k = T(np.searchsorted(self.upost, self.dn), np.int32)
        self.dn_slots = brain.slots(self.dn)
        self.d_dn_slot = self.dn_slots.to(torch.int32)
        self.ib = torch.zeros((self.B, len(self.dn)), dtype=torch.float32, device=dev) if eta_bias > 0 else None
        self._tgt = torch.zeros((self.B, len(self.names)), dtype=torch.float32, device=dev)
        self._learn = torch.zeros(self.B, dtype=torch.uint8, device=dev)
        self._coef = torch.zeros(self.B, dtype=torch.float32, device=dev)
        self._ratio = torch.ones(self.B, dtype=torch.float32, device=dev)
        self._slow = None

    def _t(self, a, dt):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(a).astype(dt))).to(self.dev)

    def _stream(self):
        return _lib.stream_handle(self.dev)

    def _pw(self):
        pw = self.brain.pw
        if pw.dtype != torch.float32 or not pw.is_contiguous():
            raise ValueError("brain.pw must be contiguous float32")
        return pw

    def reset(self, slots):
        s = torch.as_tensor(np.atleast_1d(slots), device=self.dev)
        self.trace[s] = 0
        self.rate[s] = 0

    def _update(self, mode, sel, n_sel, pre_k, coef, group=None, post_k=None, mask=None, post_l=None, delta=None,
                n_l=0):
        P = _lib.ptr
        a = _lib.FzInstructUpdateArgs(mode, self.n_flies, n_sel, self._pw().shape[1], self.trace.shape[1],
                                      self.rate.shape[1], len(self.names), n_l,
                                      P(sel) if sel is not None else None, P(pre_k),
                                      P(group) if group is not None else None, P(post_k) if post_k is not None else None,
                                      P(mask) if mask is not None else None, P(post_l) if post_l is not None else None,
                                      P(self.d_fly_ptr), P(self.d_fly_slots), P(self._learn), P(coef), P(self.trace),
                                      P(self.rate), P(self._tgt), P(delta) if delta is not None else None,
                                      P(self.sign), P(self.wmax), P(self._pw()))
        _lib.check(_lib.load().fz_instruct_update(ctypes.byref(a), self._stream()), "fz_instruct_update")

    def step(self, counts: torch.Tensor, targets, learn, window_ms: float, eta=None, report: bool = True):
        """One frame (flyzero/batch.py:283-318).  ``counts``: the window's (B, n) int32 device
        counts; ``targets``: (slots, len(GROUPS)) rates in ``GROUPS`` order or None; ``learn``:
        bool per slot; ``eta``: scalar or per-slot rate (default ``p.eta``).  Returns the mean
        |target - rate| (Hz) per slot, 0 where not learning (a host copy, which waits for the
        device; ``report=False`` returns None instead)."""
        p, brain = self.p, self.brain
        if counts.dtype != torch.int32 or not counts.is_contiguous() or tuple(counts.shape) != (self.B, brain.n):
            raise ValueError("counts must be the window's contiguous int32 (B, n) device counts")
        a_pre = np.float32(1 - np.exp(-window_ms / p.tau_pre_ms))
        a_rate = np.float32(1 - np.exp(-window_ms / p.tau_rate_ms))
        hz = np.float32(1000.0 / window_ms)
        P = _lib.ptr
        _lib.check(_lib.load().fz_instruct_traces(
            self.B, brain.n, counts.data_ptr(), self.trace.shape[1], P(self.d_upre), P(self.trace), a_pre,
            np.float32(p.tau_pre_ms / 1000.0), self.rate.shape[1], P(self.d_upost), P(self.rate), a_rate, hz,
            self._stream()), "fz_instruct_traces")
        learn = np.broadcast_to(np.asarray(learn, bool), (self.B,))
        if targets is None or not learn.any():
            return np.zeros(self.B)
        tg = np.asarray(targets, np.float32).reshape(self.B, len(self.names))
        self._tgt.copy_(torch.from_numpy(np.ascontiguousarray(tg)))
        self._learn.copy_(torch.from_numpy(learn.astype(np.uint8)))
        eta_b = np.broadcast_to(np.asarray(p.eta if eta is None else eta, np.float64).ravel(), (self.B,))
        self._learn_step(learn, eta_b)
        if self.ib is not None:
            ratio = eta_b / p.eta if eta is not None else np.ones(self.B)
            self._ratio.copy_(torch.from_numpy(ratio.astype(np.float32)))
            _lib.check(_lib.load().fz_instruct_bias(
                self.n_flies, len(self.dn), P(self.d_fly_ptr), P(self.d_fly_slots), P(self._learn),
                P(self.d_dn_group), P(self.d_dn_rate_k), P(self._tgt), len(self.names), P(self.rate),
                self.rate.shape[1], np.float32(self.eta_bias), P(self._ratio), np.float32(self.bias_limit),
                P(self.ib), P(brain.bias_ext), brain.bias_ext.shape[1], P(self.d_dn_slot), self._stream()),
                "fz_instruct_bias")
        if not report:
            return None
        rate = self.rate.cpu().numpy()
        out = np.zeros(self.B)
        for k, mem in enumerate(self.group_members):
            if len(mem):
                out += np.abs(tg[:, k] - rate[:, mem].mean(1))
        return np.where(learn, out / len(self.names), 0.0)

    def _learn_step(self, learn, eta_b):
        """The delta rule on every plastic synapse onto an instructed group (batch.py:320-322)."""
        self._coef.copy_(torch.from_numpy(eta_b.astype(np.float32)))
        self._update(0, None, self._pw().shape[1], self.d_pre_k, self._coef, group=self.d_group,
                     post_k=self.d_post_k, mask=self.d_mask)

    # consolidation: a slow copy of the plastic weights and DN biases follows the fast ones
    # (flyzero/batch.py:174-195)
    def consolidate(self, rate: float):
        if self._slow is None:
            self._slow = (self.brain.pw.clone(), self.brain.bias_ext.clone())
        self._slow[0].add_(np.float32(rate) * (self.brain.pw - self._slow[0]))
        self._slow[1].add_(np.float32(rate) * (self.brain.bias_ext - self._slow[1]))

    def swap_slow(self):
        """Exchange the fast and the consolidated weights (call again to swap back)."""
        if self._slow is None:
            return
        pw, bias = self._slow
        fast = (self.brain.pw.clone(), self.brain.bias_ext.clone())
        self.brain.pw.copy_(pw)
        self.brain.bias_ext.copy_(bias)
        self._slow = fast
        if self.ib is not None:
            self.ib += (bias - fast[1])[:, self.dn_slots]

    def drift(self) -> np.ndarray:
        """Mean relative change from the initial weights, per slot (flyzero/batch.py:221-223)."""
        w0 = torch.from_numpy(np.asarray(self.brain.w0, np.float32)).to(self.dev)
        return (torch.abs(self.brain.pw - w0) / self.wmax).mean(1).cpu().numpy()


class BatchInstructDeep(BatchInstruct):
    """Instructed learning one layer deeper (flyzero/batch.py:336-403): the DN synapses learn as
    in ``BatchInstruct``; each L1 neuron gets ``delta_i = sum_j n_ij err_j`` through its current
    synapse counts onto the DNs, and the synapses onto it learn ``eta_deep * trace * delta``.
    Needs a brain whose plastic positions include ``deep_positions``."""

    def __init__(self, brain, conn, params: InstructParams | None = None, fly_of_slot=None,
                 eta_deep: float = 1e-6, deep_scale=None, eta_bias: float = 0.0):
        super().__init__(brain, conn, params, fly_of_slot, eta_bias=eta_bias)
        self.eta_deep = eta_deep
        self.deep_scale = np.ones(self.B) if deep_scale is None else np.asarray(deep_scale, float)
        pre, post = np.asarray(brain.pl_pre), np.asarray(brain.pl_post)
        is_dn = np.zeros(conn.n, bool)
        for g in self.names:
            is_dn[conn.find(*GROUPS[g])] = True
        dn_syn = self.mask_np
        self.l1 = np.unique(pre[dn_syn])
        l1_index = np.full(conn.n, -1, np.int64)
        l1_index[self.l1] = np.arange(len(self.l1))
        self.dn_sel_np = np.flatnonzero(dn_syn)                          # L1 -> DN synapses
        self.dn_pre_l1_np = l1_index[pre[dn_syn]]
        deep = (l1_index[post] >= 0) & ~is_dn[post]
        self.deep_sel_np = np.flatnonzero(deep)                           # -> L1 synapses
        self.deep_post_l1_np = l1_index[post[deep]]
        self.n_deep = int(deep.sum())
        T = self._t
        self.d_dsyn = T(self.dn_sel_np, np.int32)
        self.d_dsyn_pre_k = T(self.pre_k_np[self.dn_sel_np], np.int32)
        self.d_dsyn_group = T(self.post_group_np[self.dn_sel_np], np.int32)
        self.d_dsyn_post_k = T(self.post_k_np[self.dn_sel_np], np.int32)
        # the L1 neurons' DN synapses as a CSR (ascending synapse order within each L1 neuron)
        order = np.argsort(self.dn_pre_l1_np, kind="stable")
        self.l_ptr = np.concatenate([[0], np.cumsum(np.bincount(self.dn_pre_l1_np, minlength=len(self.l1)))])
        sel_o = self.dn_sel_np[order]
        self.d_l_ptr = T(self.l_ptr, np.int32)
        self.d_l_syn = T(sel_o, np.int32)
        self.d_l_group = T(self.post_group_np[sel_o], np.int32)
        self.d_l_post_k = T(self.post_k_np[sel_o], np.int32)
        self.d_deep_sel = T(self.deep_sel_np, np.int32)
        self.d_deep_pre_k = T(self.pre_k_np[self.deep_sel_np], np.int32)
        self.d_deep_post_l1 = T(self.deep_post_l1_np, np.int32)
        self.delta = torch.zeros((self.B, len(self.l1)), dtype=torch.float32, device=self.dev)
        self._dcoef = torch.zeros(self.B, dtype=torch.float32, device=self.dev)

    def _deep_coef(self, learn, eta_b, eta_layer):
        """coef = per-slot rate / learning slots of its fly, 0 where not learning, the rate taken
        through float32 first (flyzero/batch.py:394-397)."""
        rate = eta_b.astype(np.float32).astype(np.float64) * (eta_layer / self.p.eta) * self.deep_scale
        n_act = np.bincount(self.fly_of_slot, weights=learn, minlength=self.n_flies)
        return np.where(learn, rate / np.maximum(n_act[self.fly_of_slot], 1), 0.0).astype(np.float32)

    def _learn_step(self, learn, eta_b):
        P = _lib.ptr
        # retrograde error from the weights before this frame's DN update
        _lib.check(_lib.load().fz_instruct_delta(
            self.B, len(self.l1), P(self.d_l_ptr), P(self.d_l_syn), P(self.d_l_group), P(self.d_l_post_k),
            P(self._tgt), len(self.names), P(self.rate), self.rate.shape[1], P(self._pw()), self._pw().shape[1],
            np.float32(self.brain.p.w_syn), P(self.delta), self._stream()), "fz_instruct_delta")
        self._coef.copy_(torch.from_numpy(eta_b.astype(np.float32)))
        self._update(0, self.d_dsyn, len(self.dn_sel_np), self.d_dsyn_pre_k, self._coef, group=self.d_dsyn_group,
                     post_k=self.d_dsyn_post_k)
        self._dcoef.copy_(torch.from_numpy(self._deep_coef(learn, eta_b, self.eta_deep)))
        self._update(1, self.d_deep_sel, self.n_deep, self.d_deep_pre_k, self._dcoef, post_l=self.d_deep_post_l1,
                     delta=self.delta, n_l=len(self.l1))
        self._deeper(learn, eta_b)

    def _deeper(self, learn, eta_b):
        """Hook for a further layer (``BatchInstructDeep2``)."""


class BatchInstructDeep2(BatchInstructDeep):
    """``BatchInstructDeep`` plus one more layer (flyzero/batch.py:424-451): synapses onto L2
    learn from the L1 errors passed back through the fly's own L2 -> L1 synapses (their weights
    after this frame's L1 update): ``delta2_k = sum_i n_ki delta1_i``,
    ``dw = eta_deep2 * trace * delta2``."""

    def __init__(self, brain, conn, params: InstructParams | None = None, fly_of_slot=None,
                 eta_deep: float = 1e-6, deep_scale=None, eta_bias: float = 0.0, eta_deep2: float = 1e-8):
        super().__init__(brain, conn, params, fly_of_slot, eta_deep, deep_scale, eta_bias)
        self.eta_deep2 = eta_deep2
        pre, post = np.asarray(brain.pl_pre), np.asarray(brain.pl_post)
        l1_index = np.full(conn.n, -1, np.int64)
        l1_index[self.l1] = np.arange(len(self.l1))
        deep_sel = self.deep_sel_np
        self.l2 = np.unique(pre[deep_sel])
        l2_index = np.full(conn.n, -1, np.int64)
        l2_index[self.l2] = np.arange(len(self.l2))
        back_pre = l2_index[pre[deep_sel]]                               # L2 -> L1 synapses
        is_sel = np.zeros(len(post), bool)
        is_sel[self.dn_sel_np] = True
        is_sel[deep_sel] = True
        sel2 = np.flatnonzero((l2_index[post] >= 0) & ~is_sel)
        self.n_deep2 = len(sel2)
        T = self._t
        # delta2 is the delta kernel over each L2 neuron's synapses onto L1 (a CSR in ascending
        # synapse order) with delta1 in the role of the target table (column = the L1 neuron)
        # and a zero rate, so that its "error" is delta1 itself
        order = np.argsort(back_pre, kind="stable")
        self.b_ptr = np.concatenate([[0], np.cumsum(np.bincount(back_pre, minlength=len(self.l2)))])
        self.d_b_ptr = T(self.b_ptr, np.int32)
        self.d_b_syn = T(deep_sel[order], np.int32)
        self.d_b_l1 = T(self.deep_post_l1_np[order], np.int32)
        self.d_k2_sel = T(sel2, np.int32)
        self.d_k2_pre = T(self.pre_k_np[sel2], np.int32)
        self.d_k2_post = T(l2_index[post[sel2]], np.int32)
        self._zero_rate = torch.zeros((self.B, 1), dtype=torch.float32, device=self.dev)
        self._zero_k = torch.zeros(max(len(deep_sel), 1), dtype=torch.int32, device=self.dev)
        self.delta2 = torch.zeros((self.B, len(self.l2)), dtype=torch.float32, device=self.dev)
        self._d2coef = torch.zeros(self.B, dtype=torch.float32, device=self.dev)

    def _deeper(self, learn, eta_b):
        P = _lib.ptr
        # delta2[b][k] = sum over k's synapses q onto L1 of (pw[b][syn_q] / w_syn) * delta1[b][l1_q],
        # with pw after this frame's L1 update (flyzero/batch.py:437-440)
        _lib.check(_lib.load().fz_instruct_delta(
            self.B, len(self.l2), P(self.d_b_ptr), P(self.d_b_syn), P(self.d_b_l1), P(self._zero_k),
            P(self.delta), len(self.l1), P(self._zero_rate), 1, P(self._pw()), self._pw().shape[1],
            np.float32(self.brain.p.w_syn), P(self.delta2), self._stream()), "fz_instruct_delta")
        self._d2coef.copy_(torch.from_numpy(self._deep_coef(learn, eta_b, self.eta_deep2)))
        self._update(1, self.d_k2_sel, self.n_deep2, self.d_k2_pre, self._d2coef, post_l=self.d_k2_post,
                     delta=self.delta2, n_l=len(self.l2))
