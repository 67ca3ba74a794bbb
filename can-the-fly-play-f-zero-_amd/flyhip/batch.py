"""Batched learning rule and readout on a ``BatchBrain``: drop-ins for flyzero/batch.py.

* ``BatchReward``: the exploratory-Hebbian reward rule (flyzero/learning.py:89-131, batched as
  flyzero/batch.py:454-495 with the per-fly averaging of batch.py:141-172), one fused HIP kernel
  per frame (``fz_reward_step``) that reads the window's spike counts on the device.  The rule
  runs in float64 with weights in synapse units; the simulation uses the integer synapse count
  ``rint(w)`` (``pw = count * w_syn``), so the counts the brain sees change in whole synapses.
* ``DNReadout``: smoothed descending-neuron rates for the 8 groups of flyzero/instruct.py:43-48
  (the rate half of ``BatchMotor.update``, flyzero/batch.py:107-111), on the device.
* ``BatchMotor``: the whole readout (flyzero/batch.py:77-138): the DN rates of the 9 READOUT
  groups on the device, then button dicts (with the fleet's sigma-delta taps) on the host.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib


@dataclass
class LearnParams:
    """The rule's fields of flyzero/learning.py:36-53 (same names and defaults), plus
    ``w_abs_max``: an absolute cap in synapses (inf: only the reference's cap applies)."""
    eta: float = 2e-4             # mV per unit of (reward x eligibility)
    tau_elig_ms: float = 500.0
    tau_post_ms: float = 500.0
    tau_reward_ms: float = 1000.0
    cap: float = 3.0              # |w| <= cap * max(|w0|, one synapse)
    noise_hz: float = 20.0
    noise_tau_ms: float = 300.0
    targets: tuple = ("DNa02", "DNa01", "DNg02*", "DNp09", "MDN")
    w_abs_max: float = float("inf")


class FzRewardArgs(ctypes.Structure):
    _fields_ = [("B", ctypes.c_int), ("n", ctypes.c_int), ("n_pl", ctypes.c_int), ("n_post", ctypes.c_int),
                ("n_flies", ctypes.c_int), ("any_learn", ctypes.c_int)] + \
               [(f, ctypes.c_void_p) for f in ("counts", "pre", "post", "post_k", "upost", "sign", "wmax", "elig",
                                               "w", "count", "pw", "post_mean", "r_mean", "reward", "learn",
                                               "fly_ptr", "fly_slots")] + \
               [("ae", ctypes.c_double), ("ap", ctypes.c_double), ("ar", ctypes.c_double), ("eta", ctypes.c_double),
                ("w_syn", ctypes.c_float), ("adv_out", ctypes.c_void_p), ("eta_slot", ctypes.c_void_p)]


def rule_constants(p: LearnParams, window_ms: float, w_syn: float) -> dict:
    """flyzero/batch.py:484-486; eta converted from mV to synapses."""
    return {"ae": float(np.exp(-window_ms / p.tau_elig_ms)), "ap": float(1 - np.exp(-window_ms / p.tau_post_ms)),
            "ar": float(1 - np.exp(-window_ms / p.tau_reward_ms)),
            "eta": float(p.eta) / float(np.float32(w_syn))}


def fly_csr(fly_of_slot: np.ndarray):
    fos = np.asarray(fly_of_slot, np.int64)
    n_flies = int(fos.max()) + 1
    order = np.argsort(fos, kind="stable")            # ascending slots within each fly
    ptr = np.concatenate([[0], np.cumsum(np.bincount(fos, minlength=n_flies))]).astype(np.int32)
    return n_flies, ptr, order.astype(np.int32)


class BatchReward:
    """Exploratory-Hebbian reward learning on the brain's plastic synapses
    (flyzero/batch.py:454-495; the rule of flyzero/learning.py:116-131)."""

    def __init__(self, brain, conn, params: LearnParams | None = None, fly_of_slot=None, seed: int = 0):
        self.p = p = params or LearnParams()
        self.brain = brain
        self.B = brain.B
        dev = brain.device
        self.rng = np.random.default_rng(seed)
        self.targets = np.unique(np.concatenate([conn.find(t) for t in p.targets]))
        self.upre, pre_inv = np.unique(brain.pl_pre, return_inverse=True)
        self.upost, post_inv = np.unique(brain.pl_post, return_inverse=True)
        self.pre_k, self.post_k_np = pre_inv, post_inv
        w0 = np.asarray(brain.w0_syn, np.float64)
        self.sign_np = np.sign(w0).astype(np.int8)
        self.wmax_np = np.minimum(p.cap * np.maximum(np.abs(w0), 1.0), p.w_abs_max)
        fos = np.arange(self.B) if fly_of_slot is None else np.asarray(fly_of_slot)
        self.fly_of_slot = fos
        self.n_flies, fptr, fslots = fly_csr(fos)
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(dev)  # noqa: E731
        self.d_pre = t(brain.pl_pre, np.int32)
        self.d_post = t(brain.pl_post, np.int32)
        self.d_post_k = t(post_inv, np.int32)
        self.d_upost = t(self.upost, np.int32)
        self.d_sign = t(self.sign_np, np.int8)
        self.d_wmax = t(self.wmax_np, np.float64)
        self.d_fly_ptr = t(fptr, np.int32)
        self.d_fly_slots = t(fslots, np.int32)
        n_pl = len(brain.plastic_pos)
        self.elig = torch.zeros((self.B, n_pl), dtype=torch.float32, device=dev)
        self.w = t(np.tile(w0, (self.B, 1)), np.float64)
        self.count = torch.round(self.w).to(torch.int32)
        self.post_mean = torch.zeros((self.B, len(self.upost)), dtype=torch.float64, device=dev)
        self.r_mean = torch.zeros(self.B, dtype=torch.float64, device=dev)
        self.noise = np.zeros((self.B, len(self.targets)))
        self._reward = torch.zeros(self.B, dtype=torch.float64, device=dev)
        self._learn = torch.zeros(self.B, dtype=torch.uint8, device=dev)
        self._adv = torch.zeros(self.B, dtype=torch.float64, device=dev)
        self._eta_slot = torch.zeros(self.B, dtype=torch.float64, device=dev)

    def reset(self, slots):
        s = np.atleast_1d(slots)
        self.elig[torch.as_tensor(s, device=self.brain.device)] = 0
        self.noise[s] = 0

    def exploration(self, window_ms: float) -> np.ndarray:
        """Extra Poisson rates (slots, len(self.targets)) for the target neurons (batch.py:476-480)."""
        a = window_ms / self.p.noise_tau_ms
        self.noise += -a * self.noise + np.sqrt(2 * a) * self.rng.standard_normal(self.noise.shape)
        return np.maximum(self.p.noise_hz * self.noise, 0.0)

    def step(self, counts: torch.Tensor, reward, learn, window_ms: float, return_adv: bool = True,
             eta_slot=None):
        """One frame: traces, means and (if any slot learns) weights.  ``counts``: the window's
        (B, n) int32 device counts; ``reward``, ``learn``: per slot; ``eta_slot``: None (every slot
        learns at ``p.eta``) or a per-slot learning rate in mV per unit, like ``p.eta``
        (flyzero/school.py:289-304).  Returns the advantage (a host copy, which waits for the
        device; ``return_adv=False`` returns None instead)."""
        brain = self.brain
        if counts.dtype != torch.int32 or not counts.is_contiguous() or tuple(counts.shape) != (self.B, brain.n):
            raise ValueError("counts must be the window's contiguous int32 (B, n) device counts")
        learn = np.broadcast_to(np.asarray(learn, bool), (self.B,))
        if isinstance(reward, torch.Tensor):  # device rewards stay on the device
            self._reward.copy_(reward.to(self._reward.device, torch.float64).expand(self.B))
        else:
            self._reward.copy_(torch.from_numpy(np.broadcast_to(np.asarray(reward, np.float64), (self.B,)).copy()))
        self._learn.copy_(torch.from_numpy(learn.astype(np.uint8)))
        c = rule_constants(self.p, window_ms, brain.p.w_syn)
        if eta_slot is not None:   # mV -> synapses per unit, as rule_constants does for p.eta
            e = np.broadcast_to(np.asarray(eta_slot, np.float64), (self.B,)) / float(np.float32(brain.p.w_syn))
            self._eta_slot.copy_(torch.from_numpy(np.ascontiguousarray(e)))
        if not brain.pw.is_contiguous() or brain.pw.dtype != torch.float32:
            raise ValueError("brain.pw must be contiguous float32")
        P = _lib.ptr
        a = FzRewardArgs(self.B, brain.n, len(brain.plastic_pos), len(self.upost), self.n_flies, int(learn.any()),
                         counts.data_ptr(), P(self.d_pre), P(self.d_post), P(self.d_post_k), P(self.d_upost),
                         P(self.d_sign), P(self.d_wmax), P(self.elig), P(self.w), P(self.count), P(brain.pw),
                         P(self.post_mean), P(self.r_mean), P(self._reward), P(self._learn), P(self.d_fly_ptr),
                         P(self.d_fly_slots), c["ae"], c["ap"], c["ar"], c["eta"], np.float32(brain.p.w_syn),
                         P(self._adv), P(self._eta_slot) if eta_slot is not None else None)
        _lib.check(_lib.load().fz_reward_step(ctypes.byref(a), _lib.stream_handle(brain.device)), "fz_reward_step")
        return self._adv.cpu().numpy() if return_adv else None

    def drift(self) -> np.ndarray:
        """Mean relative change from the initial weights, per slot (flyzero/batch.py:221-223)."""
        w0 = torch.from_numpy(np.asarray(self.brain.w0_syn, np.float64)).to(self.w.device)
        return (torch.abs(self.w - w0) / self.d_wmax).mean(1).cpu().numpy()


class DNReadout:
    """Smoothed rates (Hz) of the descending-neuron groups (flyzero/batch.py:107-111), on the GPU."""

    def __init__(self, conn, batch: int, groups: dict | None = None, tau_ms: float = 80.0, device=None):
        from .connectome import GROUPS

        self.groups = dict(GROUPS if groups is None else groups)
        self.names = list(self.groups)
        members = [np.asarray(conn.find(t, s), np.int64) for t, s in self.groups.values()]
        ptr = np.concatenate([[0], np.cumsum([len(m) for m in members])]).astype(np.int32)
        dev = torch.device(device if device is not None else "cuda")
        self.d_ptr = torch.from_numpy(ptr).to(dev)
        self.d_idx = torch.from_numpy(np.concatenate(members).astype(np.int32) if len(members) else
                                      np.zeros(0, np.int32)).to(dev)
        self.B, self.tau_ms, self.n = batch, tau_ms, conn.n
        self.rates = torch.zeros((batch, len(self.names)), dtype=torch.float32, device=dev)

    def reset(self, slots=None):
        if slots is None:
            self.rates.zero_()
        else:
            self.rates[torch.as_tensor(np.atleast_1d(slots), device=self.rates.device)] = 0

    def update(self, counts: torch.Tensor, window_ms: float) -> torch.Tensor:
        alpha = np.float32(1.0 - np.exp(-window_ms / self.tau_ms))
        hz = np.float32(1000.0 / window_ms)
        _lib.check(_lib.load().fz_dn_rates(self.B, self.n, counts.data_ptr(), len(self.names), _lib.ptr(self.d_ptr),
                                           _lib.ptr(self.d_idx), hz, alpha, self.rates.data_ptr(),
                                           _lib.stream_handle(self.rates.device)), "fz_dn_rates")
        return self.rates


BUTTONS = ["B", "Y", "SELECT", "START", "UP", "DOWN", "LEFT", "RIGHT", "A", "X", "L", "R"]
READOUT = ["steer_left", "steer_right", "lean_left", "lean_right", "wing_left", "wing_right", "accelerate", "brake",
           "boost"]


@dataclass
class MotorParams:
    """flyzero/motor.py:39-66 (same names and defaults)."""
    tau_ms: float = 80.0
    steer_threshold: float = 10.0
    steer_bias: float = 0.0
    lean_threshold: float = 10.0
    accel_threshold: float = 5.0
    brake_threshold: float = 15.0
    boost_threshold: float = 1.0
    taps: bool = False
    steer_span: float = 150.0
    lean_span: float = 70.0
    groups: dict = None

    def __post_init__(self):
        if self.groups is None:
            self.groups = {"steer_left": ("DNa02", "left"), "steer_right": ("DNa02", "right"),
                           "lean_left": ("DNa01", "left"), "lean_right": ("DNa01", "right"),
                           "wing_left": ("DNg02*", "left"), "wing_right": ("DNg02*", "right"),
                           "accelerate": ("DNp09", None), "brake": ("MDN", None), "boost": ("DNp01", None)}


class BatchMotor:
    """``MotorReadout`` for every slot (flyzero/batch.py:77-138): smoothed DN rates on the device
    (the ``dn_rates`` kernel over the READOUT groups), then button dicts on the host, with the
    sigma-delta taps of the GPU fleet when ``params.taps``."""

    def __init__(self, conn, batch: int, params: MotorParams | None = None, device=None):
        self.p = p = params or MotorParams()
        self.B = batch
        self.readout = DNReadout(conn, batch, groups={k: p.groups[k] for k in READOUT}, tau_ms=p.tau_ms,
                                 device=device)
        self.size = np.array([len(conn.find(*p.groups[k])) for k in READOUT], np.float32)
        self.rates = np.zeros((batch, len(READOUT)), np.float32)
        self.acc = np.zeros((batch, 2))   # sigma-delta accumulators for taps (steer, lean)

    def reset(self, slots=None):
        self.readout.reset(slots)
        if slots is None:
            self.rates[:] = 0
            self.acc[:] = 0
        else:
            self.rates[np.atleast_1d(slots)] = 0
            self.acc[np.atleast_1d(slots)] = 0

    def update(self, counts: torch.Tensor, window_ms: float) -> list:
        """Rates from the window's (B, n) device counts, then one button dict per slot (a host copy)."""
        self.rates = self.readout.update(counts, window_ms).cpu().numpy()
        return [dict(zip(BUTTONS, row)) for row in self.masks(self.rates).tolist()]

    def _taps(self, rows, j, x, threshold, span):
        """+1 (left), -1 (right) or 0 this frame per slot; duty = (|x| - threshold) / span into a
        sigma-delta accumulator (flyzero/batch.py:114-120)."""
        duty = np.sign(x) * np.clip((np.abs(x) - threshold) / span, 0.0, 1.0)
        acc = self.acc[rows, j] + duty
        t = np.where(acc >= 0.5, 1, np.where(acc <= -0.5, -1, 0))
        self.acc[rows, j] = acc - t
        return t

    def masks(self, r, rows=None) -> np.ndarray:
        """Pressed buttons (m, len(BUTTONS)) for the readout rates ``r`` (m, 9) of slots ``rows``
        (default 0 .. m-1), all slots at once (flyzero/batch.py:122-138)."""
        p = self.p
        r = np.asarray(r)
        rows = np.arange(len(r)) if rows is None else np.atleast_1d(rows)
        steer = (r[:, 0] - r[:, 1]) + (r[:, 5] - r[:, 4]) + p.steer_bias
        lean = r[:, 2] - r[:, 3]
        m = np.zeros((len(r), len(BUTTONS)), bool)
        col = {k: i for i, k in enumerate(BUTTONS)}
        if p.taps:
            s = self._taps(rows, 0, steer, p.steer_threshold, p.steer_span)
            le = self._taps(rows, 1, lean, p.lean_threshold, p.lean_span)
            m[:, col["LEFT"]], m[:, col["RIGHT"]], m[:, col["L"]], m[:, col["R"]] = s > 0, s < 0, le > 0, le < 0
        else:
            m[:, col["LEFT"]] = steer > p.steer_threshold
            m[:, col["RIGHT"]] = steer < -p.steer_threshold
            m[:, col["L"]] = lean > p.lean_threshold
            m[:, col["R"]] = lean < -p.lean_threshold
        m[:, col["B"]] = r[:, 6] > p.accel_threshold
        m[:, col["Y"]] = r[:, 7] > p.brake_threshold
        m[:, col["A"]] = r[:, 8] > p.boost_threshold
        return m

    def buttons(self, r, k: int = 0) -> dict:
        """One slot's button dict (flyzero/batch.py:122-138)."""
        return dict(zip(BUTTONS, self.masks(np.asarray(r)[None], [k])[0].tolist()))


def step_scaled(rp: BatchReward, counts: torch.Tensor, reward, learn, window_ms: float, eta_slot):
    """``BatchReward.step`` with a per-slot learning rate (flyzero/school.py:289-304,
    ``_step_scaled``: one rule shared by several flies that learn at different rates).
    ``eta_slot``: (B,) in mV per unit of (reward x eligibility), like ``LearnParams.eta``."""
    return rp.step(counts, reward, np.asarray(learn, bool), window_ms, eta_slot=eta_slot)
