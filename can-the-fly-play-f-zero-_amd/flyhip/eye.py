"""Batched motion-vision front end on the GPU: a drop-in for ``MotionEye`` (flyzero/vision.py:64-191)
for many flies at once.

The reference runs one ``MotionEye`` per emulator worker on the CPU (flyzero/pool.py:41-44) and
copies the rates into the brain's input probabilities every frame (flyzero/fleet.py:76-80).
Here every fly's frame goes through one kernel (``fz_eye_step``): grid luminance, the ON / OFF
split against a running low-pass, Hassenstein-Reichardt correlators and the T4/T5 rates, which
can be written straight into ``BatchBrain.p_ext``.  The default ``MotionParams`` (motion
channels only) are supported; the optional medulla and colour channels are not.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib


@dataclass
class MotionParams:
    """flyzero/vision.py:37-56 (same names and defaults)."""
    max_rate: float = 150.0   # Hz
    gain: float = 1000.0      # correlator output -> fraction of max rate
    cell: int = 4             # pixels per sampling point
    tau_hp: float = 4.0       # frames, high-pass of the ON/OFF split
    tau_delay: float = 2.0    # frames, low-pass "delay" arm of the correlator
    shift: int = 1            # sampling points between correlator arms
    overlap: float = 0.1      # binocular overlap (fraction of screen width)
    y_min: float = 0.0        # ignore the top of the screen (fraction of height)
    transient_gain: float = 0.0
    sustained_gain: float = 0.0
    surround: int = 3
    chromatic_gain: float = 0.0


# subtype -> direction in eye coordinates ("front" = towards the midline), flyzero/vision.py:60
SUBTYPES = {"a": "front_to_back", "b": "back_to_front", "c": "up", "d": "down"}
POLARITIES = {"T4": "on", "T5": "off"}
DIRECTIONS = ("right", "left", "down", "up")   # kernel key = polarity * 4 + direction


def screen_direction(direction: str, side: str) -> str:
    """Screen direction of a T4/T5 subtype's preferred motion for an eye (flyzero/vision.py:112-121):
    front-to-back is leftwards for the left eye."""
    if direction == "front_to_back":
        return "left" if side == "left" else "right"
    if direction == "back_to_front":
        return "right" if side == "left" else "left"
    return direction


def field_to_screen(u, v, side, w: int, h: int, overlap: float):
    """Screen pixel of visual-field coordinates u (0 medial .. 1 lateral), v (0 dorsal .. 1) for an
    eye (flyzero/eyes.py:61-69): the left eye covers [0, 0.5 + o] of the width, the right eye
    [0.5 - o, 1], lateral at the outer edge."""
    o = overlap / 2
    left = np.asarray(side) == "left"
    sx = np.where(left, (1 - u) * (0.5 + o), 0.5 - o + u * (0.5 + o))
    px = np.clip((sx * w).astype(int), 0, w - 1)
    py = np.clip((np.asarray(v) * h).astype(int), 0, h - 1)
    return px, py


def synthetic_motion_cells(fly, frame_shape, overlap: float = 0.1, seed: int = 9):
    """T4/T5 cells of the synthetic fly: its visual input neurons, each given a subtype (T4/T5 x
    a-d) from a seeded draw and an eye from its retinotopic x (x < 0.5: left); the visual-field
    coordinates are u = |2 x - 1| (0 medial .. 1 lateral) and v = y, standing in for the principal
    axes that flyzero/eyes.py:72-88 fits to FlyWire neuron positions.  Returns (idx, px, py,
    side, kind) like the reference's retinotopy plus the cell type."""
    h, w = frame_shape
    rng = np.random.default_rng(seed)
    idx = np.asarray(fly.input_idx, np.int64)
    xy = np.asarray(fly.input_xy, np.float64)
    kinds = np.array([k + s for k in POLARITIES for s in SUBTYPES])
    kind = kinds[rng.integers(0, len(kinds), len(idx))]
    side = np.where(xy[:, 0] < 0.5, "left", "right")
    px, py = field_to_screen(np.abs(2 * xy[:, 0] - 1), xy[:, 1], side, w, h, overlap)
    return idx, px, py, side, kind


class BatchMotionEye:
    """``MotionEye`` for B flies: ``idx`` (the T4/T5 neurons), ``see(frames) -> rates``."""

    def __init__(self, cells, frame_shape, batch: int, params: MotionParams | None = None, device=None):
        """``cells`` = (idx, px, py, side, kind) as ``synthetic_motion_cells`` returns (kind: T4a .. T5d)."""
        self.p = p = params or MotionParams()
        if p.transient_gain or p.sustained_gain or p.chromatic_gain:
            raise ValueError("the medulla and colour channels are not supported (MotionParams defaults only)")
        self.h, self.w = frame_shape
        self.gh, self.gw = self.h // p.cell, self.w // p.cell
        self.B = batch
        self.dev = dev = torch.device(device if device is not None else "cuda")
        idx, px, py, side, kind = cells
        self.idx = np.asarray(idx, np.int64)
        self.px, self.py = np.asarray(px, np.int64), np.asarray(py, np.int64)
        self.side = np.asarray(side)
        key = np.empty(len(self.idx), np.int32)
        for n, (k, s) in enumerate(zip(kind, self.side)):
            pol = 0 if POLARITIES[k[:2]] == "on" else 1
            key[n] = pol * 4 + DIRECTIONS.index(screen_direction(SUBTYPES[k[2]], s))
        self.key = key
        self.gx = np.clip(self.px // p.cell, 0, self.gw - 1).astype(np.int32)
        self.gy = np.clip(self.py // p.cell, 0, self.gh - 1).astype(np.int32)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a).astype(np.int32)).to(dev)  # noqa: E731
        self.d_key, self.d_gx, self.d_gy, self.d_py = t(self.key), t(self.gx), t(self.gy), t(self.py)
        G = (batch, self.gh, self.gw)
        self.lp = torch.zeros(G, dtype=torch.float32, device=dev)
        self.d_on = torch.zeros(G, dtype=torch.float32, device=dev)
        self.d_off = torch.zeros(G, dtype=torch.float32, device=dev)
        self.first = torch.ones(batch, dtype=torch.uint8, device=dev)
        self.rates = torch.zeros((batch, len(self.idx)), dtype=torch.float32, device=dev)
        self._slot = None

    @classmethod
    def for_fly(cls, fly, frame_shape, batch: int, params: MotionParams | None = None, device=None):
        p = params or MotionParams()
        return cls(synthetic_motion_cells(fly, frame_shape, p.overlap), frame_shape, batch, p, device)

    def reset(self, slots=None):
        """The given flies' filters restart from their next frame (a respawned emulator)."""
        if slots is None:
            self.first.fill_(1)
        else:
            self.first[torch.as_tensor(np.atleast_1d(slots), device=self.dev)] = 1

    def see(self, frames: torch.Tensor, brain=None) -> torch.Tensor:
        """One frame per fly: ``frames`` uint8 (B, h, w, 3) on the device.  Returns the (B, n) T4/T5
        rates (Hz); with ``brain``, also writes them into ``brain.p_ext`` as per-step
        probabilities (rate * dt / 1000, flyzero/fleet.py:76-80)."""
        if frames.dtype != torch.uint8 or tuple(frames.shape) != (self.B, self.h, self.w, 3):
            raise ValueError(f"frames must be uint8 ({self.B}, {self.h}, {self.w}, 3)")
        frames = frames.contiguous()
        p, P = self.p, _lib.ptr
        p_ext, n_ext, slot, scale = None, 0, None, 0.0
        if brain is not None:
            if self._slot is None or self._slot[0] is not brain:
                self._slot = (brain, brain.slots(self.idx).to(torch.int32))
            p_ext, n_ext, slot = P(brain.p_ext), brain.p_ext.shape[1], P(self._slot[1])
            scale = np.float32(brain.p.dt / 1000.0)
        a = _lib.FzEyeArgs(self.B, self.h, self.w, p.cell, self.gh, self.gw, p.shift, len(self.idx), P(self.first),
                           frames.data_ptr(), P(self.lp), P(self.d_on), P(self.d_off), P(self.d_key), P(self.d_gy),
                           P(self.d_gx), P(self.d_py), np.float32(p.y_min * self.h), np.float32(1.0 / p.tau_hp),
                           np.float32(1.0 / p.tau_delay), np.float32(p.gain), np.float32(p.max_rate), P(self.rates),
                           p_ext, n_ext, slot, scale)
        _lib.check(_lib.load().fz_eye_step(ctypes.byref(a), _lib.stream_handle(self.dev)), "fz_eye_step")
        self.first.zero_()
        return self.rates
