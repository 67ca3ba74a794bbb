"""Batched vision front end on the GPU: a drop-in for ``MotionEye`` (flyzero/vision.py:64-191)
for many flies at once.

The reference runs one ``MotionEye`` per emulator worker on the CPU (flyzero/pool.py:41-44) and
copies the rates into the brain's input probabilities every frame (flyzero/fleet.py:76-80).
Here every fly's frame goes through one kernel (``fz_eye_step``): grid luminance, the ON / OFF
split against a running low-pass, Hassenstein-Reichardt correlators for the T4/T5 cells, the
medulla transient / sustained-contrast channels and the R7/R8 colour channels, with the rates
written straight into ``BatchBrain.p_ext``.  Cells are placed on the screen by the retinotopy of
flyzero/eyes.py:37-88 (FlyWire) or by the synthetic fly's own layout.
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
# medulla columnar inputs and what they encode (flyzero/vision.py:64-71), colour photoreceptors
COLUMNAR = {"Mi1": "on_transient", "Tm3": "on_transient", "Tm1": "off_transient", "Tm2": "off_transient",
            "Tm4": "off_transient", "Mi4": "on_sustained", "Mi9": "off_sustained", "Tm9": "off_sustained"}
COLOUR = {"R7": "blue", "R8": "green"}
# kernel keys of the feature channels (FZ_EYE_* of include/flyhip.h)
FEATURE_KEYS = {"on_transient": 8, "off_transient": 9, "on_sustained": 10, "off_sustained": 11, "blue": 12,
                "green": 13}


def cell_key(kind: str, side: str) -> int:
    """Kernel key of a cell: T4a .. T5d -> polarity * 4 + screen direction; medulla and colour
    cell types -> their feature channel."""
    if kind in COLUMNAR:
        return FEATURE_KEYS[COLUMNAR[kind]]
    if kind in COLOUR:
        return FEATURE_KEYS[COLOUR[kind]]
    pol = 0 if POLARITIES[kind[:2]] == "on" else 1
    return pol * 4 + DIRECTIONS.index(screen_direction(SUBTYPES[kind[2]], side))


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


def _corr(a: np.ndarray, b: np.ndarray) -> float:
    """Pearson correlation; 0 when either side is constant."""
    da, db = a - a.mean(), b - b.mean()
    den = float(np.sqrt((da * da).sum() * (db * db).sum()))
    return float((da * db).sum()) / den if den > 0 else 0.0


def _unit(a: np.ndarray) -> np.ndarray:
    lo = a.min()
    return (a - lo) / max(float(a.max() - lo), 1e-9)


def eye_plane(xyz: np.ndarray, mid_x: float) -> np.ndarray:
    """Visual-field coordinates (u, v) in [0, 1]^2 of one eye's cells from their 3-D positions,
    the PCA retinotopy of flyzero/eyes.py:72-88: the eye's cells lie on a curved sheet whose two
    widest directions (the top two eigenvectors of the position covariance) span the visual
    field.  The direction whose coordinate follows brain y most closely is vertical (v grows with
    y, which grows ventrally, so v = 0 is dorsal); the other is horizontal, oriented so that u
    grows with the distance from the midline x = ``mid_x`` (u = 0 medial, 1 lateral)."""
    centred = xyz - xyz.mean(axis=0)
    _, vecs = np.linalg.eigh(centred.T @ centred)          # eigenvalues ascending
    coords = centred @ vecs[:, [-1, -2]]                   # along the two widest directions
    vert = 1 if abs(_corr(coords[:, 1], xyz[:, 1])) > abs(_corr(coords[:, 0], xyz[:, 1])) else 0
    v, u = coords[:, vert], coords[:, 1 - vert]
    if _corr(v, xyz[:, 1]) < 0:
        v = -v
    if _corr(u, np.abs(xyz[:, 0] - mid_x)) < 0:
        u = -u
    return np.stack([_unit(u), _unit(v)], axis=1)


def retinotopy(conn, cells, w: int, h: int, overlap: float = 0.1):
    """Screen pixel (px, py), eye and a placed-mask for each neuron in ``cells`` of a connectome
    with positions (the contract of flyzero/eyes.py:37-69).  Each eye's cells with finite
    positions get visual-field coordinates from ``eye_plane``; a cell without a side label
    belongs to the eye on its side of the median x of all neurons (above it: left).  An eye with
    fewer than 3 placed cells is not placed (mask False, pixel of field (0, 0))."""
    cells = np.asarray(cells, np.int64)
    table = conn.neurons
    xyz = table[["pos_x", "pos_y", "pos_z"]].to_numpy(float)[cells]
    label = table["side"].to_numpy(dtype=object, na_value="")[cells]
    mid_x = float(np.nanmedian(table["pos_x"].to_numpy(float)))
    finite = np.isfinite(xyz).all(axis=1)
    by_midline = np.where(xyz[:, 0] > mid_x, "left", "right")
    eye = np.where((label == "left") | (label == "right"), label, by_midline).astype(str)
    field_uv = np.zeros((len(cells), 2))
    placed = np.zeros(len(cells), bool)
    for side in ("left", "right"):
        members = np.flatnonzero((eye == side) & finite)
        if members.size >= 3:
            field_uv[members] = eye_plane(xyz[members], mid_x)
            placed[members] = True
    px, py = field_to_screen(field_uv[:, 0], field_uv[:, 1], eye, w, h, overlap)
    return px, py, eye, placed


def motion_cells(conn, frame_shape, params: "MotionParams | None" = None):
    """The eye's cells of a connectome in ``MotionEye``'s order (flyzero/vision.py:67-103): T4a-d
    then T5a-d, R7 and R8 with the colour channel, the medulla columnar types with the
    transient or sustained channel.  Returns (idx, px, py, side, kind)."""
    p = params or MotionParams()
    h, w = frame_shape
    kinds = [pol + sub for pol in POLARITIES for sub in SUBTYPES]
    if p.chromatic_gain > 0:
        kinds += list(COLOUR)
    if p.transient_gain > 0 or p.sustained_gain > 0:
        kinds += list(COLUMNAR)
    out = [[], [], [], [], []]
    for kind in kinds:
        cells = conn.find(kind)
        if len(cells) == 0:
            continue
        px, py, side, keep = retinotopy(conn, cells, w, h, p.overlap)
        for lst, a in zip(out, (cells[keep], px[keep], py[keep], side[keep], np.full(keep.sum(), kind))):
            lst.append(a)
    if not out[0] or not any(k[:2] in POLARITIES for k in np.concatenate(out[4])):
        raise ValueError("connectome has no T4/T5 neurons")
    return tuple(np.concatenate(lst) for lst in out)


# FlyWire cell counts standing in for the extra channels of the synthetic fly (both sides)
SYNTHETIC_COUNTS = {"R7": 1500, "R8": 1500, **{k: 1600 for k in COLUMNAR}}


def synthetic_motion_cells(fly, frame_shape, overlap: float = 0.1, seed: int = 9, params=None):
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
    extra = []
    if params is not None and params.chromatic_gain > 0:
        extra += list(COLOUR)
    if params is not None and (params.transient_gain > 0 or params.sustained_gain > 0):
        extra += list(COLUMNAR)
    if extra:
        # further cells, drawn from the neurons that are neither inputs nor descending neurons,
        # at FlyWire's counts scaled to the fly's size, uniformly placed
        from .connectome import DN_TYPES

        taken = set(idx.tolist()) | {int(fly.drive_idx)}
        for t in DN_TYPES:
            taken |= set(np.asarray(fly.find(t)).tolist())
        free = np.setdiff1d(np.arange(fly.weights.shape[0]), np.fromiter(taken, np.int64))
        sizes = [max(2, round(SYNTHETIC_COUNTS[k] * fly.weights.shape[0] / 138_639)) for k in extra]
        pick = rng.choice(free, size=min(sum(sizes), len(free)), replace=False)
        ekind = np.repeat(extra, sizes)[:len(pick)]
        exy = rng.random((len(pick), 2))
        eside = np.where(exy[:, 0] < 0.5, "left", "right")
        epx, epy = field_to_screen(np.abs(2 * exy[:, 0] - 1), exy[:, 1], eside, w, h, overlap)
        idx, px, py = np.r_[idx, pick], np.r_[px, epx], np.r_[py, epy]
        side, kind = np.r_[side, eside], np.r_[kind, ekind]
    return idx, px, py, side, kind


class BatchMotionEye:
    """``MotionEye`` for B flies: ``idx`` (the eye's neurons), ``see(frames) -> rates``."""

    def __init__(self, cells, frame_shape, batch: int, params: MotionParams | None = None, device=None):
        """``cells`` = (idx, px, py, side, kind) as ``motion_cells`` / ``synthetic_motion_cells``
        return (kind: T4a .. T5d, a ``COLUMNAR`` or ``COLOUR`` cell type)."""
        self.p = p = params or MotionParams()
        if (p.transient_gain or p.sustained_gain or p.chromatic_gain) and not p.gain > 0:
            raise ValueError("the feature channels need a positive gain")
        self.h, self.w = frame_shape
        self.gh, self.gw = self.h // p.cell, self.w // p.cell
        self.B = batch
        self.dev = dev = torch.device(device if device is not None else "cuda")
        idx, px, py, side, kind = cells
        self.idx = np.asarray(idx, np.int64)
        self.px, self.py = np.asarray(px, np.int64), np.asarray(py, np.int64)
        self.side = np.asarray(side)
        self.kind = np.asarray(kind).astype(str)
        self.key = np.array([cell_key(k, s) for k, s in zip(self.kind, self.side)], np.int32)
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
        """The eye of a synthetic fly (``flywire_like``) or of a connectome with positions and cell
        types (``flywire.Connectome``)."""
        p = params or MotionParams()
        if hasattr(fly, "input_xy"):
            cells = synthetic_motion_cells(fly, frame_shape, p.overlap, params=p)
        else:
            cells = motion_cells(fly, frame_shape, p)
        return cls(cells, frame_shape, batch, p, device)

    def reset(self, slots=None):
        """The given flies' filters restart from their next frame (a respawned emulator)."""
        if slots is None:
            self.first.fill_(1)
        else:
            self.first[torch.as_tensor(np.atleast_1d(slots), device=self.dev)] = 1

    def see(self, frames: torch.Tensor, brain=None) -> torch.Tensor:
        """One frame per fly: ``frames`` uint8 (B, h, w, 3) on the device.  Returns the (B, n) cell
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
                           p_ext, n_ext, slot, scale, np.float32(p.transient_gain), np.float32(p.sustained_gain),
                           np.float32(p.chromatic_gain), int(p.surround))
        _lib.check(_lib.load().fz_eye_step(ctypes.byref(a), _lib.stream_handle(self.dev)), "fz_eye_step")
        self.first.zero_()
        return self.rates
