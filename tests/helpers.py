"""Shared test helpers (inputs only; every expected value comes from the oracle or mathematics)."""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def ref_weights(n=300, seed=2, density=0.05, mean=40.0, sd=30.0):
    """The random matrix of the reference's GPU tests (tests/test_gpu.py:16-20)."""
    rng = np.random.default_rng(seed)
    w = sp.random(n, n, density=density, random_state=seed, format="csr")
    w.data = rng.normal(mean, sd, w.nnz)
    return w


def pack_bits(mask: np.ndarray) -> np.ndarray:
    """bool [..., m] -> uint32 [..., ceil(m/32)], bit j%32 of word j//32."""
    m = mask.shape[-1]
    pad = (-m) % 32
    if pad:
        mask = np.concatenate([mask, np.zeros(mask.shape[:-1] + (pad,), bool)], -1)
    by = np.packbits(mask, axis=-1, bitorder="little")
    return np.ascontiguousarray(by).view(np.uint32)


def unpack_bits(words: np.ndarray, m: int) -> np.ndarray:
    by = np.ascontiguousarray(words).view(np.uint8)
    return np.unpackbits(by, axis=-1, bitorder="little")[..., :m].astype(bool)


def poisson_hits(p: np.ndarray, steps: int, batch: int, seed: int) -> np.ndarray:
    """Precomputed Poisson hits: bool draws U < p per (step, fly, ext slot), packed."""
    rng = np.random.default_rng(seed)
    return pack_bits(rng.random((steps, batch, len(p))) < np.asarray(p, np.float32)[None, None, :])


def fly_inputs(fly):
    idx = np.r_[fly.input_idx, fly.drive_idx].astype(np.int64)
    rate = np.r_[fly.input_rate, np.float32(fly.drive_hz)].astype(np.float32)
    return idx, rate


def frame_rates(fly, frames: int, gain=None, drive=None) -> np.ndarray:
    """[frames, B, n_inputs + 1] Hz: the smooth retinotopic field of each frame times each fly's
    visual gain, and each fly's drive-neuron rate in the last column."""
    from flyhip.connectome import smooth_field

    vis = np.stack([smooth_field(fly.input_xy, f) for f in range(frames)])
    gain = np.ones(1, np.float32) if gain is None else np.asarray(gain, np.float32)
    drive = np.full(len(gain), fly.drive_hz, np.float32) if drive is None else np.asarray(drive, np.float32)
    out = np.concatenate([vis[:, None, :] * gain[None, :, None],
                          np.broadcast_to(drive[None, :, None], (frames, len(gain), 1))], -1)
    return np.minimum(out, 1e4).astype(np.float32)
