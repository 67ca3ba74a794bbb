"""Batched motion-vision front end (flyzero/vision.py:64-191): host mapping and the CPU
restatement on the CPU (closed forms, direction selectivity); the HIP kernel against the
restatement bit for bit on the GPU."""

import numpy as np
import pytest

from flyhip.eye import DIRECTIONS, MotionParams, field_to_screen, screen_direction
from oracle.eye_oracle import OracleEye

f32 = np.float32
H, W = 224, 256


def test_screen_mapping_follows_the_reference():
    # flyzero/eyes.py:61-69: left eye on [0, 0.5 + o], lateral at the outer edge
    px, py = field_to_screen(np.array([0.0, 1.0, 0.0, 1.0]), np.array([0.0, 0.5, 0.999, 0.25]),
                             np.array(["left", "left", "right", "right"]), W, H, 0.1)
    assert list(px) == [int(0.55 * W), 0, int(0.45 * W), W - 1]
    assert list(py) == [0, H // 2, int(0.999 * H), H // 4]
    # flyzero/vision.py:112-121: front-to-back is leftwards for the left eye
    assert screen_direction("front_to_back", "left") == "left"
    assert screen_direction("front_to_back", "right") == "right"
    assert screen_direction("back_to_front", "left") == "right"
    assert screen_direction("up", "right") == "up"


def test_grid_luminance_of_a_flat_frame():
    eye = OracleEye([0], [0], [0], [100], (H, W))
    frame = np.empty((H, W, 3), np.uint8)
    frame[...] = (30, 60, 90)
    g = eye.grid(frame)
    assert g.shape == (H // 4, W // 4)
    np.testing.assert_allclose(g, (30 + 60 + 90) / 3 / 255, rtol=1e-6)


def _all_cells(gh, gw):
    """One neuron per (grid point, key)."""
    gy, gx = np.meshgrid(np.arange(gh), np.arange(gw), indexing="ij")
    gy, gx = gy.ravel(), gx.ravel()
    key = np.repeat(np.arange(8), len(gy))
    return key, np.tile(gy, 8), np.tile(gx, 8)


def test_moving_bar_excites_the_preferred_direction_only():
    # a bright vertical bar moving right one grid point per frame: ON right-motion detectors fire
    # behind it, ON left-motion detectors stay silent; the mirror movie gives the mirror result
    gh, gw = H // 4, W // 4
    key, gy, gx = _all_cells(gh, gw)
    for step, pref, null in ((1, "right", "left"), (-1, "left", "right")):
        eye = OracleEye(key, gy, gx, gy * 4, (H, W), gain=1000.0)
        for t in range(12):
            frame = np.zeros((H, W, 3), np.uint8)
            x = 32 + step * t
            frame[:, 4 * x:4 * x + 4] = 255
            r = eye.see(frame)
        on_pref = r[key == DIRECTIONS.index(pref)]
        on_null = r[key == DIRECTIONS.index(null)]
        assert on_pref.max() > 100.0 and on_null.max() == 0.0
        assert r[key == DIRECTIONS.index("down")].max() == 0.0 and r[key == DIRECTIONS.index("up")].max() == 0.0


def test_oracle_matches_the_reference_arithmetic():
    # the reference's own formulation (flyzero/vision.py:137-191 with numpy means and shifted
    # slices) against the restatement's loops, on a random movie
    rng = np.random.default_rng(3)
    gh, gw = H // 4, W // 4
    key, gy, gx = _all_cells(gh, gw)
    eye = OracleEye(key, gy, gx, gy * 4, (H, W), gain=300.0, y_min=0.1)
    lp = d_on = d_off = None
    for t in range(6):
        frame = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
        got = eye.see(frame)
        lum = (frame.astype(f32).mean(2) / 255.0).reshape(gh, 4, gw, 4).mean((1, 3))
        if lp is None:
            lp, d_on, d_off = lum.copy(), np.zeros_like(lum), np.zeros_like(lum)
        lp += 0.25 * (lum - lp)
        hp = lum - lp
        maps = {}
        for pol, sig, d in ((0, np.maximum(hp, 0), d_on), (1, np.maximum(-hp, 0), d_off)):
            right = np.zeros_like(sig)
            down = np.zeros_like(sig)
            right[:, :-1] = d[:, :-1] * sig[:, 1:] - sig[:, :-1] * d[:, 1:]
            down[:-1] = d[:-1] * sig[1:] - sig[:-1] * d[1:]
            maps.update({pol * 4: np.maximum(right, 0), pol * 4 + 1: np.maximum(-right, 0),
                         pol * 4 + 2: np.maximum(down, 0), pol * 4 + 3: np.maximum(-down, 0)})
            d += 0.5 * (sig - d)
        m = np.array([maps[k][y, x] for k, y, x in zip(key, gy, gx)], f32)
        want = 150.0 * np.clip(300.0 * m, 0, 1)
        want[gy * 4 < 0.1 * H] = 0
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-3)


@pytest.mark.gpu
def test_batched_eye_matches_oracle_and_feeds_the_brain():
    import torch

    from flyhip import BatchBrain, LIFParams
    from flyhip.connectome import flywire_like
    from flyhip.eye import BatchMotionEye

    fly = flywire_like(n=6000, n_edges=6000 * 50, seed=4)
    B = 3
    p = MotionParams(gain=200.0, y_min=0.05)
    eye = BatchMotionEye.for_fly(fly, (H, W), B, p)
    gb = BatchBrain(fly.weights, LIFParams(), batch=B, engine="tile", tile=512)
    oracles = [OracleEye(eye.key, eye.gy, eye.gx, eye.py, (H, W), gain=200.0, y_min=0.05) for _ in range(B)]
    rng = np.random.default_rng(1)
    base = rng.integers(0, 256, (B, H, W, 3)).astype(np.uint8)
    for t in range(8):
        frames = np.stack([np.roll(base[b], 4 * ((t * (b + 1)) % 7), axis=1) for b in range(B)])
        if t == 5:                                     # fly 1 respawns: its filters restart
            eye.reset([1])
            oracles[1] = OracleEye(eye.key, eye.gy, eye.gx, eye.py, (H, W), gain=200.0, y_min=0.05)
        got = eye.see(torch.from_numpy(frames).cuda(), brain=gb).cpu().numpy()
        for b in range(B):
            np.testing.assert_array_equal(got[b], oracles[b].see(frames[b]), err_msg=f"frame {t} fly {b}")
        assert t < 2 or got.max() > 0                  # the delay arm is empty for two frames
    slots = gb.slots(eye.idx).cpu().numpy()
    np.testing.assert_array_equal(gb.p_ext.cpu().numpy()[:, slots], got * f32(gb.p.dt / 1000.0))
    counts = gb.run(16.6).cpu().numpy()
    assert counts.sum() > 0
