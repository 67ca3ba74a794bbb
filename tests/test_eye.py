"""Batched vision front end (flyzero/vision.py:64-191, flyzero/eyes.py:37-88): host mapping,
retinotopy and the CPU restatement on the CPU (closed forms, direction selectivity, the
reference's own formulation); the HIP kernel against the restatement bit for bit on the GPU."""

import numpy as np
import pytest

from flyhip.eye import (DIRECTIONS, FEATURE_KEYS, MotionParams, cell_key, field_to_screen, motion_cells, retinotopy,
                        screen_direction)
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


def _feature_cells(gh, gw):
    """One cell per (grid point, feature key 8..13)."""
    gy, gx = np.meshgrid(np.arange(gh), np.arange(gw), indexing="ij")
    gy, gx = gy.ravel(), gx.ravel()
    key = np.repeat(np.arange(8, 14), len(gy))
    return key, np.tile(gy, 6), np.tile(gx, 6)


def _rates_of(r, key, gy, gx, k, gh, gw):
    out = np.zeros((gh, gw), f32)
    sel = key == k
    out[gy[sel], gx[sel]] = r[sel]
    return out


def test_feature_channels_closed_forms():
    gh, gw = H // 4, W // 4
    key, gy, gx = _feature_cells(gh, gw)
    tg, sg, cg, gain = 0.5, 2.0, 0.8, 50.0
    eye = OracleEye(key, gy, gx, gy * 4, (H, W), gain=gain, transient_gain=tg, sustained_gain=sg,
                    chromatic_gain=cg, surround=3)
    # a flat frame: no contrast; colour = channel / 255; first frame: no transient
    frame = np.empty((H, W, 3), np.uint8)
    frame[...] = (10, 100, 200)
    r = eye.see(frame)
    assert r[key == 8].max() == 0.0 and r[key == 9].max() == 0.0
    assert r[key == 10].max() < 1e-3 and r[key == 11].max() < 1e-3     # the window mean's rounding
    np.testing.assert_allclose(r[key == 12], 150 * cg * 200 / 255, rtol=1e-5)
    np.testing.assert_allclose(r[key == 13], 150 * cg * 100 / 255, rtol=1e-5)
    # a brighter second frame: hp = (1 - 1/tau_hp) * step on the ON transient channel only
    frame2 = frame.copy()
    frame2[..., 0] += 30
    r = eye.see(frame2)
    np.testing.assert_allclose(r[key == 8], 150 * min(1.0, tg * 0.75 * 10 / 255), rtol=1e-4)
    assert r[key == 9].max() == 0.0
    # one bright grid point: centre minus the 7x7 mean, clipped to the grid at the corner
    eye = OracleEye(key, gy, gx, gy * 4, (H, W), gain=gain, sustained_gain=sg, surround=3)
    frame = np.zeros((H, W, 3), np.uint8)
    frame[40:44, 80:84] = 51                      # grid point (10, 20), luminance 0.2
    frame[0:4, 0:4] = 51                          # the corner: a 4x4 window
    r = eye.see(frame)
    on = _rates_of(r, key, gy, gx, FEATURE_KEYS["on_sustained"], gh, gw)
    off = _rates_of(r, key, gy, gx, FEATURE_KEYS["off_sustained"], gh, gw)
    np.testing.assert_allclose(on[10, 20], 150 * sg * 0.2 * 48 / 49, rtol=1e-5)
    np.testing.assert_allclose(off[12, 22], 150 * sg * 0.2 / 49, rtol=1e-5)
    assert off[10, 24] == 0.0 and on[12, 22] == 0.0   # outside the window; inside, darker than the mean
    np.testing.assert_allclose(on[0, 0], 150 * sg * 0.2 * 15 / 16, rtol=1e-5)


def test_oracle_feature_channels_match_the_reference_formulation():
    # flyzero/vision.py:159-183: transient from hp, sustained contrast from a float32 integral image
    # finished in float64, colour from one channel's cell mean
    rng = np.random.default_rng(4)
    gh, gw = H // 4, W // 4
    key, gy, gx = _feature_cells(gh, gw)
    p = dict(gain=10 ** 2.7, transient_gain=4.0, sustained_gain=4.0, chromatic_gain=1.0, surround=3, y_min=0.2)
    eye = OracleEye(key, gy, gx, gy * 4, (H, W), **p)
    lp = None
    for t in range(4):
        frame = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
        got = eye.see(frame)
        img = frame.astype(f32)
        lum = (img.mean(2) / 255.0).reshape(gh, 4, gw, 4).mean((1, 3))
        lp = lum.copy() if lp is None else lp
        lp += 0.25 * (lum - lp)
        hp = lum - lp
        r = p["surround"]
        ii = np.pad(lum, ((1, 0), (1, 0))).cumsum(0).cumsum(1)
        yy, xx = np.mgrid[0:gh, 0:gw]
        y0, y1 = np.clip(yy - r, 0, gh), np.clip(yy + r + 1, 0, gh)
        x0, x1 = np.clip(xx - r, 0, gw), np.clip(xx + r + 1, 0, gw)
        sur = (ii[y1, x1] - ii[y0, x1] - ii[y1, x0] + ii[y0, x0]) / ((y1 - y0) * (x1 - x0))
        con = lum - sur
        chan = lambda c: (img[..., c] / 255.0).reshape(gh, 4, gw, 4).mean((1, 3))  # noqa: E731
        feats = {8: 4.0 * np.maximum(hp, 0), 9: 4.0 * np.maximum(-hp, 0), 10: 4.0 * np.maximum(con, 0),
                 11: 4.0 * np.maximum(-con, 0), 12: 1.0 * chan(2), 13: 1.0 * chan(1)}
        m = np.array([feats[k][y, x] / p["gain"] for k, y, x in zip(key, gy, gx)], f32)
        want = 150.0 * np.clip(p["gain"] * m, 0, 1)
        want[gy * 4 < 0.2 * H] = 0
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=0.02)


def _brain_cloud(rng, n_side, tilt):
    """Cells on a full lattice of a tilted plane per side (so the in-plane axes are exactly the
    principal axes): lateral along |x - mid|, dorsal-ventral along y; shuffled."""
    nu = 10
    u, v = np.meshgrid(np.linspace(0, 1, nu), np.linspace(0, 1, n_side // nu), indexing="ij")
    perm = rng.permutation(n_side)
    u, v = u.ravel()[perm], v.ravel()[perm]
    pos = []
    for sgn in (1, -1):          # x above the midline: the left eye
        x = 500 + sgn * (100 + 200 * u)
        y = 300 + 400 * v
        z = 50 + tilt * x
        pos.append(np.c_[x, y, z])
    return np.vstack(pos), u, v


def test_retinotopy_recovers_the_field_axes():
    import pandas as pd

    from flyhip.flywire import Connectome

    rng = np.random.default_rng(0)
    n_side = 200
    pos, u, v = _brain_cloud(rng, n_side, tilt=0.3)
    side = np.r_[["left"] * n_side, [None] * n_side]          # right-eye labels missing: the midline decides
    neurons = pd.DataFrame({"cell_type": pd.array(["T4a"] * 2 * n_side, dtype="string"),
                            "side": pd.array(side, dtype="string"),
                            "pos_x": pos[:, 0], "pos_y": pos[:, 1], "pos_z": pos[:, 2]})
    import scipy.sparse as sp
    conn = Connectome(sp.csr_matrix((2 * n_side, 2 * n_side), dtype=np.float32), neurons)
    px, py, s, keep = retinotopy(conn, np.arange(2 * n_side), W, H, overlap=0.1)
    assert keep.all() and list(s[:n_side]) == ["left"] * n_side and list(s[n_side:]) == ["right"] * n_side
    want_l = field_to_screen(u, v, np.array(["left"] * n_side), W, H, 0.1)
    want_r = field_to_screen(u, v, np.array(["right"] * n_side), W, H, 0.1)
    # u = 0 medial .. 1 lateral, v = 0 dorsal .. 1; rounding can move a cell by one pixel
    assert np.abs(px[:n_side] - want_l[0]).max() <= 1 and np.abs(py[:n_side] - want_l[1]).max() <= 1
    assert np.abs(px[n_side:] - want_r[0]).max() <= 1 and np.abs(py[n_side:] - want_r[1]).max() <= 1
    assert px[:n_side].max() <= int(0.55 * W) and px[n_side:].min() >= int(0.45 * W)


def test_motion_cells_of_a_connectome_follow_the_reference_order():
    import pandas as pd
    import scipy.sparse as sp

    from flyhip.flywire import Connectome

    rng = np.random.default_rng(1)
    types = ["T4a", "T5c", "Mi1", "Tm9", "R7", "R8", "T4d", "KC"]
    n_side = 20
    rows = []
    for t in types:
        pos, _, _ = _brain_cloud(rng, n_side, tilt=0.1)
        for k in range(2 * n_side):
            rows.append((t, "left" if k < n_side else "right", *pos[k]))
    df = pd.DataFrame(rows, columns=["cell_type", "side", "pos_x", "pos_y", "pos_z"])
    df["cell_type"] = df["cell_type"].astype("string")
    df["side"] = df["side"].astype("string")
    conn = Connectome(sp.csr_matrix((len(df), len(df)), dtype=np.float32), df)
    idx, px, py, side, kind = motion_cells(conn, (H, W), MotionParams())
    assert list(dict.fromkeys(kind)) == ["T4a", "T4d", "T5c"]          # motion only by default
    vis = MotionParams(transient_gain=4.0, sustained_gain=4.0, chromatic_gain=1.0)
    idx, px, py, side, kind = motion_cells(conn, (H, W), vis)
    assert list(dict.fromkeys(kind)) == ["T4a", "T4d", "T5c", "R7", "R8", "Mi1", "Tm9"]
    keys = {k: cell_key(k, "left") for k in ("T4a", "Mi1", "Tm9", "R7", "R8")}
    assert keys == {"T4a": DIRECTIONS.index("left"), "Mi1": 8, "Tm9": 11, "R7": 12, "R8": 13}
    assert cell_key("T4a", "right") == DIRECTIONS.index("right") and cell_key("T5c", "left") == 4 + 3
    assert "KC" not in set(kind) and len(idx) == 7 * 2 * n_side


VISION = dict(transient_gain=4.0, sustained_gain=4.0, chromatic_gain=1.0)   # flyzero/train.py:21


@pytest.mark.gpu
@pytest.mark.parametrize("channels", ["motion", "vision"])
def test_batched_eye_matches_oracle_and_feeds_the_brain(channels):
    import torch

    from flyhip import BatchBrain, LIFParams
    from flyhip.connectome import flywire_like
    from flyhip.eye import BatchMotionEye

    fly = flywire_like(n=6000, n_edges=6000 * 50, seed=4)
    B = 3
    extra = VISION if channels == "vision" else {}
    p = MotionParams(gain=200.0, y_min=0.05, **extra)
    eye = BatchMotionEye.for_fly(fly, (H, W), B, p)
    assert (eye.key >= 8).any() == (channels == "vision")
    gb = BatchBrain(fly.weights, LIFParams(), batch=B, engine="tile", tile=512)

    def oracle():
        return OracleEye(eye.key, eye.gy, eye.gx, eye.py, (H, W), gain=200.0, y_min=0.05, **extra)

    oracles = [oracle() for _ in range(B)]
    rng = np.random.default_rng(1)
    base = rng.integers(0, 256, (B, H, W, 3)).astype(np.uint8)
    for t in range(8):
        frames = np.stack([np.roll(base[b], 4 * ((t * (b + 1)) % 7), axis=1) for b in range(B)])
        if t == 5:                                     # fly 1 respawns: its filters restart
            eye.reset([1])
            oracles[1] = oracle()
        got = eye.see(torch.from_numpy(frames).cuda(), brain=gb).cpu().numpy()
        for b in range(B):
            np.testing.assert_array_equal(got[b], oracles[b].see(frames[b]), err_msg=f"frame {t} fly {b}")
        assert t < 2 or got.max() > 0                  # the delay arm is empty for two frames
    slots = gb.slots(eye.idx).cpu().numpy()
    np.testing.assert_array_equal(gb.p_ext.cpu().numpy()[:, slots], got * f32(gb.p.dt / 1000.0))
    counts = gb.run(16.6).cpu().numpy()
    assert counts.sum() > 0


@pytest.mark.gpu
def test_batched_eye_on_a_frame_that_is_not_a_multiple_of_the_cell():
    # 101 x 131 pixels, cell 4: a 25 x 32 grid, the last rows and columns of pixels unused
    import torch

    from flyhip.connectome import flywire_like
    from flyhip.eye import BatchMotionEye

    fly = flywire_like(n=3000, n_edges=3000 * 40, seed=8)
    shape = (101, 131)
    p = MotionParams(gain=300.0, surround=2, **VISION)
    eye = BatchMotionEye.for_fly(fly, shape, 2, p)
    assert (eye.gh, eye.gw) == (25, 32)
    oracles = [OracleEye(eye.key, eye.gy, eye.gx, eye.py, shape, gain=300.0, surround=2, **VISION) for _ in range(2)]
    rng = np.random.default_rng(2)
    for t in range(4):
        frames = rng.integers(0, 256, (2,) + shape + (3,)).astype(np.uint8)
        got = eye.see(torch.from_numpy(frames).cuda()).cpu().numpy()
        for b in range(2):
            np.testing.assert_array_equal(got[b], oracles[b].see(frames[b]), err_msg=f"frame {t} fly {b}")
