"""The fleet (flyzero/fleet.py:22-136): progress unwrapping, the summary and the synthetic race on
the CPU; on the GPU, ``think`` against its parts and a closed-loop exam against the CPU chain of
oracles (eye, brain, DN readout) driving the same game, button for button."""

import numpy as np
import pytest
import torch

from flyhip.batch import READOUT, BatchMotor, MotorParams
from flyhip.fleet import FRAME_SHAPE, SyntheticRace, Progress, buttons_to_mask, summarize

H, W = FRAME_SHAPE


def test_progress_unwraps_the_lap():
    # flyzero/tune.py:65-71: steps are taken modulo the lap, in [-S/2, S/2)
    p = Progress(59)
    assert [p.update(s, i) for i, s in enumerate([57, 58, 0, 1, 0, 58])] == [0, 1, 2, 3, 2, 1]
    # brute force: any walk of steps shorter than half a lap is recovered exactly
    rng = np.random.default_rng(0)
    for S in (59, 84, 10):
        steps = rng.integers(-(S // 2) + 1, S // 2, 400)
        pos = np.cumsum(np.r_[7, steps])
        p = Progress(S)
        got = [p.update(int(x) % S, i) for i, x in enumerate(pos)]
        assert got[-1] == pos[-1] - pos[0] and got == list(pos - pos[0])
    # exactly half a lap reads as backwards
    p = Progress(60)
    p.update(0, 0)
    assert p.update(30, 1) == -30


def test_summary_of_hand_made_results():
    res = [{"progress": 10, "lap": 0, "frames": 100, "finished": False, "laps5": False},
           {"progress": 300, "lap": 5, "frames": 900, "finished": True, "laps5": True},
           {"progress": -3, "lap": 0, "frames": 51, "finished": False}]
    assert summarize(res) == {"mean_progress": 102.3, "best": 300, "worst": -3, "laps_best": 5, "finished": 1,
                              "laps5": 1, "n": 3, "mean_frames": 350}
    mask = buttons_to_mask({"B": True, "LEFT": True, "R": True})
    assert mask.dtype == np.uint8 and list(np.flatnonzero(mask)) == [0, 6, 11]


def _press(*keys):
    return {k: True for k in keys}


def test_synthetic_race_dynamics_and_telemetry():
    race = SyntheticRace(3, segments=59, max_frames=500)
    frames, infos = race.load([SyntheticRace.start_state()] * 3)
    assert frames.shape == (3, H, W, 3) and frames.dtype == torch.uint8
    assert infos[0] == {"segment": 0, "lap": 0, "energy": 100.0, "done": False, "finished": False, "rank": None,
                        "frame": 0}
    for _ in range(10):
        frames, infos = race.step([_press("B"), _press("B", "LEFT"), {}])
    v = race.st[:, 2]
    assert v[0] == pytest.approx(0.04) and v[2] == 0.0           # 0.004 per frame on the road; no gas
    for _ in range(90):
        frames, infos = race.step([_press("B"), _press("B", "LEFT"), {}])
    s, x = race.st[:, 0], race.st[:, 1]
    assert s[0] > 5 and s[2] == 0.0 and infos[0]["segment"] == int(s[0]) % 59
    assert x[1] < x[0]                                            # steering left moves the car left
    # off the road: energy drains until the race ends; a finished race stops changing
    race.load([SyntheticRace.start_state(x=1.9, v=0.2, energy=1.0)], which=[2])
    for _ in range(3):
        frames, infos = race.step([{}, {}, _press("RIGHT")])
    assert infos[2]["done"] and infos[2]["energy"] <= 0 and not infos[0]["done"]
    before = race.st[2].copy()
    race.step([{}, {}, _press("B")])
    np.testing.assert_array_equal(race.st[2], before)
    # five laps: finished
    race.load([SyntheticRace.start_state(s=5 * 59 - 0.1, v=0.2)], which=[0])
    _, infos = race.step([_press("B"), {}, {}])
    assert infos[0]["finished"] and infos[0]["done"] and infos[0]["lap"] == 5
    # save / load round trip
    saved = race.save([1])
    race2 = SyntheticRace(1)
    f2, _ = race2.load(saved)
    np.testing.assert_array_equal(race2.st[0, :4], race.st[1, :4])
    np.testing.assert_array_equal(f2[0].numpy(), race.render([1])[0].numpy())


def test_synthetic_race_view():
    race = SyntheticRace(2)
    frames, _ = race.load([SyntheticRace.start_state(), SyntheticRace.start_state(x=0.5)], flips=[False, True])
    f = frames.numpy()
    assert (f[:, :SyntheticRace.SKY] == SyntheticRace.PALETTE[0]).all()
    # bottom row (p = 1): road of half-width 0.45 W + 2 centred at W / 2 - x hw, white rims
    hw = 0.45 * W + 2
    row = f[0, H - 1]
    road = np.flatnonzero((row != SyntheticRace.PALETTE[1]).any(1) & (row != SyntheticRace.PALETTE[2]).any(1))
    assert road[0] == max(0, int(np.floor(W / 2 - hw))) and road[-1] == min(W - 1, int(np.floor(W / 2 + hw)) - 1)
    assert (row[road[0]] == 250).all() and (row[road[-1]] == 250).all()
    # the second player sits right of centre (road shifted left) and sees the view mirrored
    unflipped = race.render([1])
    race.flip[1] = False
    np.testing.assert_array_equal(unflipped[0].numpy(), race.render([1])[0].numpy()[:, ::-1])
    row = race.render([1])[0, H - 1].numpy()
    rim = np.flatnonzero((row == 250).all(1))     # only the right rim is on screen
    assert rim.max() == int(np.floor(W / 2 - 0.5 * hw + hw)) - 1 and len(rim) == int(np.floor(hw / 20))
    # the road's shading moves as the car advances: the eye sees motion
    a = race.render([0])
    race.st[0, 0] += 0.1
    assert (race.render([0]) != a).any()


def test_game_pool_runs_one_game_per_process_like_the_batched_race():
    # flyzero/pool.py:133-203: one game per spawned process; here the frames come back through
    # shared memory and must equal the in-process batched game's, as must the telemetry
    from functools import partial

    from flyhip.fleet import GamePool

    pool = GamePool(partial(SyntheticRace, 1, max_frames=30), 3)
    ref = SyntheticRace(3, max_frames=30)
    try:
        states = [SyntheticRace.start_state(v=0.05 * k, x=0.3 * k) for k in range(3)]
        f1, i1 = pool.load(states, flips=[False, True, False])
        f2, i2 = ref.load(states, flips=[False, True, False])
        np.testing.assert_array_equal(f1, f2.numpy())
        assert i1 == i2
        rng = np.random.default_rng(0)
        for _ in range(8):
            buttons = [{k: bool(rng.random() < 0.5) for k in ("B", "LEFT", "RIGHT", "Y", "L")} for _ in range(3)]
            f1, i1 = pool.step(buttons)
            f2, i2 = ref.step(buttons)
            np.testing.assert_array_equal(f1, f2.numpy())
            assert i1 == i2
        assert pool.save([1]) == ref.save([1])
        f1, i1 = pool.load([states[0]], which=[2])
        f2, i2 = ref.load([states[0]], which=[2])
        assert f1.shape == (1, H, W, 3)
        np.testing.assert_array_equal(f1, f2.numpy())
        assert i1 == i2
        with pytest.raises(ValueError, match="one message"):
            pool.load(states[:2])
    finally:
        pool.close()
    assert all(not p.is_alive() for p in pool.procs)
    pool.close()                                   # idempotent
    with GamePool(partial(SyntheticRace, 1), 1) as one:
        f, info = one.load([SyntheticRace.start_state()])
        assert f.shape == (1, H, W, 3) and info[0]["segment"] == 0


# ---------------------------------------------------------------- GPU

def _schema_connectome(n=5000, seed=0):
    """A small connectome in the FlyWire schema (flywire.Connectome): cell types, sides and
    positions for the eye's T4/T5, medulla and R7/R8 cells and the descending neurons, random
    Dale-signed whole-synapse counts."""
    import pandas as pd
    import scipy.sparse as sp

    from flyhip.flywire import Connectome

    rng = np.random.default_rng(seed)
    types = ([f"T{p}{d}" for p in (4, 5) for d in "abcd"] * 60 + ["Mi1", "Tm3", "Tm1", "Tm2", "Tm4", "Mi4", "Mi9",
                                                                   "Tm9", "R7", "R8"] * 30)
    dns = ["DNa02", "DNa01", "DNg02_a", "DNp09", "MDN", "DNp01"] * 4
    ct = np.array(types + dns + [None] * (n - len(types) - len(dns)), dtype=object)
    side = np.where(rng.random(n) < 0.5, "left", "right").astype(object)
    pos = rng.random((n, 3)) * 1000
    pos[:, 0] = np.where(side == "left", 600 + 300 * rng.random(n), 100 + 300 * rng.random(n))
    neurons = pd.DataFrame({"root_id": np.arange(n), "cell_type": pd.array(ct, dtype="string"),
                            "hemibrain_type": pd.array([None] * n, dtype="string"),
                            "side": pd.array(side, dtype="string"), "pos_x": pos[:, 0], "pos_y": pos[:, 1],
                            "pos_z": pos[:, 2]})
    w = sp.random(n, n, density=0.015, format="csr", random_state=seed, dtype=np.float32)
    w.data = np.rint(1 + 5 * w.data).astype(np.float32)
    sign = np.where(rng.random(n) < 0.6, 1.0, -1.0).astype(np.float32)
    w = sp.csr_matrix(sp.diags(sign) @ w)
    w.setdiag(0)
    w.eliminate_zeros()
    w.sort_indices()
    return Connectome(w, neurons, name="schema test")


def _small_fleet(B, seed, pool, conn=None, **kw):
    from flyhip.connectome import flywire_like
    from flyhip.fleet import Fleet

    fly = conn if conn is not None else flywire_like(n=6000, n_edges=6000 * 50, seed=4)
    return Fleet(pool, B, seed=seed, conn=fly, tile=512, tune=False, **kw), fly


@pytest.mark.gpu
def test_think_is_eye_then_brain_and_host_frames_match_device_frames():
    from flyhip import BatchBrain, LIFParams
    from flyhip.eye import BatchMotionEye
    from flyhip.fleet import vision_params

    B = 3
    race = SyntheticRace(B, device="cuda")
    fleet, fly = _small_fleet(B, 5, race)
    twin, _ = _small_fleet(B, 5, SyntheticRace(B))         # host frames
    brain = BatchBrain(fly.weights, LIFParams(dt=0.25), batch=B, seed=5, plastic_pos=fleet.pos, engine="tile",
                       tile=512)
    brain.set_bias(fleet.bias_idx, 7.8)
    eye = BatchMotionEye.for_fly(fly, FRAME_SHAPE, B, vision_params())
    frames, _ = race.load([SyntheticRace.start_state(v=0.1)] * B)
    host, _ = twin.pool.load([SyntheticRace.start_state(v=0.1)] * B)
    assert fleet.frames_on_device(frames) is frames          # device frames are used in place
    total = 0
    for t in range(6):
        c1 = fleet.think(frames).cpu().numpy()
        c2 = twin.think(host.numpy()).cpu().numpy()
        brain.p_ext.zero_()
        eye.see(frames, brain=brain)
        c3 = brain.run(fleet.window).cpu().numpy()
        np.testing.assert_array_equal(c1, c3, err_msg=f"frame {t}")
        np.testing.assert_array_equal(c1, c2, err_msg=f"frame {t}")
        total += c1.sum()
        buttons = [_press("B", "LEFT")] * B
        frames, _ = race.step(buttons)
        host, _ = twin.pool.step(buttons)
    assert total > 0
    # extra inputs add on top of the eye's
    fleet.extra_idx = fly.find("DNp01")
    fleet.think(frames, extra=np.full((B, len(fleet.extra_idx)), 200.0, np.float32))
    j = fleet.brain.slots(fleet.extra_idx).cpu().numpy()
    np.testing.assert_allclose(fleet.brain.p_ext[:, j].cpu().numpy(), 200.0 * 0.25 / 1000, rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("taps,schema", [(False, False), (True, False), (False, True)])
def test_exam_matches_the_cpu_chain_of_oracles(taps, schema):
    # schema: a connectome in the FlyWire schema, whose eye comes from cell types and the
    # retinotopy of neuron positions
    from flyhip.fleet import Fleet  # noqa: F401
    from oracle.eye_oracle import OracleEye
    from oracle.oracle import OracleBrain
    from oracle.plasticity_oracle import dn_rates_step

    B, frames_n = 3, 14
    start = SyntheticRace.start_state(x=0.6, v=0.15, energy=3.0)
    fleet, fly = _small_fleet(B, 11, SyntheticRace(B, device="cuda", max_frames=12), taps=taps,
                              steer_threshold=4.0, lean_threshold=4.0, conn=_schema_connectome() if schema else None)
    if schema:
        assert len(fleet.eye.idx) == 8 * 60 + 10 * 30 and set(fleet.eye.key) == set(range(14))
    got_buttons = []
    res = fleet.exam(start, frames_n, on_frame=lambda i, c, b, inf: got_buttons.append(b), record=True)

    # the same drive on the CPU: oracle eye per fly -> oracle brain -> readout -> buttons -> game
    from flyhip import LIFParams

    ob = OracleBrain(fly.weights, LIFParams(dt=0.25), batch=B, seed=11, plastic_pos=fleet.pos,
                     frac_bits=fleet.brain.frac_bits)
    ob.set_bias(fleet.bias_idx, 7.8)
    e = fleet.eye
    ej = ob.slots(e.idx)
    vis = dict(gain=10 ** 2.7, y_min=0.2, transient_gain=4.0, sustained_gain=4.0, chromatic_gain=1.0)
    eyes = [OracleEye(e.key, e.gy, e.gx, e.py, FRAME_SHAPE, **vis) for _ in range(B)]
    mp = fleet.motor.p
    groups = [fly.find(*mp.groups[k]) for k in READOUT]
    host_motor = BatchMotor.__new__(BatchMotor)
    host_motor.p, host_motor.acc = mp, np.zeros((B, 2))
    rates = np.zeros((B, len(READOUT)), np.float32)
    race = SyntheticRace(B, max_frames=12)
    frames, infos = race.load([start] * B)
    alive = np.ones(B, bool)
    progs = [Progress(59) for _ in range(B)]
    want = [None] * B
    for i in range(frames_n):
        ob.p_ext[:] = 0
        for b in range(B):
            ob.p_ext[b, ej] = eyes[b].see(frames[b].numpy()) * np.float32(0.25 / 1000.0)
        ob.alive = alive.astype(np.uint8)
        counts = ob.run(fleet.window)
        rates = dn_rates_step(rates, counts, groups, fleet.window, mp.tau_ms)
        buttons = [host_motor.buttons(rates[k], k) for k in range(B)]
        for k in np.flatnonzero(alive):
            assert buttons[k] == got_buttons[i][k], (i, k)
        frames, infos = race.step(buttons)
        for k in np.flatnonzero(alive):
            progs[k].update(infos[k]["segment"], i)
            if infos[k]["done"] or i == frames_n - 1:
                alive[k] = False
                want[k] = (progs[k].total, infos[k]["lap"], i + 1, infos[k]["energy"])
        if not alive.any():
            break
    for k in range(B):
        assert (res[k]["progress"], res[k]["lap"], res[k]["frames"], res[k]["energy"]) == want[k], k
        assert res[k]["masks"].shape == (res[k]["frames"], 12)
    pressed = sum(int(m.sum()) for r in res for m in r["masks"])
    assert (pressed > 0 or schema) and fleet.brain.alive.all()


@pytest.mark.gpu
def test_retiring_finished_slots_changes_no_result():
    race_a = SyntheticRace(4, device="cuda", max_frames=10)
    race_b = SyntheticRace(4, device="cuda", max_frames=10)
    start = SyntheticRace.start_state(x=0.9, v=0.2, energy=1.5)
    fa, _ = _small_fleet(4, 3, race_a)
    fb, _ = _small_fleet(4, 3, race_b)
    ra = fa.exam(start, 12, retire=True, record=True)
    rb = fb.exam(start, 12, retire=False, record=True)
    for a, b in zip(ra, rb):
        assert a["frames"] == b["frames"] and a["progress"] == b["progress"] and a["energy"] == b["energy"]
        np.testing.assert_array_equal(a["masks"], b["masks"])
        np.testing.assert_array_equal(a["rates"], b["rates"])


@pytest.mark.gpu
def test_exam_with_games_in_subprocesses_equals_the_device_game():
    from functools import partial

    from flyhip.fleet import GamePool

    start = SyntheticRace.start_state(x=0.3, v=0.12, energy=2.0)
    pool = GamePool(partial(SyntheticRace, 1, max_frames=9), 2)
    try:
        fa, _ = _small_fleet(2, 7, pool)
        ra = fa.exam(start, 10, record=True)
    finally:
        pool.close()
    fb, _ = _small_fleet(2, 7, SyntheticRace(2, device="cuda", max_frames=9))
    rb = fb.exam(start, 10, record=True)
    for a, b in zip(ra, rb):
        assert (a["frames"], a["progress"], a["energy"]) == (b["frames"], b["progress"], b["energy"])
        np.testing.assert_array_equal(a["masks"], b["masks"])
        np.testing.assert_array_equal(a["rates"], b["rates"])


@pytest.mark.gpu
@pytest.mark.parametrize("pool", ["device", "processes"])
def test_a_fleet_in_parts_drives_exactly_like_one_brain(pool):
    # two brains of 2 flies each, run in turn with the games of one part stepping while the other
    # part simulates, against one brain of 4 flies: same noise per fly, same results
    from functools import partial

    from flyhip.fleet import GamePool

    start = SyntheticRace.start_state(x=0.5, v=0.12, energy=2.5)

    def make():
        if pool == "device":
            return SyntheticRace(4, device="cuda", max_frames=11)
        return GamePool(partial(SyntheticRace, 1, max_frames=11), 4)

    pa, pb = make(), make()
    try:
        fa, _ = _small_fleet(4, 9, pa, parts=2)
        fb, _ = _small_fleet(4, 9, pb)
        assert fa.brain is None and [p.nb for p in fa.parts] == [2, 2]
        ra = fa.exam(start, 12, record=True)
        rb = fb.exam(start, 12, record=True)
    finally:
        for p in (pa, pb):
            p.close()
    for a, b in zip(ra, rb):
        assert (a["frames"], a["progress"], a["energy"]) == (b["frames"], b["progress"], b["energy"])
        np.testing.assert_array_equal(a["masks"], b["masks"])
        np.testing.assert_array_equal(a["rates"], b["rates"])
    assert sum(int(r["masks"].sum()) for r in ra) > 0
