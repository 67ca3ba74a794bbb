"""Reward-modulated plasticity (flyzero/learning.py:116-131, flyzero/batch.py:454-495) and the
DN-rate readout (flyzero/batch.py:107-111).

The reference has no test of ``RewardPlasticity`` / ``BatchReward`` (parity unpinned there), so
the FP64 oracle is pinned here by hand-computed values, closed forms of the recurrences and the
rule's invariants; the GPU kernel is then compared with the oracle exactly.
"""

import numpy as np
import pytest

from oracle.plasticity_oracle import RewardOracle, dn_rates_step

AE, AP, AR, ETA = float(np.exp(-16.64 / 200)), float(1 - np.exp(-16.64 / 500)), float(1 - np.exp(-16.64 / 1000)), 0.05


def _one_synapse(w0=4.0, cap=3.0, wabs=np.inf):
    # two neurons: 0 -> 1, one plastic synapse of w0 synapses
    return RewardOracle([0], [1], [w0], B=1, cap=cap, w_abs_max=wabs)


def test_hand_computed_two_frames():
    o = _one_synapse()
    counts = np.array([[3, 5]])
    adv = o.step(counts, reward=2.0, learn=True, ae=0.5, ap=0.25, ar=0.1, eta=0.01)
    # frame 1: dev = 5 - 0; e = 0.5*0 + 3*5 = 15; adv = 2 - 0; w = 4 + 0.01*2*15 = 4.3
    assert adv[0] == 2.0 and o.elig[0, 0] == 15.0
    assert o.w[0, 0] == pytest.approx(4.3, abs=1e-15)
    assert o.post_mean[0, 0] == 1.25 and o.r_mean[0] == pytest.approx(0.2)
    adv = o.step(counts, reward=2.0, learn=True, ae=0.5, ap=0.25, ar=0.1, eta=0.01)
    # frame 2: dev = 5 - 1.25 = 3.75; e = 7.5 + 11.25 = 18.75; adv = 1.8; w = 4.3 + 0.01*1.8*18.75
    assert o.elig[0, 0] == 18.75 and adv[0] == pytest.approx(1.8)
    assert o.w[0, 0] == pytest.approx(4.3 + 0.3375, abs=1e-12)
    assert o.count()[0, 0] == 5


def test_recurrences_match_closed_forms():
    o = _one_synapse(w0=100.0, cap=1e9)
    cpre, cpost, R = 2, 3, 1.5
    for t in range(1, 31):
        o.step(np.array([[cpre, cpost]]), reward=R, learn=False, ae=AE, ap=AP, ar=AR, eta=ETA)
        # <post>_t = c (1 - (1-ap)^t);  e_t = cpre cpost (ae^t - (1-ap)^t) / (ae - (1-ap));  <R>_t = R (1 - (1-ar)^t)
        q = 1 - AP
        assert o.post_mean[0, 0] == pytest.approx(cpost * (1 - q ** t), rel=1e-12)
        assert o.elig[0, 0] == pytest.approx(cpre * cpost * (AE ** t - q ** t) / (AE - q), rel=1e-6)
        assert o.r_mean[0] == pytest.approx(R * (1 - (1 - AR) ** t), rel=1e-12)
    assert o.w[0, 0] == 100.0                                     # learn=False: weights untouched


def test_dale_sign_and_caps_hold():
    rng = np.random.default_rng(0)
    w0 = rng.integers(-40, 40, 200).astype(np.float64)
    w0[w0 == 0] = 1
    o = RewardOracle(rng.integers(0, 50, 200), rng.integers(50, 60, 200), w0, B=3, cap=3.0, w_abs_max=50.0)
    for _ in range(40):
        c = rng.integers(0, 6, (3, 60))
        o.step(c, reward=rng.normal(0, 5, 3), learn=True, ae=AE, ap=AP, ar=AR, eta=0.5)
    s = np.sign(w0)
    assert np.all(s * o.w >= 0)                                        # never changes sign
    assert np.all(np.abs(o.w) <= np.minimum(3 * np.abs(w0), 50) + 1e-12)
    assert np.abs(o.w - w0).max() > 1                                  # it did learn


def test_zero_advantage_leaves_weights():
    o = RewardOracle([0, 1], [2, 2], [3.0, -2.0], B=2)
    o.r_mean[:] = 1.0
    o.step(np.ones((2, 3), int), reward=1.0, learn=True, ae=AE, ap=AP, ar=AR, eta=1.0)
    assert np.array_equal(o.w, [[3.0, -2.0], [3.0, -2.0]])


def test_slots_of_one_fly_share_the_averaged_change():
    # flyzero/batch.py:158-172: the change is averaged over the fly's learning slots and applied
    # to every slot of the fly; a slot that does not learn contributes nothing
    o = RewardOracle([0], [1], [10.0], B=3, cap=1e9, fly_of_slot=[0, 0, 0])
    counts = np.array([[1, 2], [3, 4], [9, 9]])
    o.step(counts, reward=[1.0, 2.0, 100.0], learn=[True, True, False], ae=0.0, ap=0.5, ar=0.5, eta=0.1)
    # e = cpre * cpost: 2 and 12 (slot 2 learns nothing); mean change = (0.1*1*2 + 0.1*2*12) / 2 = 1.3
    assert np.allclose(o.w[:, 0], 11.3, atol=1e-12)


def test_per_slot_learning_rate_by_hand():
    # flyzero/school.py:289-304: slot b's change is eta_b * adv_b * e_b before the fly's average
    o = RewardOracle([0], [1], [10.0], B=3, cap=1e9, fly_of_slot=[0, 0, 1])
    counts = np.array([[1, 2], [3, 4], [2, 5]])
    o.step(counts, reward=[1.0, 2.0, -1.0], learn=True, ae=0.0, ap=0.5, ar=0.5, eta=[0.1, 0.3, 0.2])
    # fly 0: e = 2 and 12, (0.1*1*2 + 0.3*2*12) / 2 = 3.7; fly 1: e = 10, 0.2 * -1 * 10 = -2
    assert np.allclose(o.w[:, 0], [13.7, 13.7, 8.0], atol=1e-12)
    # one eta for every slot is the scalar rule
    a = RewardOracle([0], [1], [10.0], B=2, cap=1e9)
    b = RewardOracle([0], [1], [10.0], B=2, cap=1e9)
    a.step(counts[:2], reward=[1.0, 2.0], learn=True, ae=0.0, ap=0.5, ar=0.5, eta=0.25)
    b.step(counts[:2], reward=[1.0, 2.0], learn=True, ae=0.0, ap=0.5, ar=0.5, eta=[0.25, 0.25])
    assert np.array_equal(a.w, b.w)


def test_dn_readout_closed_form():
    counts = np.array([[2, 4, 0, 6]])
    r = dn_rates_step(np.zeros((1, 2), np.float32), counts, [np.array([0, 1]), np.array([3])], window_ms=10.0,
                      tau_ms=10.0)
    a = 1 - np.exp(-1.0)
    assert r[0, 0] == pytest.approx(a * 3 * 100, rel=1e-6) and r[0, 1] == pytest.approx(a * 600, rel=1e-6)


# ---------------------------------------------------------------- GPU: kernel vs oracle


def _np(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def small_fly():
    from flyhip.connectome import flywire_like

    return flywire_like(n=8000, n_edges=8000 * 109, seed=13, plastic_target=800)


@pytest.mark.gpu
def test_reward_kernel_equals_fp64_oracle(small_fly):
    import torch

    from flyhip import BatchBrain, LIFParams
    from flyhip.batch import BatchReward, LearnParams, rule_constants
    from flyhip.connectome import plastic_positions

    pos = plastic_positions(small_fly)
    B = 6
    gb = BatchBrain(small_fly.weights, LIFParams(), batch=B, plastic_pos=pos)
    p = LearnParams(tau_elig_ms=200.0, eta=2e-2, w_abs_max=50.0)
    fos = [0, 0, 1, 2, 2, 2]
    br = BatchReward(gb, small_fly, p, fly_of_slot=fos)
    o = RewardOracle(gb.pl_pre, gb.pl_post, gb.w0_syn, B, cap=p.cap, w_abs_max=p.w_abs_max, fly_of_slot=fos)
    rng = np.random.default_rng(3)
    c = rule_constants(p, 16.64, gb.p.w_syn)
    for frame in range(25):
        counts = rng.integers(0, 4, (B, gb.n)).astype(np.int32)
        reward = rng.normal(0, 1, B)
        learn = rng.random(B) < 0.7
        adv = br.step(torch.from_numpy(counts).cuda(), reward, learn, 16.64)
        want = o.step(counts, reward, learn, c["ae"], c["ap"], c["ar"], c["eta"])
        np.testing.assert_array_equal(adv, want)
        np.testing.assert_array_equal(_np(br.elig), o.elig)
        np.testing.assert_array_equal(_np(br.w), o.w, err_msg=f"frame {frame}")
        np.testing.assert_array_equal(_np(br.count), o.count())
        np.testing.assert_array_equal(_np(gb.pw), o.pw(gb.p.w_syn))
        np.testing.assert_array_equal(_np(br.post_mean), o.post_mean)
        np.testing.assert_array_equal(_np(br.r_mean), o.r_mean)
    assert np.abs(o.w - gb.w0_syn[None]).max() >= 1.0          # integer counts moved


@pytest.mark.gpu
def test_reward_kernel_per_slot_eta_equals_fp64_oracle(small_fly):
    # step_scaled (flyzero/school.py:289-304): slots of one fly learn at their own rates
    import torch

    from flyhip import BatchBrain, LIFParams
    from flyhip.batch import BatchReward, LearnParams, rule_constants, step_scaled
    from flyhip.connectome import plastic_positions

    pos = plastic_positions(small_fly)
    B = 5
    gb = BatchBrain(small_fly.weights, LIFParams(), batch=B, plastic_pos=pos)
    p = LearnParams(tau_elig_ms=200.0, eta=2e-2, w_abs_max=50.0)
    fos = [0, 0, 1, 1, 1]
    br = BatchReward(gb, small_fly, p, fly_of_slot=fos)
    o = RewardOracle(gb.pl_pre, gb.pl_post, gb.w0_syn, B, cap=p.cap, w_abs_max=p.w_abs_max, fly_of_slot=fos)
    eta_mv = np.array([2e-2, 6e-2, 0.0, 1e-2, 4e-2])
    eta_syn = eta_mv / float(np.float32(gb.p.w_syn))
    rng = np.random.default_rng(9)
    c = rule_constants(p, 16.64, gb.p.w_syn)
    for frame in range(15):
        counts = rng.integers(0, 4, (B, gb.n)).astype(np.int32)
        reward = rng.normal(0, 1, B)
        learn = rng.random(B) < 0.8
        adv = step_scaled(br, torch.from_numpy(counts).cuda(), reward, learn, 16.64, eta_mv)
        np.testing.assert_array_equal(adv, o.step(counts, reward, learn, c["ae"], c["ap"], c["ar"], eta_syn))
        np.testing.assert_array_equal(_np(br.w), o.w, err_msg=f"frame {frame}")
        np.testing.assert_array_equal(_np(br.count), o.count())
    assert np.abs(o.w - gb.w0_syn[None]).max() >= 1.0


@pytest.mark.gpu
def test_closed_loop_simulation_and_learning_match_oracles(small_fly):
    # config-4 shape at small scale: batched LIF + reward rule, frame by frame; spike counts and
    # integer synapse counts identical to the C LIF oracle + FP64 rule oracle
    import torch

    from flyhip import BatchBrain, LIFParams
    from flyhip.batch import BatchReward, DNReadout, LearnParams, rule_constants
    from flyhip.connectome import plastic_positions
    from oracle.oracle import OracleBrain

    pos = plastic_positions(small_fly)
    B, window = 4, 1000.0 / 60.0988
    gb = BatchBrain(small_fly.weights, LIFParams(), batch=B, seed=2, plastic_pos=pos)
    ob = OracleBrain(small_fly.weights, LIFParams(), batch=B, seed=2, plastic_pos=pos, frac_bits=gb.frac_bits)
    idx = np.r_[small_fly.input_idx, small_fly.drive_idx]
    rate = np.r_[small_fly.input_rate, 60.0].astype(np.float32)
    for b in (gb, ob):
        b.set_input(idx, rate)
    p = LearnParams(tau_elig_ms=200.0, eta=5e-2, w_abs_max=50.0)
    br = BatchReward(gb, small_fly, p)
    ro = RewardOracle(gb.pl_pre, gb.pl_post, gb.w0_syn, B, cap=p.cap, w_abs_max=p.w_abs_max)
    dn = DNReadout(small_fly, B)
    rates = np.zeros((B, 8), np.float32)
    groups = small_fly.group_members()
    c = rule_constants(p, window, gb.p.w_syn)
    rng = np.random.default_rng(6)
    for frame in range(12):
        reward = rng.normal(0, 1, B)
        cg = gb.run(window)
        co = ob.run(window)
        np.testing.assert_array_equal(_np(cg), co, err_msg=f"frame {frame}")
        r_gpu = _np(dn.update(cg, window))
        rates = dn_rates_step(rates, co, groups, window)
        np.testing.assert_array_equal(r_gpu, rates)
        br.step(cg, reward, True, window)
        ro.step(co, reward, True, c["ae"], c["ap"], c["ar"], c["eta"])
        ob.pw[...] = ro.pw(gb.p.w_syn)
        np.testing.assert_array_equal(_np(br.count), ro.count())
    assert co.sum() > 0
