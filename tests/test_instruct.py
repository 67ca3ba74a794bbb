"""Instructed learning (flyzero/batch.py:226-451, flyzero/instruct.py:124-148): host logic and the
CPU restatement on the CPU; the four HIP kernels against it (bit for bit) and against the
reference tests' plain formulas (tests/test_gpu.py:75-149,186-219) on the GPU."""

import numpy as np
import pytest
import scipy.sparse as sp

from flyhip.connectome import flywire_like
from flyhip.instruct import (GROUPS, InstructParams, deep2_positions, deep_positions, intent, plastic_positions,
                             targets, targets_from_buttons, targets_from_intent)
from oracle import instruct_oracle as io

f32 = np.float32


# ------------------------------------------------------------------ host logic (CPU)

def test_targets_closed_forms():
    p = InstructParams()
    t = targets_from_intent([1.0, 0.0, 0.5, 0.0, 0.0], p)          # full right, half gas
    assert (t["a02L"], t["a02R"], t["g02L"], t["g02R"]) == (p.low_hz, p.high_hz, p.high_hz, p.low_hz)
    assert t["gas"] == p.low_hz + 0.5 * (p.high_hz - p.low_hz) and t["brake"] == p.low_hz
    t = targets_from_intent([0.0, -1.0, 0.0, 1.0, 0.0], p)         # straight, lean left, brake
    assert t["a02L"] == t["a02R"] == t["g02L"] == t["g02R"] == p.rest_hz
    assert (t["a01L"], t["a01R"], t["brake"]) == (p.high_hz, p.low_hz, p.high_hz)
    b = targets_from_buttons({"LEFT": True, "B": True}, p)
    assert (b["a02L"], b["a02R"], b["g02R"], b["g02L"], b["gas"]) == (p.high_hz, p.low_hz, p.high_hz, p.low_hz,
                                                                      p.high_hz)
    both = targets_from_buttons({"LEFT": True, "RIGHT": True, "L": True}, p)
    assert (both["a02L"], both["a02R"], both["g02L"], both["g02R"], both["a01L"]) == (p.rest_hz,) * 4 + (p.high_hz,)
    # the button form is the intent form at full commands
    for buttons, x in (({"RIGHT": True, "Y": True}, [1.0, 0.0, 0.0, 1.0]), ({"LEFT": True, "R": True}, [-1.0, 1.0, 0.0, 0.0])):
        assert targets_from_buttons(buttons, p) == targets_from_intent(x, p)
    tl = targets([0.0, 0.0, 0.0, 0.0, 1.0], p)
    assert len(tl) == len(GROUPS) == 9 and tl[-1] == p.high_hz     # the giant fiber follows boost


def test_intent_is_a_causal_exponential_filter():
    buttons = ["LEFT", "RIGHT", "L", "R", "B", "Y", "A"]
    m = np.zeros((40, len(buttons)), bool)
    m[10:, buttons.index("RIGHT")] = True
    x = intent(m, buttons, tau_frames=4.0)
    assert (x[:10, 0] == 0).all()
    np.testing.assert_allclose(x[10:, 0], 1 - (1 - 0.25) ** np.arange(1, 31), rtol=1e-6)


def _tiny():
    # pre x post; DNs are 5 and 6, L1 = {1, 2} (pre of DN synapses), L2 = {0, 3} (pre of L1 synapses)
    rows = [0, 1, 2, 3, 1, 2, 4, 0]
    cols = [1, 5, 6, 2, 6, 2, 0, 3]
    w = sp.csr_matrix((np.ones(len(rows), np.float32), (rows, cols)), shape=(7, 7))
    w.sort_indices()
    return w


def test_deep_positions_by_brute_force():
    w = _tiny()

    class C:
        n = 7
        weights = w

    rows = np.repeat(np.arange(7), np.diff(w.indptr))
    dn_pos = np.flatnonzero(np.isin(w.indices, [5, 6]))
    l1 = set(rows[dn_pos])
    assert l1 == {1, 2}
    want = [k for k in range(w.nnz) if w.indices[k] in l1]
    np.testing.assert_array_equal(deep_positions(C, dn_pos), want)
    l2 = {rows[k] for k in want} - l1 - {5, 6}
    want2 = [k for k in range(w.nnz) if w.indices[k] in l2]
    np.testing.assert_array_equal(deep2_positions(C, dn_pos), want2)


# ------------------------------------------------------------------ oracle pins (CPU)

def test_oracle_traces_converge_geometrically():
    B, n = 2, 5
    counts = np.full((B, n), 3, np.int32)
    tr, ra = np.zeros((B, 2), f32), np.zeros((B, 1), f32)
    a_pre, a_rate, hz = f32(0.2), f32(0.1), f32(1000 / 16.6)
    for k in range(1, 30):
        io.traces(counts, [0, 4], [2], tr, ra, a_pre, f32(0.05), a_rate, hz)
        np.testing.assert_allclose(tr, 3 * hz * 0.05 * (1 - 0.8 ** k), rtol=2e-5)
        np.testing.assert_allclose(ra, 3 * hz * (1 - 0.9 ** k), rtol=2e-5)


def test_oracle_delta_rule_averages_learning_slots_and_clips():
    # 2 flies x 2 slots, 3 synapses; fly 1's second slot does not learn
    pw = np.array([[1.0, -0.1, 0.5]] * 4, f32)
    sign, wmax = np.array([1, -1, 1], f32), np.array([3.0, 3.0, 0.6], f32)
    trace = np.array([[1.0], [2.0], [1.0], [5.0]], f32)
    rate = np.zeros((4, 1), f32)
    tgt = np.array([[10.0], [20.0], [40.0], [40.0]], f32)
    learn = np.array([True, True, True, False])
    eta = np.full(4, 0.01)
    io.delta_rule(pw, None, np.zeros(3, int), np.zeros(3, int), np.zeros(3, int), None, trace, rate, tgt, eta, learn,
                  [0, 0, 1, 1], sign, wmax)
    d0 = (0.01 * 1 * 10 + 0.01 * 2 * 20) / 2                        # fly 0: mean of both slots
    d1 = 0.01 * 1 * 40                                              # fly 1: its learning slot only
    np.testing.assert_allclose(pw[:2, 0], 1 + d0, rtol=1e-6)
    np.testing.assert_allclose(pw[2:, 0], 1 + d1, rtol=1e-6)
    assert (pw[:, 1] == 0).all()                                     # Dale: an inhibitory weight pushed up stops at 0
    assert (pw[:, 2] == f32(0.6)).all()                              # cap


def test_oracle_deep_update_skips_zero_changes_and_clips():
    pw = np.array([[5.0, 1.0]] * 2, f32)                             # 5.0 is above the cap
    sign, wmax = np.ones(2, f32), np.array([3.0, 3.0], f32)
    trace = np.ones((2, 1), f32)
    delta = np.array([[0.0, 2.0], [0.0, 2.0]], f32)
    io.deep_update(pw, [0, 1], np.zeros(2, int), np.array([0, 1]), trace, delta, np.array([0.5, 0.0], f32), [0, 0],
                   sign, wmax)
    assert (pw[:, 0] == 5.0).all()                                   # no change: not clipped either
    assert (pw[:, 1] == 2.0).all()                                   # 1 + 0.5 * 1 * 2, both slots of the fly


def test_oracle_intrinsic_bias_is_bounded():
    ib = np.zeros((2, 1), f32)
    bias = np.full((2, 3), 7.8, f32)
    rate = np.full((2, 1), 100.0, f32)
    tgt = np.zeros((2, 1), f32)
    for _ in range(50):
        io.intrinsic_bias(ib, bias, np.array([2]), np.array([0]), np.array([0]), tgt, rate, np.array([True, False]),
                          [0, 0], 0.01, np.ones(2), 2.0)
    np.testing.assert_allclose(ib, -2.0)
    np.testing.assert_allclose(bias[:, 2], 5.8, rtol=1e-6)
    assert (bias[:, :2] == f32(7.8)).all()


# ------------------------------------------------------------------ GPU

@pytest.fixture(scope="module")
def fly():
    return flywire_like(n=4000, n_edges=4000 * 60, seed=21)


def _brain(fly, B, layers=2):
    from flyhip import BatchBrain, LIFParams

    p1 = plastic_positions(fly)
    pos = p1
    if layers >= 1:
        pos = np.union1d(pos, deep_positions(fly, p1))
    if layers >= 2:
        pos = np.union1d(pos, deep2_positions(fly, p1))
    return BatchBrain(fly.weights, LIFParams(dt=0.25), batch=B, plastic_pos=pos, engine="tile", tile=512)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.gpu
def test_instruct_step_matches_oracle_and_plain_formulas(fly):
    import torch

    from flyhip.instruct import BatchInstructDeep2

    B = 4
    gb = _brain(fly, B, layers=2)
    bi = BatchInstructDeep2(gb, fly, fly_of_slot=[0, 0, 1, 1], eta_deep=1e-3, eta_deep2=1e-4, eta_bias=0.01)
    assert bi.n_deep > 0 and bi.n_deep2 > 0
    rng = np.random.default_rng(0)
    bi.trace.copy_(torch.from_numpy(rng.uniform(0, 5, tuple(bi.trace.shape)).astype(f32)))
    bi.rate.copy_(torch.from_numpy(rng.uniform(0, 60, tuple(bi.rate.shape)).astype(f32)))
    counts = torch.from_numpy(rng.integers(0, 3, (B, gb.n)).astype(np.int32)).cuda()
    tgt = rng.uniform(0, 80, (B, len(bi.names))).astype(f32)
    learn = np.array([True, False, True, True])
    eta = np.array([2e-4, 1e-4, 3e-4, 2e-4])
    # oracle state before the step
    tr, ra, pw0 = _np(bi.trace).copy(), _np(bi.rate).copy(), _np(gb.pw).copy()
    ib, bias = _np(bi.ib).copy(), _np(gb.bias_ext).copy()
    p = bi.p
    err_out = bi.step(counts, tgt, learn, 16.6, eta=eta)
    # ---- oracle: traces, retrograde delta, DN update, deep update, second layer, biases
    c = _np(counts)
    io.traces(c, bi.upre, bi.upost, tr, ra, f32(1 - np.exp(-16.6 / p.tau_pre_ms)), f32(p.tau_pre_ms / 1000.0),
              f32(1 - np.exp(-16.6 / p.tau_rate_ms)), f32(1000.0 / 16.6))
    np.testing.assert_array_equal(_np(bi.trace), tr)
    np.testing.assert_array_equal(_np(bi.rate), ra)
    w = pw0.copy()
    sel, pg, pk, qk = bi.dn_sel_np, bi.post_group_np, bi.pre_k_np, bi.post_k_np
    lists = [[] for _ in bi.l1]
    for s in sel:                                      # ascending plastic index within each L1 neuron
        lists[int(np.searchsorted(bi.l1, gb.pl_pre[s]))].append((s, pg[s], qk[s]))
    d1 = io.retrograde(w, lists, tgt, ra, gb.p.w_syn)
    np.testing.assert_array_equal(_np(bi.delta), d1)
    io.delta_rule(w, sel, pk[sel], pg[sel], qk[sel], None, tr, ra, tgt, eta.astype(f32), learn, bi.fly_of_slot,
                  bi.sign_np, bi.wmax_np)
    dsel = bi.deep_sel_np
    coef1 = io.deep_coef(learn, eta.astype(f32), bi.eta_deep, p.eta, bi.deep_scale, bi.fly_of_slot)
    io.deep_update(w, dsel, pk[dsel], bi.deep_post_l1_np, tr, d1, coef1, bi.fly_of_slot, bi.sign_np, bi.wmax_np)
    l2_index = {int(v): k for k, v in enumerate(bi.l2)}
    l1_index = {int(v): k for k, v in enumerate(bi.l1)}
    back = [[] for _ in bi.l2]
    for s in dsel:
        back[l2_index[int(gb.pl_pre[s])]].append((s, l1_index[int(gb.pl_post[s])], 0))
    d2 = io.retrograde(w, back, d1, np.zeros((B, 1), f32), gb.p.w_syn)
    np.testing.assert_array_equal(_np(bi.delta2), d2)
    sel2 = _np(bi.d_k2_sel)
    coef2 = io.deep_coef(learn, eta.astype(f32), bi.eta_deep2, p.eta, bi.deep_scale, bi.fly_of_slot)
    io.deep_update(w, sel2, pk[sel2], _np(bi.d_k2_post), tr, d2, coef2, bi.fly_of_slot, bi.sign_np, bi.wmax_np)
    np.testing.assert_array_equal(_np(gb.pw), w)
    dn_slot = _np(bi.d_dn_slot)
    io.intrinsic_bias(ib, bias, dn_slot, _np(bi.d_dn_group), _np(bi.d_dn_rate_k), tgt, ra, learn, bi.fly_of_slot,
                      bi.eta_bias, (eta / p.eta).astype(f32), bi.bias_limit)
    np.testing.assert_array_equal(_np(bi.ib), ib)
    np.testing.assert_array_equal(_np(gb.bias_ext), bias)
    assert np.abs(w - pw0)[:, dsel].max() > 1e-4 and np.abs(w - pw0)[:, sel2].max() > 1e-6
    # ---- the reference test's plain formula for the deep synapses (tests/test_gpu.py:90-99)
    err = tgt[:, pg[sel]] - ra[:, qk[sel]]
    delta = np.zeros((B, len(bi.l1)), f32)
    np.add.at(delta.T, bi.dn_pre_l1_np, (pw0[:, sel] / gb.p.w_syn * err).T)
    rate_b = eta * (bi.eta_deep / p.eta)
    change = rate_b[:, None] * tr[:, pk[dsel]] * delta[:, bi.deep_post_l1_np]
    per_fly = np.stack([change[0], change[2:].mean(0)])
    sg, mx = bi.sign_np[dsel], bi.wmax_np[dsel]
    plain = sg * np.clip(sg * (pw0[:, dsel] + per_fly[[0, 0, 1, 1]]), 0, mx)
    np.testing.assert_allclose(_np(gb.pw)[:, dsel], plain, rtol=1e-4, atol=1e-5)
    assert err_out[1] == 0 and (err_out[[0, 2, 3]] > 0).all()


@pytest.mark.gpu
def test_intrinsic_plasticity_moves_bias_towards_target(fly):
    # tests/test_gpu.py:105-124 (without the save/load round trip)
    import torch

    from flyhip.instruct import BatchInstruct

    gb = _brain(fly, 2, layers=0)
    gb.set_bias(fly.find("DNp09"), 7.8)
    bi = BatchInstruct(gb, fly, fly_of_slot=[0, 0], eta_bias=0.01, bias_limit=2.0)
    j = bi.dn_slots
    before = _np(gb.bias_ext[:, j]).copy()
    tgt = np.zeros((2, len(bi.names)), f32)
    zeros = torch.zeros((2, gb.n), dtype=torch.int32, device="cuda")
    for _ in range(50):
        bi.rate.fill_(100.0)                               # every DN far above its target
        bi.step(zeros, tgt, np.array([True, False]), 16.6)
    after = _np(gb.bias_ext[:, j])
    assert (after[0] < before[0]).all() and (after[0] >= before[0] - 2.0 - 1e-5).all()
    np.testing.assert_allclose(after[1], after[0])         # same fly: every slot changes together


@pytest.mark.gpu
def test_consolidated_weights_average_and_swap_back(fly):
    # tests/test_gpu.py:127-149
    from flyhip.instruct import BatchInstruct

    gb = _brain(fly, 1, layers=0)
    bi = BatchInstruct(gb, fly, eta_bias=0.01)
    w0, b0 = _np(gb.pw).copy(), _np(gb.bias_ext).copy()
    bi.consolidate(0.5)
    gb.pw.mul_(0.5)
    bi.ib += 1.0
    gb.bias_ext[:, bi.dn_slots] += 1.0
    bi.consolidate(0.5)
    fast_w, fast_b, fast_ib = _np(gb.pw).copy(), _np(gb.bias_ext).copy(), _np(bi.ib).copy()
    bi.swap_slow()
    np.testing.assert_allclose(_np(gb.pw), 0.75 * w0, rtol=1e-6)
    np.testing.assert_allclose(_np(bi.ib), fast_ib - 0.5, atol=1e-6)
    bi.swap_slow()
    np.testing.assert_allclose(_np(gb.pw), fast_w)
    np.testing.assert_allclose(_np(gb.bias_ext), fast_b)
    np.testing.assert_allclose(_np(bi.ib), fast_ib, atol=1e-6)
    assert b0.shape == fast_b.shape


@pytest.mark.gpu
def test_instructed_learning_drives_a_group_towards_its_target(fly):
    # closed loop: the engine runs, the rule reads its counts; slot 0 is taught "full gas"
    # (DNp09 at 80 Hz), slot 1 (another fly, same input) is not taught
    from flyhip.instruct import BatchInstruct

    gb = _brain(fly, 2, layers=0)
    idx = np.r_[fly.input_idx, fly.drive_idx]
    gb.set_input(idx, np.r_[fly.input_rate, f32(60.0)])
    bi = BatchInstruct(gb, fly, fly_of_slot=[0, 1], params=InstructParams(eta=2e-3))
    p = bi.p
    tgt = np.array([targets([0, 0, 1.0, 0, 0], p)] * 2, f32)
    gas = bi.names.index("gas")
    for f in range(80):
        c = gb.run(16.6)
        bi.step(c, tgt, np.array([True, False]), 16.6)
    rate = _np(bi.rate)
    mem = bi.group_members[gas]
    taught, control = rate[0, mem].mean(), rate[1, mem].mean()
    assert taught > control + 10.0, (taught, control)
    w = _np(gb.pw)
    onto_gas = np.isin(gb.pl_post, fly.find("DNp09"))
    exc = onto_gas & (bi.sign_np > 0)
    assert (w[0, exc] >= w[1, exc]).all() and (w[0, exc] > w[1, exc]).any()


@pytest.mark.gpu
def test_batch_motor_rates_and_buttons(fly):
    # flyzero/batch.py:77-138: the 9 readout groups (the giant fiber drives A) from device counts
    import torch

    from flyhip.batch import READOUT, BatchMotor, MotorParams

    B = 3
    motor = BatchMotor(fly, B, MotorParams(boost_threshold=0.0))
    rng = np.random.default_rng(2)
    ema = np.zeros((B, len(READOUT)), f32)
    a = f32(1.0 - np.exp(-16.6 / 80.0))
    for t in range(4):
        c = rng.integers(0, 3, (B, fly.n)).astype(np.int32)
        buttons = motor.update(torch.from_numpy(c).cuda(), 16.6)
        for k, name in enumerate(READOUT):
            ids = fly.find(*motor.p.groups[name])
            inst = c[:, ids].sum(1).astype(f32) / f32(len(ids)) * f32(1000.0 / 16.6)
            ema[:, k] = ema[:, k] + a * (inst - ema[:, k])
        np.testing.assert_allclose(motor.rates, ema, rtol=1e-5)
        assert len(buttons) == B and all(b["A"] == (motor.rates[i, 8] > 0.0) for i, b in enumerate(buttons))


@pytest.mark.gpu
def test_default_rates_change_the_weights_the_simulation_uses(fly):
    # the reference's default rates (InstructParams.eta 2e-4, eta_deep 1e-6) make edits far below
    # half a synapse per frame; with sub-synapse fixed point for plastic weights (frac_bits 8,
    # w_syn / 256) they reach the quantised weights the engine adds, rint(pw * 2^8 / w_syn)
    import torch

    from flyhip.instruct import BatchInstructDeep

    gb = _brain(fly, 4, layers=1)
    assert gb.frac_bits == 8
    gb.set_input(np.r_[fly.input_idx, fly.drive_idx], np.r_[fly.input_rate, f32(60.0)])
    bi = BatchInstructDeep(gb, fly, fly_of_slot=[0, 0, 1, 1])
    assert bi.p.eta == 2e-4 and bi.eta_deep == 1e-6
    qs = float(gb.pw_qscale)
    q0 = torch.round(gb.pw * qs)
    tgt = np.array([targets([1.0, 0.0, 1.0, 0.0, 0.0], bi.p)] * 4, f32)
    for f in range(40):
        bi.step(gb.run(16.6), tgt, np.ones(4, bool), 16.6, report=False)
    moved = _np(torch.round(gb.pw * qs) != q0)
    assert moved[:, bi.dn_sel_np].any(), "DN synapse edits never reach the simulation"
    assert moved[:, bi.deep_sel_np].any(), "deep synapse edits never reach the simulation"
    # edits the engine sees now that whole-synapse rounding (frac_bits 0) would have hidden
    small = np.abs(_np(gb.pw) - _np(q0 / qs)) < 0.5 * float(gb.p.w_syn)
    assert (moved & small)[:, bi.deep_sel_np].any()
