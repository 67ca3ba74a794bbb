"""The batched, plastic and surrogate workloads at their full sizes, against the oracles.

* B = 256 flies on the full 138,639-neuron connectome (tile engine, per-tile bucket source,
  the window graph's two fly-half chains, the autotuned launch shape), per-fly drive
  U(20, 120) Hz and visual gain U(0.5, 2.0), uneven episodes with retire / respawn: the
  reference's ``BatchBrain`` cannot hold this batch (flyzero/gpu.py:168 asserts B * N < 2^24).
* B = 64 with the 7,387 synapses onto the four motor DN types plastic and the reward rule
  (flyzero/batch.py:454-495) every frame.
* The full 2048 -> 1024 -> 1024 + GRU(256) -> 8 surrogate, forward and backward.

The CPU oracle simulates a sample of the flies (its ``alive`` mask skips the others; flies are
independent, so a sampled fly's spikes do not depend on which others are simulated), chosen
from both graph chains (flies 0..127 and 128..255).
"""

import numpy as np
import pytest
import torch

from helpers import frame_rates

pytestmark = pytest.mark.gpu

from flyhip import BatchBrain, LIFParams, autotune  # noqa: E402
from oracle.oracle import OracleBrain  # noqa: E402
from oracle.plasticity_oracle import RewardOracle  # noqa: E402

FRAME_MS = 1000.0 / 60.0988


def _np(t):
    return t.detach().cpu().numpy()


def _respawn_oracle(ob, flies):
    for b in flies:
        ob.u[b] = 0.0
        ob.g[b] = 0.0
        ob.ref[b] = 0
        ob.ring[b] = 0


@pytest.mark.parametrize("parts,acc_bits", [(1, 0), (2, 0), (1, 32)])
def test_b256_full_connectome_uneven_episodes_match_oracle(full_fly, parts, acc_bits):
    fly, p, B = full_fly, LIFParams(), 256
    sample = np.array([0, 127, 128, 255])
    rng = np.random.default_rng(5)
    drive = rng.uniform(20.0, 120.0, B).astype(np.float32)
    gain = rng.uniform(0.5, 2.0, B).astype(np.float32)
    frames = 4
    rates = frame_rates(fly, frames, gain, drive)                 # [frames, B, n_in + 1]
    idx = np.r_[fly.input_idx, fly.drive_idx]
    gb = BatchBrain(fly.weights, p, batch=B, seed=11, engine="tile", tile=1024)
    gb.set_input(idx, rates[0])
    shape = dict(autotune.tune(gb, use_cache=False, reps=1))
    shape["tile_source"] = 1                                      # the bucket source: two graph chains
    shape["block_parts"] = parts                                  # blocks in one part, or two (half the LDS)
    shape["acc_bits"] = acc_bits                                  # two 16-bit step sums per LDS word where allowed, or one int
    gb.set_tune(**shape)
    ob = OracleBrain(fly.weights, p, batch=B, seed=11)

    # episode schedule: the bench's (lengths U{20..600} frames, random start phase), plus
    # forced retirements of sampled flies in both chains (128 after frame 0, 255 after frame 1)
    erng = np.random.default_rng(105)
    length = erng.integers(20, 601, B)
    pos = erng.integers(0, length)
    alive = np.ones(B, bool)
    force = {0: [128], 1: [255]}
    events = 0
    for f in range(frames):
        gb.set_input(idx, rates[f])
        ob.set_input(idx, rates[f])
        ob.alive = (alive & np.isin(np.arange(B), sample)).astype(np.uint8)
        cg = _np(gb.run(FRAME_MS))                               # one graph replay per frame
        co = ob.run(FRAME_MS)
        np.testing.assert_array_equal(cg[sample], co[sample], err_msg=f"frame {f}")
        assert cg[~alive].sum() == 0                             # retired flies skipped
        assert (cg[alive].sum(1) > 0).all()                      # every live fly spikes
        # end of frame: respawn last frame's retirees at rest, retire finished episodes
        back = np.flatnonzero(~alive)
        if len(back):
            gb.reset(back)
            _respawn_oracle(ob, back)
            length[back] = erng.integers(20, 601, len(back))
            pos[back] = 0
            alive[back] = True
        pos[alive] += 1
        done = alive & (pos >= length)
        done[force.get(f, [])] = True
        alive &= ~done
        events += int(done.sum())
        gb.alive.copy_(torch.from_numpy(alive.astype(np.uint8)))
    assert events >= 2
    np.testing.assert_array_equal(_np(gb.u)[sample], ob.u[sample])
    np.testing.assert_array_equal(_np(gb.g)[sample], ob.g[sample])
    np.testing.assert_array_equal(_np(gb.ref)[sample], ob.ref[sample])
    gb.check()


def test_b64_reward_plasticity_on_7387_synapses_matches_oracles(full_fly):
    from flyhip.batch import BatchReward, LearnParams, rule_constants
    from flyhip.connectome import plastic_positions

    fly, p, B = full_fly, LIFParams(), 64
    pos = plastic_positions(fly)
    assert 7000 <= len(pos) <= 7600
    sample = np.array([0, 21, 42, 63])
    rng = np.random.default_rng(7)
    drive = rng.uniform(20.0, 120.0, B).astype(np.float32)
    gain = rng.uniform(0.5, 2.0, B).astype(np.float32)
    frames = 6
    rates = frame_rates(fly, frames, gain, drive)
    idx = np.r_[fly.input_idx, fly.drive_idx]
    gb = BatchBrain(fly.weights, p, batch=B, seed=3, plastic_pos=pos, engine="tile", tile=1024)
    gb.set_input(idx, rates[0])
    shape = dict(autotune.tune(gb, use_cache=False, reps=1))
    shape["plastic_pull"] = 1                                    # plastic synapses pulled in-window
    gb.set_tune(**shape)
    ob = OracleBrain(fly.weights, p, batch=B, seed=3, plastic_pos=pos, frac_bits=gb.frac_bits)
    ob.alive = np.isin(np.arange(B), sample).astype(np.uint8)
    lp = LearnParams(tau_elig_ms=200.0, eta=5e-2, w_abs_max=50.0)
    br = BatchReward(gb, fly, lp)                                 # fly_of_slot = arange(B)
    ro = RewardOracle(gb.pl_pre, gb.pl_post, gb.w0_syn, B, cap=lp.cap, w_abs_max=lp.w_abs_max)
    c = rule_constants(lp, FRAME_MS, gb.p.w_syn)
    count0 = ro.count().copy()
    rewards = rng.normal(0.0, 1.0, (frames, B))
    for f in range(frames):
        gb.set_input(idx, rates[f])
        ob.set_input(idx, rates[f])
        cg = gb.run(FRAME_MS)
        co = ob.run(FRAME_MS)
        np.testing.assert_array_equal(_np(cg)[sample], co[sample], err_msg=f"frame {f}")
        adv = br.step(cg, rewards[f], True, FRAME_MS)
        # the rule on every slot, from the GPU's counts (equal to the oracle's on the sample)
        np.testing.assert_array_equal(adv, ro.step(_np(cg), rewards[f], True, c["ae"], c["ap"], c["ar"], c["eta"]))
        np.testing.assert_array_equal(_np(br.elig), ro.elig, err_msg=f"frame {f}")
        np.testing.assert_array_equal(_np(br.w), ro.w, err_msg=f"frame {f}")
        np.testing.assert_array_equal(_np(br.count), ro.count(), err_msg=f"frame {f}")
        np.testing.assert_array_equal(_np(gb.pw), ro.pw(gb.p.w_syn))
        ob.pw[...] = ro.pw(gb.p.w_syn)
    assert int((ro.count() != count0).sum()) > 0                # whole-synapse edits happened
    np.testing.assert_array_equal(_np(gb.u)[sample], ob.u[sample])


def test_full_size_surrogate_forward_backward_matches_oracle():
    from flyhip.surrogate import N_OUT_PAD, RateSurrogate
    from oracle.surrogate_oracle import bf16, forward_backward

    n_in, hid, H, n_out, T, B = 2048, 1024, 256, 8, 4, 16      # training takes T * B in multiples of 32
    m = RateSurrogate(n_in, hid, H, n_out, seed=21)
    P = m.params()
    rng = np.random.default_rng(22)
    X = bf16(rng.random((T, B, n_in)) * 2.0)                      # pooled rates, scaled
    h0 = rng.normal(0, 0.3, (B, H)).astype(np.float32)
    Y = rng.random((T, B, n_out)).astype(np.float32)
    x16 = torch.from_numpy(np.asarray(X, np.float32)).cuda().to(torch.bfloat16)
    y, _ = m.forward(x16, torch.from_numpy(h0).cuda(), train=True)
    loss = float(m.loss_backward(x16, torch.from_numpy(Y).cuda()).item())
    Yp = np.zeros((T, B, N_OUT_PAD))
    Yp[..., :n_out] = Y
    want_loss, g, want_y = forward_backward(P, X, h0.astype(np.float64), Yp, n_out, bf16_ops=True)
    np.testing.assert_allclose(y.cpu().numpy(), want_y[..., :n_out], rtol=1e-3, atol=1e-3)
    assert abs(loss - want_loss) < 1e-4 * abs(want_loss) + 1e-6
    G = m.grads()
    for name in P:
        a, b = G[name], g[name]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel < 2e-3, (name, rel)
