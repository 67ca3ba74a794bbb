"""Pins for the CPU oracle (CPU only).

The reference ships no golden vectors for this path (its tests are property and equivalence
tests), so the oracle is pinned by: the reference's own property tests restated on it, closed
forms of the model (Euler series, delay, refractory period, Poisson renewal rate), a library
SpMV for the integer push, the published inverse of the hash, and an independent numpy
restatement.  Each check is chosen so a plausible slip fails it (comments say which).
"""

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import pack_bits, ref_weights
from oracle import lif_numpy
from oracle.oracle import OracleBrain, OracleParams, constants, mix


def test_derived_constants_dt_0_1_and_0_25():
    # flyzero/brain.py:92-96 evaluated by hand
    c = constants(OracleParams())
    assert c["n_delay"] == 18 and c["n_ref"] == 22
    assert c["decay_m"] == np.float32(0.005)
    assert c["decay_syn"] == np.float32(np.exp(-0.02))
    assert abs(float(c["decay_syn"]) - 0.9801987) < 1e-7
    assert c["kick"] == np.float32(68.75)
    c = constants(OracleParams(dt=0.25))
    assert c["n_delay"] == 7 and c["n_ref"] == 9
    assert c["decay_m"] == np.float32(0.0125)
    assert abs(float(c["decay_syn"]) - 0.95122945) < 1e-7


def _lone_neuron(bias):
    ob = OracleBrain(sp.csr_matrix((1, 1), dtype=np.float32))
    ob.set_bias([0], bias)
    return ob


def test_euler_geometric_series_and_first_spike():
    # v_k = v_rest + b (1 - (1 - dt/tau_m)^k): forward Euler, not exponential integration
    # (a 1 - exp(-dt/tau) leak, or a leak of the wrong sign, fails the trajectory check)
    b = 10.0
    ob = _lone_neuron(b)
    dm = 0.1 / 20.0
    for k in range(1, 200):
        ob.run_steps(1)
        want = -52.0 + b * (1.0 - (1.0 - dm) ** k)
        assert abs(float(ob.v[0, 0]) - want) < 2e-4, k
    # threshold -45 is crossed when 1 - 0.995^k >= 0.7: k = 241 (k = 240 gives -45.0054)
    ob = _lone_neuron(b)
    counts, ras = ob.run_steps(300, raster=True)
    first = int(np.flatnonzero(ras[:, 0, 0] & 1)[0])
    assert first == 240 and counts[0, 0] == 1


def test_bias_below_threshold_never_fires_and_refractory_holds_v():
    ob = _lone_neuron(6.9)                       # v_inf = -45.1 mV < v_th
    assert ob.run_steps(3000)[0, 0] == 0
    ob = _lone_neuron(10.0)                      # period = 241 + 22 held steps
    counts = ob.run_steps(241 + 263 * 3)
    assert counts[0, 0] == 4


def test_synaptic_delay_and_decay():
    # one spike at step 0 reaches the target exactly n_delay steps later as count * w_syn, then
    # decays by decay_syn per step (an off-by-one slot, or decaying the arrival, fails this)
    w = np.zeros((2, 2), np.float32)
    w[0, 1] = 10
    ob = OracleBrain(sp.csr_matrix(w))
    ob.v[0, 0] = -40.0
    gs = []
    for _ in range(25):
        ob.run_steps(1)
        gs.append(float(ob.g[0, 1]))
    nd = 18
    assert all(x == 0.0 for x in gs[:nd])
    assert gs[nd] == float(np.float32(10.0) * np.float32(0.275))
    assert gs[nd + 1] == float(np.float32(gs[nd]) * np.float32(np.exp(-0.02)))
    # the input moves v in the same step it arrives (g is updated before v)
    assert float(ob.v[0, 1]) > -52.0


def test_spike_interval_is_n_ref_under_constant_drive():
    # numba/GPU semantics (brain.py:50): the spike test follows the decrement, and kicks land
    # while refractory, so a neuron kicked every step fires every n_ref steps (the numpy
    # backend's pre-decrement test would give n_ref + 1)
    ob = OracleBrain(sp.csr_matrix((1, 1), dtype=np.float32))
    ob.set_input([0], 1e6)                        # p >= 1: a kick every step
    counts, ras = ob.run_steps(220, raster=True)
    steps = np.flatnonzero(ras[:, 0, 0] & 1)
    assert list(steps) == list(range(0, 220, 22))
    ob2 = OracleBrain(sp.csr_matrix((1, 1), dtype=np.float32))
    ob2.set_input([0], 1e6)
    ob2.run_steps(2)                              # spike at step 0, one kick while refractory
    assert float(ob2.v[0, 0]) == float(np.float32(-52.0) + np.float32(68.75))


def test_poisson_renewal_rate_closed_form():
    # unconnected neurons, rate r: ISI = n_ref if any kick lands in the n_ref refractory steps,
    # else n_ref + Geometric(p); E[ISI] = n_ref + (1-p)^n_ref / p steps
    dt, r, n = 0.25, 100.0, 200
    ob = OracleBrain(sp.csr_matrix((n, n), dtype=np.float32), OracleParams(dt=dt), seed=5)
    ob.set_input(np.arange(n), r)
    secs = 10
    c = ob.run_steps(int(secs * 1000 / dt))
    p = float(np.float32(r * dt / 1000.0))
    isi = 9 + (1 - p) ** 9 / p
    want = 1000.0 / (isi * dt)
    got = c.mean() / secs
    assert abs(got - want) / want < 0.01, (got, want)
    # the reference's own bounds (tests/test_gpu.py:65-72)
    assert 70 < c.mean() / secs < 105


def test_reference_driven_neuron_properties():
    # tests/test_flyzero.py:16-24
    w = np.zeros((2, 2))
    w[0, 1] = 100
    ob = OracleBrain(sp.csr_matrix(w))
    ob.set_input([0], 200.0)
    c = ob.run(200)[0]
    assert c[0] > 10 and c[1] > 5
    ob = OracleBrain(sp.csr_matrix(-w))
    ob.set_input([0], 200.0)
    assert ob.run(200)[0][1] == 0


def test_integer_push_is_the_library_spmv():
    # one step with neurons forced above threshold: the ring slot equals the scipy row sum of
    # the quantised matrix (a transposed matrix or a pre/post swap fails; w is asymmetric)
    w = ref_weights(n=200, seed=7)
    ob = OracleBrain(w, frac_bits=8)
    spk = np.array([3, 17, 44, 150, 199])
    ob.v[0, spk] = -40.0
    ob.run_steps(1)
    net = ob.net
    q = sp.csr_matrix((net.w, net.col, net.indptr), shape=(200, 200))
    want = np.asarray(q[spk].sum(0)).ravel()
    assert np.array_equal(ob.ring[0, 0], want.astype(np.int32))
    # and in mV it is the reference's dense check within float rounding (test_flyzero.py:27-34)
    ws = sp.csr_matrix(w, dtype=np.float32)
    dense = np.asarray(ws[spk].sum(0)).ravel() * 0.275
    np.testing.assert_allclose(ob.ring[0, 0] * float(net.in_scale), dense, rtol=1e-4, atol=1e-3)


def test_hash_matches_published_inverse():
    # lowbias32 and its published inverse (hash-prospector): the multipliers are inverses mod
    # 2^32 and the round trip is the identity, which pins every constant of mix()
    assert (0x7FEB352D * 0x1D69E2A5) % 2 ** 32 == 1
    assert (0x846CA68B * 0x43021123) % 2 ** 32 == 1

    def inv(x):
        x ^= x >> 16
        x = (x * 0x43021123) & 0xFFFFFFFF
        x ^= (x >> 15) ^ (x >> 30)
        x = (x * 0x1D69E2A5) & 0xFFFFFFFF
        x ^= x >> 16
        return x

    rng = np.random.default_rng(0)
    for x in [0, 1, 2, 0xFFFFFFFF] + list(rng.integers(0, 2 ** 32, 200)):
        x = int(x)
        assert inv(mix(x)) == x
        assert int(lif_numpy.mix(np.uint64(x))) == mix(x)


def test_hash_uniform_hits():
    # per-step hit frequency over many (neuron, step) pairs matches p (chi-square-ish bound)
    n = 4000
    ob = OracleBrain(sp.csr_matrix((n, n), dtype=np.float32), seed=3)
    ob.set_input(np.arange(n), 1000.0)     # p = 0.1 per step; a kick always fires (then refractory)
    ob.ref[:] = 0
    # count first-step hits only: fresh neurons, one step
    c = ob.run_steps(1)
    frac = c.mean()
    assert abs(frac - 0.1) < 4 * np.sqrt(0.1 * 0.9 / n)


def _numpy_twin(ob, steps, hits=None):
    net = ob.net
    v, g, ref, ring = ob.u.copy(), ob.g.copy(), ob.ref.copy(), ob.ring.copy()
    counts = lif_numpy.run(net, ob.ext_idx, v, g, ref, ring, ob.t, ob.step_no, steps, ob.p_ext, ob.bias_ext,
                           ob.pw if len(net.plastic_pos) else None, ob.seed, hits=hits, alive=ob.alive)
    return counts, (v, g, ref, ring)


@pytest.mark.parametrize("use_hits", [False, True])
def test_c_oracle_equals_numpy_restatement(use_hits):
    w = ref_weights(n=300, seed=2)
    ws = sp.csr_matrix(w, dtype=np.float32)
    ws.sort_indices()
    pos = np.flatnonzero(ws.indices < 40)
    ob = OracleBrain(w, OracleParams(dt=0.25), batch=3, seed=4, plastic_pos=pos)
    ob.pw[2] *= 1.7
    ob.set_input(np.arange(0, 300, 7), 150.0)
    ob.set_bias(np.arange(5, 300, 11), 3.0)
    ob.v[:, :40] = -40.0
    ob.alive = np.array([1, 1, 0], np.uint8)
    steps = 120
    hits = None
    if use_hits:
        hits = pack_bits(np.random.default_rng(1).random((steps, 3, ob.p_ext.shape[1])) < 0.05)
    want, (v, g, ref, ring) = _numpy_twin(ob, steps, hits)
    got = ob.run_steps(steps, hits=hits)
    assert want.sum() > 100
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(ob.u, v)
    np.testing.assert_array_equal(ob.g, g)
    np.testing.assert_array_equal(ob.ref, ref)
    np.testing.assert_array_equal(ob.ring, ring)


def test_openmp_sweep_equals_serial():
    w = ref_weights(n=1000, seed=5, density=0.02)
    a = OracleBrain(w, seed=2)
    b = OracleBrain(w, seed=2)
    for o in (a, b):
        o.set_input(np.arange(0, 1000, 3), 120.0)
    ca, ra = a.run_steps(300, raster=True, nthreads=1)
    cb, rb = b.run_steps(300, raster=True, nthreads=4)
    np.testing.assert_array_equal(ca, cb)
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(a.u, b.u)


def test_one_step_fp64_reading_by_hand():
    # one neuron, one step: the oracle's FP64 step from the FP32 state against numpy's float32
    # and float64 arithmetic of the same step (u = v - v_rest; brain.py:40-45)
    w = sp.csr_matrix((1, 1), dtype=np.float32)
    ob = OracleBrain(w)
    ob.set_bias([0], 7.3)
    ob.u[0, 0] = np.float32(5.123456)
    ob.g[0, 0] = np.float32(0.3)
    u, g, bias = ob.u[0, 0], ob.g[0, 0], np.float32(7.3)
    ds, dm = ob.c["decay_syn"], ob.c["decay_m"]
    g1 = np.float32(g * ds)
    u1 = np.float32(u + np.float32(np.float32(np.float32(g1 + bias) - u) * dm))
    gd = float(g) * float(ds)
    ud = float(u) + (gd + float(bias) - float(u)) * float(dm)
    ob.run_steps(1, one_step=True)
    rep = ob.last_report
    assert ob.u[0, 0] == u1 and ob.g[0, 0] == g1
    assert rep["step_max_dv"] == abs(ud - float(u1)) > 0
    assert rep["step_max_dg"] == abs(gd - float(g1))
    # dyadic state and constants that every float32 operation represents exactly: no difference
    ob2 = OracleBrain(w, OracleParams(tau_m=25.6, tau_syn=1e9, dt=0.1))   # decay_m = 2^-8, decay_syn = 1
    ob2.u[0, 0] = 1.5
    ob2.g[0, 0] = 2.0
    ob2.run_steps(1, one_step=True)
    assert ob2.c["decay_m"] == np.float32(2.0 ** -8) and ob2.c["decay_syn"] == 1.0
    assert ob2.last_report["step_max_dv"] == 0.0 and ob2.last_report["step_max_dg"] == 0.0


def test_fp32_membrane_step_within_1e5_mv_of_fp64_at_full_size(full_fly):
    # North-star tolerance, read per step: every step, the FP32 state advanced once in FP64 (the
    # same integer arrivals, kicks and refractory state) differs from the FP32 step by less than
    # 1e-5 mV, over all 138,639 neurons and 1,000 steps (measured: 4.05e-6 mV, no threshold
    # decision that differs).  Reported beside it, not bounded by the 1e-5 reading: the drift of
    # a separate FP64 trajectory that follows the FP32 spikes (measured max 1.12e-5 mV in the
    # decision band -64 <= v < -45 mV, RMS 4.9e-7 mV, one flipped decision in 1.4e8 neuron-steps).
    fly = full_fly
    ob = OracleBrain(fly.weights)
    ob.set_input(np.r_[fly.input_idx, fly.drive_idx], np.r_[fly.input_rate, 60.0])
    ob.enable_shadow()
    c = ob.run_steps(1000, one_step=True, nthreads=4)
    rep = ob.last_report
    assert c.sum() > 100_000
    assert rep["step_max_dv"] < 1e-5, rep
    assert rep["step_flips"] == 0, rep
    # trajectory drift (a regression bound on the measured figures above)
    d = np.abs(ob.u.astype(np.float64) - ob.u64)
    band = (ob.u64 >= -12) & (ob.u64 < float(ob.c["u_th"]))
    assert np.sqrt((d[band] ** 2).mean()) < 1e-6
    assert rep["max_dv_sub"] < 2e-5 and rep["flips"] <= 2, rep
