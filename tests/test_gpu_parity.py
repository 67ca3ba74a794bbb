"""GPU parity: the HIP engine against the CPU oracle, bit for bit.

Restates the reference's GPU tests (flyzero tests/test_gpu.py:39-72,168-183) with the oracle in
the role of the reference's CPU ``Brain``, and adds raster/state identity on the FlyWire-shaped
workload.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from helpers import fly_inputs, poisson_hits, ref_weights

pytestmark = pytest.mark.gpu

from flyhip import BatchBrain, Brain, LIFParams  # noqa: E402
from flyhip.connectome import flywire_like, smooth_field  # noqa: E402
from oracle.oracle import OracleBrain  # noqa: E402

FRAME_MS = 1000.0 / 60.0988


def _np(t):
    return t.detach().cpu().numpy()


def _oracle_counts(w, dt=0.25, pw=None, plastic_pos=None, ms=50, frac_bits=None):
    ob = OracleBrain(w, LIFParams(dt=dt), batch=1 if pw is None else len(pw), plastic_pos=plastic_pos,
                     frac_bits=frac_bits)
    if pw is not None:
        ob.pw[...] = pw
    ob.v[:, :40] = -40.0
    return ob.run(ms)


ENGINES = ["push", "tile"]


@pytest.mark.parametrize("engine", ENGINES)
def test_gpu_matches_oracle_without_noise(engine):
    # tests/test_gpu.py:39-47
    w = ref_weights()
    ref = _oracle_counts(w)[0]
    gb = BatchBrain(w, LIFParams(dt=0.25), batch=3, engine=engine)
    gb.v[:, :40] = -40.0
    out = _np(gb.run(50))
    assert ref.sum() > 50
    for k in range(3):
        np.testing.assert_array_equal(out[k], ref)


@pytest.mark.parametrize("engine", ENGINES)
def test_each_fly_has_its_own_plastic_synapses(engine):
    # tests/test_gpu.py:50-62
    w = ref_weights()
    ws = sp.csr_matrix(w, dtype=np.float32)
    ws.sort_indices()
    pos = np.flatnonzero(ws.indices < 30)          # synapses onto neurons 0..29
    gb = BatchBrain(w, LIFParams(dt=0.25), batch=2, plastic_pos=pos, engine=engine)
    changed = gb.data.copy()
    changed[pos] = np.abs(changed[pos]) * 3
    gb.pw[1] = torch.from_numpy(changed[pos]).cuda()
    gb.v[:, :40] = -40.0
    out = _np(gb.run(50))
    pw = np.stack([gb.w0, changed[pos]])
    ref = _oracle_counts(w, pw=pw, plastic_pos=pos, frac_bits=gb.frac_bits)
    np.testing.assert_array_equal(out, ref)
    assert not np.array_equal(out[0], out[1])


@pytest.mark.parametrize("engine", ENGINES)
def test_plastic_path_equals_fixed_path_for_whole_synapse_counts(engine):
    # with whole synapse counts the plastic weights quantise exactly, so a fly whose plastic
    # weights are untouched spikes exactly like a brain without plastic synapses
    w = ref_weights()
    w.data = np.rint(w.data)
    ws = sp.csr_matrix(w, dtype=np.float32)
    ws.sort_indices()
    pos = np.flatnonzero(ws.indices < 150)
    gb = BatchBrain(w, LIFParams(dt=0.25), batch=1, plastic_pos=pos, engine=engine)
    assert gb.frac_bits == 8 and gb.w_bits == 32        # sub-synapse plastic weights; fixed sums exact
    gb.v[:, :40] = -40.0
    out = _np(gb.run(50))
    np.testing.assert_array_equal(out, _oracle_counts(w))


@pytest.mark.parametrize("engine", ENGINES)
def test_poisson_input_rate(engine):
    # tests/test_gpu.py:65-72; 100 Hz with 2.2 ms dead time after each spike
    w = sp.csr_matrix((10, 10), dtype=np.float32)
    gb = BatchBrain(w, LIFParams(dt=0.25), batch=2, seed=5, engine=engine)
    gb.set_input(np.arange(10), 100.0)
    c = _np(gb.run(1000))
    assert 70 < c.mean() < 105
    assert not np.array_equal(c[0], c[1])
    ob = OracleBrain(w, LIFParams(dt=0.25), batch=2, seed=5)
    ob.set_input(np.arange(10), 100.0)
    np.testing.assert_array_equal(c, ob.run(1000))


@pytest.mark.parametrize("engine", ENGINES)
def test_long_plastic_rows_cover_every_synapse(engine):
    # tests/test_gpu.py:168-183: every synapse plastic, rows of >300 synapses
    rng = np.random.default_rng(4)
    w = sp.random(600, 600, density=0.6, random_state=4, format="csr")
    w.data = rng.normal(5.0, 3, w.nnz)
    pos = np.arange(w.nnz)
    gb = BatchBrain(w, LIFParams(dt=0.25), batch=2, plastic_pos=pos, engine=engine)
    assert np.diff(gb.indptr).max() > 300
    gb.pw[1] *= 0.5
    gb.v[:, :40] = -40.0
    out = _np(gb.run(30))
    ob = OracleBrain(w, LIFParams(dt=0.25), batch=2, plastic_pos=pos, frac_bits=gb.frac_bits)
    ob.pw[1] *= 0.5
    ob.v[:, :40] = -40.0
    ref = ob.run(30)
    assert ref[0].sum() > 50
    np.testing.assert_array_equal(out, ref)
    assert not np.array_equal(out[0], out[1])


@pytest.fixture(scope="module")
def small_fly():
    return flywire_like(n=20_000, n_edges=20_000 * 109, seed=11)


def _pair(fly, batch=1, seed=0, plastic_pos=None, engine="push", tile=1024):
    p = LIFParams()
    gb = BatchBrain(fly.weights, p, batch=batch, seed=seed, plastic_pos=plastic_pos, engine=engine, tile=tile)
    ob = OracleBrain(fly.weights, p, batch=batch, seed=seed, plastic_pos=plastic_pos, frac_bits=gb.frac_bits)
    idx, rate = fly_inputs(fly)
    gb.set_input(idx, rate)
    ob.set_input(idx, rate)
    return gb, ob


def _assert_state_equal(gb, ob):
    np.testing.assert_array_equal(_np(gb.u), ob.u)
    np.testing.assert_array_equal(_np(gb.g), ob.g)
    np.testing.assert_array_equal(_np(gb.ref), ob.ref)
    if gb.engine == "push":      # the tile engine keeps in-flight spikes as lists, not a ring
        np.testing.assert_array_equal(_np(gb.ring), ob.ring)


@pytest.mark.parametrize("engine,shape", [(e, {}) for e in ENGINES] +
                         [("tile", dict(upd_threads=256, tile_source=1, acc_bits=16)),
                          ("tile", dict(upd_threads=256, tile_source=1, acc_bits=32))])
def test_raster_identical_with_precomputed_hits(small_fly, engine, shape):
    gb, ob = _pair(small_fly, batch=2, engine=engine)
    if shape:
        gb.set_tune(**shape)             # the bucket source with two 16-bit step sums per word, or one int
    steps = 400
    hits = poisson_hits(ob.p_ext[0], steps, 2, seed=3)
    cg, rg = gb.run_steps(steps, hits=torch.from_numpy(hits.view(np.int32)).cuda(), raster=True)
    co, ro = ob.run_steps(steps, hits=hits, raster=True)
    rg = _np(rg).view(np.uint32)
    assert co.sum() > 1000
    first = np.flatnonzero((rg != ro).any(axis=(1, 2)))
    assert len(first) == 0, f"first divergent step {first[:1]}"
    np.testing.assert_array_equal(_np(cg), co)
    _assert_state_equal(gb, ob)


@pytest.mark.parametrize("engine", ENGINES)
def test_hash_rng_frames_graph_and_eager_match_oracle(small_fly, engine):
    gb, ob = _pair(small_fly, batch=3, seed=9, engine=engine)
    ge, _ = _pair(small_fly, batch=3, seed=9, engine=engine)
    xy = np.r_[small_fly.input_xy, [[0.5, 0.5]]]
    for frame in range(6):
        rate = np.r_[smooth_field(small_fly.input_xy, frame), np.float32(60.0)]
        idx = np.r_[small_fly.input_idx, small_fly.drive_idx]
        for b in (gb, ob, ge):
            b.set_input(idx, rate)
        cg = _np(gb.run(FRAME_MS))                 # graph replay
        ce = _np(ge.run(FRAME_MS, graph=False))    # eager launches
        co = ob.run(FRAME_MS)
        np.testing.assert_array_equal(cg, co, err_msg=f"frame {frame}")
        np.testing.assert_array_equal(ce, co, err_msg=f"frame {frame}")
    assert xy.shape[0] == len(small_fly.input_idx) + 1
    _assert_state_equal(gb, ob)
    assert gb.t == ob.t and gb.step_no == ob.step_no


@pytest.mark.parametrize("engine", ENGINES)
def test_a_brain_of_a_slice_of_flies_draws_that_slice_s_noise(small_fly, engine):
    # fly_base: flies [2, 4) of a 4-fly batch in a brain of their own get the same Poisson draws
    # (hash key seed + fly, neuron key fly * N + i), hence the same spikes
    whole, _ = _pair(small_fly, batch=4, seed=7, engine=engine)
    lo, _ = _pair(small_fly, batch=2, seed=7, engine=engine)
    hi, _ = _pair(small_fly, batch=2, seed=7, engine=engine)
    hi.fly_base = 2
    for frame in range(3):
        cw = _np(whole.run(FRAME_MS))
        np.testing.assert_array_equal(_np(lo.run(FRAME_MS)), cw[:2], err_msg=f"frame {frame}")
        np.testing.assert_array_equal(_np(hi.run(FRAME_MS)), cw[2:], err_msg=f"frame {frame}")
    assert cw.sum() > 0
    np.testing.assert_array_equal(_np(hi.u), _np(whole.u[2:]))


@pytest.mark.parametrize("engine", ENGINES)
def test_retired_flies_are_skipped_and_others_unaffected(small_fly, engine):
    gb, ob = _pair(small_fly, batch=4, seed=2, engine=engine)
    gb.run(FRAME_MS)
    ob.run(FRAME_MS)
    v_before = _np(gb.u[1]).copy()
    gb.alive[1] = 0
    ob.alive = np.array([1, 0, 1, 1], np.uint8)
    cg = _np(gb.run(FRAME_MS))
    co = ob.run(FRAME_MS)
    assert cg[1].sum() == 0
    np.testing.assert_array_equal(_np(gb.u[1]), v_before)
    np.testing.assert_array_equal(cg[[0, 2, 3]], co[[0, 2, 3]])
    # respawn: reset the fly and revive it
    gb.reset([1])
    gb.alive[1] = 1
    ob.u[1] = 0.0
    ob.g[1] = 0
    ob.ref[1] = 0
    ob.ring[1] = 0
    ob.alive = None
    np.testing.assert_array_equal(_np(gb.run(FRAME_MS)), ob.run(FRAME_MS))


def test_brain_single_fly_api():
    # tests/test_flyzero.py:16-24 through the drop-in Brain
    w = np.zeros((2, 2))
    w[0, 1] = 100
    b = Brain(sp.csr_matrix(w))
    b.set_input([0], 200.0)
    counts = b.run(200)
    assert isinstance(counts, np.ndarray) and counts.dtype == np.int32
    assert counts[0] > 10 and counts[1] > 5
    b = Brain(sp.csr_matrix(-w))
    b.set_input([0], 200.0)
    assert b.run(200)[1] == 0


def test_brain_propagate_matches_dense_matrix():
    # tests/test_flyzero.py:27-34
    w = sp.random(50, 50, density=0.2, random_state=1, format="csr") * 10
    b = Brain(w)
    spiking = np.array([3, 7, 20, 49])
    out = np.zeros(50, np.float32)
    b._propagate(spiking, out)
    expected = np.asarray(w[spiking].sum(0)).ravel() * b.p.w_syn
    # integer sums of weights quantised to 2**-frac_bits synapses: half a unit per term
    atol = len(spiking) * 0.5 * 2.0 ** -b._bb.frac_bits * b.p.w_syn
    np.testing.assert_allclose(out, expected, rtol=1e-5, atol=atol)


def test_brain_weight_edits_reach_the_engine():
    w = np.zeros((3, 3))
    w[0, 1] = 100
    b = Brain(sp.csr_matrix(w))
    b.set_input([0], 200.0)
    assert b.run(100)[2] == 0
    b.data[0] = 0.0                 # remove 0 -> 1 (mV layout, like RewardPlasticity edits)
    b.reset()
    b.set_input([0], 200.0)
    assert b.run(100)[1] == 0


@pytest.mark.parametrize("engine", ENGINES)
def test_autotuned_shapes_keep_bit_exact_results(small_fly, engine):
    from flyhip import autotune

    gb, ob = _pair(small_fly, batch=2, seed=4, engine=engine)
    shape = autotune.tune(gb, use_cache=False, reps=1)
    assert shape["upd_threads"] % 64 == 0 and shape.get("push_chunk", 64) >= 64
    for _ in range(3):
        np.testing.assert_array_equal(_np(gb.run(FRAME_MS)), ob.run(FRAME_MS))
    _assert_state_equal(gb, ob)


@pytest.mark.parametrize("shape", [dict(upd_threads=1024, push_threads=64, push_chunk=64, push_grid=7),
                                   dict(upd_threads=64, push_threads=1024, push_chunk=4096, push_grid=3000,
                                        block_steps=5)])
def test_extreme_launch_shapes_are_exact(small_fly, shape):
    gb, ob = _pair(small_fly, batch=3, seed=6)
    gb.set_tune(**shape)
    for _ in range(2):
        np.testing.assert_array_equal(_np(gb.run(FRAME_MS)), ob.run(FRAME_MS))
    _assert_state_equal(gb, ob)


@pytest.mark.parametrize("tile,shape", [(512, dict(upd_threads=64, block_steps=5)),
                                        (2048, dict(upd_threads=256)), (256, dict(upd_threads=256)),
                                        (1024, dict(upd_threads=1024, block_steps=9, tile_stage=4)),
                                        (1024, dict(upd_threads=128, tile_stage=4)),
                                        (1024, dict(upd_threads=256, tile_source=1)),
                                        (512, dict(upd_threads=64, block_steps=5, tile_source=1)),
                                        (2048, dict(upd_threads=512, block_steps=9, tile_stage=4, tile_source=1)),
                                        (256, dict(upd_threads=256, block_steps=7, tile_source=1)),
                                        (768, dict(upd_threads=384, tile_source=1)),
                                        (1024, dict(upd_threads=256, tile_source=2)),
                                        (2048, dict(upd_threads=512, block_steps=9, tile_stage=4, tile_source=2)),
                                        (1024, dict(upd_threads=512, tile_source=1, block_parts=2)),
                                        (512, dict(upd_threads=64, block_steps=5, tile_source=1, block_parts=2)),
                                        (768, dict(upd_threads=384, block_steps=9, tile_source=1, block_parts=2)),
                                        (1024, dict(upd_threads=256, tile_stage=4, block_parts=2)),
                                        (256, dict(upd_threads=256, block_steps=7, block_parts=2)),
                                        (2048, dict(upd_threads=512, block_steps=9, tile_source=2, block_parts=2)),
                                        (640, dict(upd_threads=640, tile_stage=4)),
                                        # one int per step's arrival sum, and two 16-bit sums per int
                                        (1024, dict(upd_threads=256, tile_source=1, acc_bits=32)),
                                        (512, dict(upd_threads=64, block_steps=5, tile_source=1, acc_bits=32)),
                                        (256, dict(upd_threads=256, block_steps=7, tile_source=1, acc_bits=32)),
                                        (768, dict(upd_threads=384, tile_source=1, acc_bits=16)),
                                        (2048, dict(upd_threads=256, block_steps=7, tile_source=1, acc_bits=16)),
                                        (1024, dict(upd_threads=1024, block_steps=5, tile_source=1, acc_bits=16))])
def test_tile_engine_shapes_are_exact(small_fly, tile, shape):
    # shapes that the LDS budget admits, incl. 8 neurons per thread, 5-step blocks and tiles
    # that are not powers of two
    fly = small_fly
    idx = np.r_[fly.input_idx, fly.drive_idx]
    gb, ob = _pair(fly, batch=2, seed=8, engine="tile", tile=tile)
    gb.set_tune(**shape)
    for f in range(3):
        rate = np.r_[smooth_field(fly.input_xy, f), np.float32(60.0)]
        gb.set_input(idx, rate)
        ob.set_input(idx, rate)
        np.testing.assert_array_equal(_np(gb.run(FRAME_MS)), ob.run(FRAME_MS))
    _assert_state_equal(gb, ob)


def test_two_sums_per_word_needs_its_bound_and_shape(small_fly):
    # 16-bit step sums are refused where they are not compiled (two-part blocks, the table
    # source) and where a neuron's summed |input weight| could leave int16; the automatic
    # choice then keeps one int per step, and the results stay the oracle's
    gb, _ = _pair(small_fly, batch=1, engine="tile", tile=1024)
    for bad in (dict(tile_source=1, block_parts=2, acc_bits=16), dict(tile_source=0, acc_bits=16),
                dict(tile_source=1, acc_bits=8)):
        gb.set_tune(upd_threads=256, **bad)
        with pytest.raises(ValueError, match="acc_bits"):
            gb.run(FRAME_MS, graph=False)
    rng = np.random.default_rng(17)
    n, m = 3000, 90000
    rows = np.r_[rng.integers(0, n, m), np.arange(1200)]
    cols = np.r_[rng.integers(0, n, m), np.full(1200, 7)]  # neuron 7: 1,200 more inputs of weight 30
    vals = np.r_[np.round(rng.random(m) * 12.0 - 4.0), np.full(1200, 30.0)].astype(np.float32)
    w = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    w.eliminate_zeros()
    assert np.abs(w[:, 7]).sum() >= 32768
    p = LIFParams()
    gb = BatchBrain(w, p, batch=2, seed=4, engine="tile", tile=512)
    ob = OracleBrain(w, p, batch=2, seed=4)
    idx = np.arange(0, n, 7)
    rate = np.full(len(idx), 80.0, np.float32)
    gb.set_tune(upd_threads=256, tile_source=1, acc_bits=16)
    gb.set_input(idx, rate)
    with pytest.raises(ValueError, match="32768"):
        gb.run(FRAME_MS, graph=False)
    gb.set_tune(upd_threads=256, tile_source=1)
    ob.set_input(idx, rate)
    for _ in range(2):
        np.testing.assert_array_equal(_np(gb.run(FRAME_MS)), ob.run(FRAME_MS))
    _assert_state_equal(gb, ob)


def test_tile_engine_rejects_shapes_over_the_lds_budget(small_fly):
    gb, _ = _pair(small_fly, batch=1, engine="tile", tile=2048)
    gb.set_tune(upd_threads=1024, tile_stage=4)       # 19 x 2048 x 4 B + 3 x 4 x 1024 x 4 B > 160 KB
    with pytest.raises(ValueError, match="LDS"):
        gb.run(FRAME_MS, graph=False)


@pytest.mark.parametrize("engine", ENGINES)
def test_full_connectome_raster_1000_steps_and_one_second_of_frames(full_fly, engine):
    # the full-size workload: identical rasters over the first 1,000 steps on the same
    # precomputed Poisson input, then 1 s of hash-RNG frames (rates resampled every 166
    # steps) with identical per-neuron counts
    fly = full_fly
    gb, ob = _pair(fly, batch=1, engine=engine)
    hits = poisson_hits(ob.p_ext[0], 1000, 1, seed=21)
    cg, rg = gb.run_steps(1000, hits=torch.from_numpy(hits.view(np.int32)).cuda(), raster=True)
    co, ro = ob.run_steps(1000, hits=hits, raster=True)
    rg = _np(rg).view(np.uint32)
    assert co.sum() > 10_000
    first = np.flatnonzero((rg != ro).any(axis=(1, 2)))
    assert len(first) == 0, f"first divergent step {first[:1]}"
    np.testing.assert_array_equal(_np(cg), co)
    idx = np.r_[fly.input_idx, fly.drive_idx]
    tot_g = np.zeros(fly.n, np.int64)
    tot_o = np.zeros(fly.n, np.int64)
    for frame in range(60):                               # 60 x 166 = 9,960 steps
        rate = np.r_[smooth_field(fly.input_xy, frame), np.float32(60.0)]
        gb.set_input(idx, rate)
        ob.set_input(idx, rate)
        tot_g += _np(gb.run(FRAME_MS))[0]
        tot_o += ob.run(FRAME_MS)[0]
    assert tot_o.sum() > 100_000
    np.testing.assert_array_equal(tot_g, tot_o)
    _assert_state_equal(gb, ob)


@pytest.mark.parametrize("parts", [1, 2])
@pytest.mark.parametrize("pull", [0, 1])
@pytest.mark.parametrize("source", [0, 1, 2])
def test_tile_sources_give_identical_rasters_and_plastic_arrivals(small_fly, source, pull, parts):
    # per-tile buckets versus segment-table lookups: same rasters, counts and state as the
    # oracle, with plastic synapses (pushed through the ring) and a retired fly
    fly = small_fly
    rng = np.random.default_rng(13)
    pos = np.sort(rng.choice(fly.weights.nnz, 3000, replace=False))
    gb, ob = _pair(fly, batch=3, seed=4, plastic_pos=pos, engine="tile", tile=1024)
    gb.set_tune(upd_threads=512, tile_source=source, plastic_pull=pull, block_parts=parts)
    gb.pw[1] *= 1.5
    ob.pw[1] *= 1.5
    gb.alive[2] = 0
    ob.alive = np.array([1, 1, 0], np.uint8)
    steps = 300
    hits = poisson_hits(ob.p_ext[0], steps, 3, seed=8)
    cg, rg = gb.run_steps(steps, hits=torch.from_numpy(hits.view(np.int32)).cuda(), raster=True)
    co, ro = ob.run_steps(steps, hits=hits, raster=True)
    rg = _np(rg).view(np.uint32)
    assert co[:2].sum() > 1000
    first = np.flatnonzero((rg[:, :2] != ro[:, :2]).any(axis=(1, 2)))
    assert len(first) == 0, f"first divergent step {first[:1]}"
    np.testing.assert_array_equal(_np(cg), co)
    for frame in range(3):
        np.testing.assert_array_equal(_np(gb.run(FRAME_MS)), ob.run(FRAME_MS), err_msg=f"frame {frame}")
    np.testing.assert_array_equal(_np(gb.u)[:2], ob.u[:2])
    np.testing.assert_array_equal(_np(gb.g)[:2], ob.g[:2])


@pytest.mark.parametrize("parts", [1, 2])
@pytest.mark.parametrize("pull", [0, 1])
@pytest.mark.parametrize("source", [0, 1, 2])
def test_plastic_arrivals_carry_the_weights_of_the_window_they_were_emitted_in(small_fly, source, pull, parts):
    # With plastic pulls the tile engine pulls a plastic synapse at its target when the arrival
    # falls in the window the spike was emitted in, and pushes it at spike time otherwise;
    # without, every plastic row is pushed at spike time.  Windows of 2 to 166 steps,
    # several shorter than n_delay = 18 (an arrival can land two windows later), eager and as a
    # graph, with new plastic weights before every window: each window's counts and the final
    # state equal the oracle's, which pushes every spike with the weights of its own window.
    fly = small_fly
    rng = np.random.default_rng(17)
    pos = np.sort(rng.choice(fly.weights.nnz, 6000, replace=False))
    gb, ob = _pair(fly, batch=2, seed=9, plastic_pos=pos, engine="tile", tile=1024)
    gb.set_tune(upd_threads=512, tile_source=source, plastic_pull=pull, block_parts=parts)
    base = ob.pw.copy()
    total = 0
    for wi, steps in enumerate([40, 5, 3, 17, 18, 19, 2, 166, 36, 9, 30, 166]):
        pw = (base * rng.uniform(0.0, 4.0, base.shape)).astype(np.float32)
        gb.pw.copy_(torch.from_numpy(pw))
        ob.pw[...] = pw
        cg = gb.run(FRAME_MS) if steps == 166 else gb.run_steps(steps)
        co = ob.run_steps(steps)
        np.testing.assert_array_equal(_np(cg), co, err_msg=f"window {wi} ({steps} steps)")
        total += int(co.sum())
    assert total > 2000
    _assert_state_equal(gb, ob)


@pytest.mark.parametrize("parts", [1, 2])
@pytest.mark.parametrize("t_ref", [0.5, 1.8])
@pytest.mark.parametrize("source", [0, 1, 2])
def test_tile_engine_refractory_shorter_than_or_equal_to_a_block(small_fly, t_ref, source, parts):
    # 18-step blocks: with t_ref = 0.5 ms (5 steps) a neuron can fire several times in one block
    # (the general step loop); with t_ref = 1.8 ms (18 steps) at most once, at the limit of the
    # once-per-block loop
    fly = small_fly
    p = LIFParams(t_ref=t_ref)
    gb = BatchBrain(fly.weights, p, batch=2, seed=6, engine="tile", tile=1024)
    ob = OracleBrain(fly.weights, p, batch=2, seed=6)
    idx, rate = fly_inputs(fly)
    for b in (gb, ob):
        b.set_input(idx, rate)
    gb.set_tune(upd_threads=512, block_steps=18, tile_source=source, block_parts=parts)
    steps = 360
    cg, rg = gb.run_steps(steps, raster=True)
    co, ro = ob.run_steps(steps, raster=True)
    rg = _np(rg).view(np.uint32)
    first = np.flatnonzero((rg != ro).any(axis=(1, 2)))
    assert len(first) == 0, f"first divergent step {first[:1]}"
    np.testing.assert_array_equal(_np(cg), co)
    _assert_state_equal(gb, ob)
    bits = np.unpackbits(ro.view(np.uint8), axis=2, bitorder="little")   # steps x flies x neurons
    per_block = bits.reshape(steps // 18, 18, 2, -1).sum(axis=1)
    if t_ref == 1.8:
        assert per_block.max() == 1
    else:
        assert per_block.max() >= 2


def test_tile_engine_rejects_a_bad_segment_table(small_fly):
    from flyhip import _lib

    gb = BatchBrain(small_fly.weights, LIFParams(), batch=1, engine="tile", tile=1024)
    bad = gb.d_tb.clone()
    bad[5, 3] += 1                                      # an edge moved to the wrong tile
    with pytest.raises(ValueError, match="segment table"):
        _lib.check(gb._lib.fz_set_tiles(gb._h, gb.tile, bad.shape[1] - 1, bad.data_ptr(), None, None),
                   "fz_set_tiles")


def _random_graph(n, deg, seed, hub_deg=0):
    # Dale's law per pre neuron, counts 1..4; optional hub row 0 with hub_deg targets
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(n, dtype=np.int64), deg)
    cols = rng.integers(0, n, rows.size)
    if hub_deg:
        rows = np.r_[rows, np.zeros(hub_deg, np.int64)]
        cols = np.r_[cols, rng.choice(n, hub_deg, replace=False)]
    sign = np.where(rng.random(n) < 0.55, 1.0, -1.0)
    vals = sign[rows] * rng.integers(1, 5, rows.size)
    w = sp.csr_matrix((vals.astype(np.float32), (rows, cols)), shape=(n, n))
    w.sum_duplicates()
    return w


@pytest.mark.parametrize("source", [0, 1, 2])
def test_tile_engine_row_longer_than_65535_edges(source):
    # the uint16 segment-table copy cannot hold this row: the int32 table is used instead
    w = _random_graph(80_000, 20, seed=31, hub_deg=70_000)
    assert np.diff(w.indptr).max() >= 65_536
    p = LIFParams()
    gb = BatchBrain(w, p, batch=2, seed=3, engine="tile", tile=1024)
    ob = OracleBrain(w, p, batch=2, seed=3)
    idx = np.r_[0, np.arange(1, 4000)]
    rate = np.r_[np.float32(400.0), np.full(3999, 30.0, np.float32)]
    for b in (gb, ob):
        b.set_input(idx, rate)
    gb.set_tune(upd_threads=512, tile_source=source)
    hub = 0
    for frame in range(3):
        co = ob.run(FRAME_MS)
        np.testing.assert_array_equal(_np(gb.run(FRAME_MS)), co, err_msg=f"frame {frame}")
        hub += int(co[:, 0].sum())
    assert hub > 0                                       # the hub row was pushed
    _assert_state_equal(gb, ob)


def test_tile_engine_with_more_than_2_24_edges():
    # bucket entries hold a 24-bit first-edge index: a larger graph is refused the bucket source
    # with a clear error and runs exactly on segment-table lookups
    w = _random_graph(60_000, 285, seed=32)
    assert w.nnz >= 1 << 24
    p = LIFParams()
    gb = BatchBrain(w, p, batch=1, seed=4, engine="tile", tile=1024)
    ob = OracleBrain(w, p, batch=1, seed=4)
    idx = np.arange(0, 6000)
    for b in (gb, ob):
        b.set_input(idx, 40.0)
    gb.set_tune(upd_threads=512, tile_source=1)
    with pytest.raises(ValueError, match="2\\^24"):
        gb.run(FRAME_MS, graph=False)
    gb.set_tune(upd_threads=512, tile_source=0)
    for frame in range(2):
        np.testing.assert_array_equal(_np(gb.run(FRAME_MS)), ob.run(FRAME_MS), err_msg=f"frame {frame}")
    _assert_state_equal(gb, ob)
