"""Host-side logic (CPU only): the tile engine's segment tables against a brute-force search, the
synthetic connectome against its stated construction, and the plastic-synapse selection."""

import numpy as np
import pytest
import scipy.sparse as sp

from flyhip.connectome import DN_TYPES, GROUPS, MOTOR_TYPES, flywire_like, plastic_positions, smooth_field
from flyhip.gpu import tile_tables


def _csr(n, density, seed):
    m = sp.random(n, n, density=density, format="csr", random_state=seed, dtype=np.float32)
    m.sort_indices()
    return m


@pytest.mark.parametrize("n,tile,density", [(1, 4, 1.0), (37, 8, 0.2), (200, 64, 0.05), (513, 512, 0.01)])
def test_tile_tables_match_brute_force(n, tile, density):
    w = _csr(n, density, seed=n)
    rng = np.random.default_rng(n)
    pl_post = rng.integers(0, n, size=max(1, n // 5))
    tb, ptg_ptr, ptg_idx = tile_tables(w.indptr, w.indices, n, tile, pl_post)
    nt = (n + tile - 1) // tile
    assert tb.shape == (n, nt + 1)
    for i in range(n):
        lo, hi = w.indptr[i], w.indptr[i + 1]
        for t in range(nt + 1):
            want = lo
            while want < hi and w.indices[want] < t * tile:
                want += 1
            assert tb[i, t] == want, (i, t)
        assert tb[i, 0] == lo and tb[i, nt] == hi
    # plastic posts grouped by tile, each once, sorted
    posts = np.unique(pl_post)
    assert np.array_equal(ptg_idx, posts)
    for t in range(nt):
        got = ptg_idx[ptg_ptr[t]:ptg_ptr[t + 1]]
        assert np.array_equal(got, posts[(posts >= t * tile) & (posts < (t + 1) * tile)])
    assert ptg_ptr[-1] == len(posts)


def test_tile_tables_with_empty_rows_and_no_plastic_posts():
    n = 20
    w = sp.csr_matrix((np.ones(1, np.float32), ([3], [19])), shape=(n, n))
    tb, ptg_ptr, ptg_idx = tile_tables(w.indptr, w.indices, n, 8, np.zeros(0, np.int64))
    assert (tb[:3] == 0).all() and (tb[4:] == 1).all()
    assert list(tb[3]) == [0, 0, 0, 1]     # the one edge lies in tile 2
    assert len(ptg_idx) == 0 and (ptg_ptr == 0).all()


def test_synthetic_connectome_follows_its_construction():
    n, e = 6000, 6000 * 109
    fly = flywire_like(n=n, n_edges=e, seed=5)
    w = fly.weights
    assert w.shape == (n, n) and w.has_sorted_indices
    assert abs(w.nnz - e) <= e // 1000                     # duplicates redrawn, few dropped
    rows = np.repeat(np.arange(n), np.diff(w.indptr))
    assert not (rows == w.indices).any()                  # no self-loops
    key = rows.astype(np.int64) * n + w.indices
    assert len(np.unique(key)) == w.nnz                   # no duplicate connections
    cnt = np.abs(w.data)
    assert (cnt >= 1).all() and (cnt <= 2400).all() and (cnt == np.round(cnt)).all()
    assert abs(cnt.mean() - 1 / 0.28) < 0.05              # geometric(0.28) mean
    # Dale's rule: one sign per presynaptic neuron, about 55 % excitatory
    sgn = np.sign(w.data)
    nz = np.diff(w.indptr) > 0
    first = sgn[w.indptr[:-1][nz]]
    assert (sgn == np.repeat(first, np.diff(w.indptr)[nz])).all()
    assert 0.5 < (first > 0).mean() < 0.6
    # the largest in-degree is near 5 % of N
    indeg = np.bincount(w.indices, minlength=n)
    assert 0.03 * n < indeg.max() < 0.07 * n
    # inputs, drive and descending neurons are disjoint
    assert len(np.unique(fly.input_idx)) == len(fly.input_idx) == round(12_246 * n / 138_639)
    assert fly.drive_idx not in set(fly.input_idx)
    assert ((fly.input_rate >= 0) & (fly.input_rate <= 200)).all()
    dn = np.concatenate([fly.find(t) for t in DN_TYPES])
    assert len(np.unique(dn)) == len(dn) == 10 * len(DN_TYPES)
    assert not set(dn) & (set(fly.input_idx) | {fly.drive_idx})
    for t in DN_TYPES:   # balanced sides
        assert len(fly.find(t, "left")) == len(fly.find(t, "right")) == 5
    assert all(len(g) > 0 for g in fly.group_members()) and len(GROUPS) == 8


def test_same_seed_same_connectome_other_seed_differs():
    a = flywire_like(n=2000, n_edges=2000 * 50, seed=1)
    b = flywire_like(n=2000, n_edges=2000 * 50, seed=1)
    c = flywire_like(n=2000, n_edges=2000 * 50, seed=2)
    assert (a.weights != b.weights).nnz == 0 and np.array_equal(a.input_idx, b.input_idx)
    assert (a.weights != c.weights).nnz > 0


def test_plastic_positions_are_every_synapse_onto_motor_types():
    fly = flywire_like(n=3000, n_edges=3000 * 60, seed=9)
    pos = plastic_positions(fly)
    w = fly.weights
    targets = set(np.concatenate([fly.find(t) for t in MOTOR_TYPES]).tolist())
    onto = np.array([c in targets for c in w.indices])
    assert np.array_equal(pos, np.flatnonzero(onto))


def test_smooth_field_is_bounded_and_drifts():
    xy = np.random.default_rng(0).random((500, 2))
    f0, f1 = smooth_field(xy, 0), smooth_field(xy, 30)
    assert f0.dtype == np.float32 and ((f0 >= 0) & (f0 <= 200)).all()
    assert np.array_equal(f0, smooth_field(xy, 0))
    assert not np.array_equal(f0, f1)


def test_motor_buttons_thresholds_and_taps():
    # flyzero/batch.py:114-138 (a BatchMotor without a device: buttons() and _tap() are host logic)
    from flyhip.batch import READOUT, BatchMotor, MotorParams

    m = BatchMotor.__new__(BatchMotor)
    m.p = MotorParams()
    m.acc = np.zeros((1, 2))
    r = np.zeros(len(READOUT), np.float32)
    r[READOUT.index("steer_left")] = 30.0           # left turn
    r[READOUT.index("wing_left")] = 5.0              # wings push the other way by 5 Hz
    r[READOUT.index("accelerate")] = 6.0
    r[READOUT.index("boost")] = 0.5
    b = m.buttons(r)
    assert b["LEFT"] and not b["RIGHT"] and b["B"] and not b["A"] and not b["Y"]
    r[READOUT.index("steer_left")] = 14.0            # 14 - 5 = 9 Hz: below the 10 Hz threshold
    assert not m.buttons(r)["LEFT"]
    # taps: duty (|x| - 10) / 150 accumulates; a tap fires each time the accumulator reaches 0.5
    m.p = MotorParams(taps=True)
    r[:] = 0
    r[READOUT.index("steer_left")] = 85.0            # duty 0.5: a tap every other frame
    taps = [m.buttons(r)["LEFT"] for _ in range(6)]
    assert taps == [True, False, True, False, True, False]
    r[:] = 0
    r[READOUT.index("lean_right")] = 80.0            # duty -1: lean right every frame
    assert all(m.buttons(r)["R"] for _ in range(3))


def test_motor_masks_for_all_slots_equal_one_slot_at_a_time():
    from flyhip.batch import BUTTONS, BatchMotor, MotorParams

    rng = np.random.default_rng(2)
    B = 40
    for taps in (False, True):
        a = BatchMotor.__new__(BatchMotor)
        b = BatchMotor.__new__(BatchMotor)
        for m in (a, b):
            m.p = MotorParams(taps=taps, steer_bias=1.5)
            m.acc = np.zeros((B, 2))
        for _ in range(20):
            r = (rng.random((B, 9)) * 120 - 20).astype(np.float32)
            together = a.masks(r)
            one = np.array([[d[key] for key in BUTTONS] for d in (b.buttons(r[k], k) for k in range(B))])
            np.testing.assert_array_equal(together, one)
        np.testing.assert_array_equal(a.acc, b.acc)
