"""fz_run_static (no allocation, static device scratch) against fz_run, through the C ABI.

Two calls in a row on one GPU share the static array; the second has a smaller input (fewer
flies of a smaller connectome), so stale contents of the first call must not leak into it.
"""

import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from helpers import fly_inputs

pytestmark = pytest.mark.gpu

from flyhip import BatchBrain, LIFParams, _lib  # noqa: E402
from flyhip.connectome import flywire_like  # noqa: E402

FRAME_MS = 1000.0 / 60.0988


def _np(t):
    return t.detach().cpu().numpy()


def _run_static(bb, steps, hits=None, raster=False):
    """One window through fz_run_static on the brain's own bound state."""
    inp = bb._inputs()
    ras = None
    if raster:
        ras = torch.zeros((steps, bb.B, (bb.n + 31) // 32), dtype=torch.int32, device=bb.device)
    rc = bb._lib.fz_run_static(bb._h, ctypes.byref(inp), steps, _lib.ptr(hits), _lib.ptr(ras),
                               ctypes.byref(bb.tune), _lib.stream_handle(bb.device))
    _lib.check(rc, "fz_run_static")
    bb._advance(steps)
    return (bb.counts.clone(), ras) if raster else bb.counts.clone()


def _pair(fly, batch, seed, engine="push", **tune):
    out = []
    for _ in range(2):
        bb = BatchBrain(fly.weights, LIFParams(), batch=batch, seed=seed, engine=engine)
        idx, rate = fly_inputs(fly)
        bb.set_input(idx, rate)
        if tune:
            bb.set_tune(**tune)
        out.append(bb)
    return out


def _assert_same_state(a, b):
    for name in ("u", "g", "ref", "ring", "counts"):
        np.testing.assert_array_equal(_np(getattr(a, name)), _np(getattr(b, name)), err_msg=name)


def test_static_entry_point_equals_fz_run_on_two_calls_in_a_row():
    big = flywire_like(n=20_000, n_edges=20_000 * 109, seed=11)
    small = flywire_like(n=5_000, n_edges=5_000 * 60, seed=12)
    steps = 166
    for fly, batch, seed in ((big, 3, 4), (small, 2, 9)):          # the second call is smaller
        ref, st = _pair(fly, batch, seed)
        for frame in range(2):
            want = _np(ref.run(FRAME_MS, graph=False))               # fz_run (allocated work list)
            got = _np(_run_static(st, steps))
            assert want.sum() > 0
            np.testing.assert_array_equal(got, want, err_msg=f"n={fly.n} frame {frame}")
        _assert_same_state(ref, st)
        st.check()


def test_static_entry_point_raster_and_hits_equal_fz_run():
    fly = flywire_like(n=8_000, n_edges=8_000 * 109, seed=3)
    ref, st = _pair(fly, 2, 5, upd_threads=128, push_chunk=128, block_steps=6)
    steps = 120
    hw = (len(ref.ext) + 31) // 32
    hits = torch.from_numpy(np.random.default_rng(1).integers(-2 ** 31, 2 ** 31, (steps, 2, hw), dtype=np.int64)
                            .astype(np.int32) & 0x01010101).cuda()
    cw, rw = ref.run_steps(steps, hits=hits, raster=True)
    cs, rs = _run_static(st, steps, hits=hits, raster=True)
    np.testing.assert_array_equal(_np(cs), _np(cw))
    np.testing.assert_array_equal(_np(rs), _np(rw))
    _assert_same_state(ref, st)


def test_static_entry_point_tile_engine_table_source_only():
    fly = flywire_like(n=8_000, n_edges=8_000 * 109, seed=3)
    ref, st = _pair(fly, 2, 6, engine="tile", upd_threads=256, tile_source=0)
    for _ in range(2):
        np.testing.assert_array_equal(_np(_run_static(st, 166)), _np(ref.run(FRAME_MS, graph=False)))
    st.set_tune(upd_threads=256, tile_source=1)
    with pytest.raises(ValueError, match="bucket"):
        _run_static(st, 166)


def test_static_entry_point_refuses_a_batch_over_the_limit():
    # the work list of 400,000 neurons is 8 MB per fly: 5 flies exceed the 32 MiB array
    n = 400_000
    w = sp.csr_matrix((np.ones(10, np.float32), (np.arange(10), np.arange(1, 11))), shape=(n, n))
    bb = BatchBrain(w, LIFParams(), batch=5)
    with pytest.raises(ValueError, match="static scratch"):
        _run_static(bb, 10)
    small = BatchBrain(w, LIFParams(), batch=1)                   # one fly fits
    _run_static(small, 10)
    assert _np(small.counts).sum() == 0
