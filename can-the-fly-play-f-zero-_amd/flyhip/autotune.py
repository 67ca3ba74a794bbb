"""Launch-shape autotuner for the LIF engine.

Times candidate ``fz_tune`` shapes (update workgroup size, push workgroup size, push edge
chunk, push grid) with HIP events on real windows of the caller's own workload, replayed as
the same HIP graphs ``run`` uses, then restores the brain's state, so tuning never advances
the simulation.  Search is coordinate descent:
the update kernel's shape first, then the push kernel's, each over a small grid of shapes that
respect the kernels' limits (multiples of 64 threads, chunk >= 64 edges).  The tile engine
takes a full grid instead (and, with plastic synapses, each shape with plastic rows pushed and
pulled).  Results are cached per (device, engine, neurons, connections, flies, plastic synapses,
delay steps, activity) in memory and in a JSON file; the activity tag is log2(1 + expected
Poisson kicks per step and fly) and whether any bias drive is on, so a shape tuned on a silent
brain is not reused for a driven one.
"""

from __future__ import annotations

import json
import math
import os

import torch

from . import _lib

CACHE = os.path.join(os.path.expanduser("~"), ".cache", "flyhip", "autotune.json")
_mem: dict = {}

UPD_THREADS = (256, 512, 1024)
PUSH_THREADS = (64, 128, 256, 512)
PUSH_CHUNK = (64, 128, 256, 512, 1024)
PUSH_GRID_PER_CU = (2, 4, 8, 16)


def _activity(bb) -> str:
    kicks = float(bb.p_ext.sum()) / bb.B if bb.p_ext.numel() else 0.0
    bias = bool((bb.bias_ext != 0).any()) if bb.bias_ext.numel() else False
    return f"{int(math.log2(1.0 + kicks))}{'b' if bias else ''}"


def _key(bb) -> str:
    name = torch.cuda.get_device_name(bb.device)
    eng = bb.engine + (str(bb.tile) if bb.engine == "tile" else "")
    return (f"{name}|{eng}|n={bb.n}|nnz={int(bb.d_col.numel())}|B={bb.B}|npl={len(bb.plastic_pos)}"
            f"|nd={bb.n_delay}|act={_activity(bb)}")


def _state(bb):
    extra = (bb.spk_idx, bb.spk_cnt) if bb.spk_idx is not None else ()
    return (bb.u, bb.g, bb.ref, bb.ring, bb.counts) + extra


def _snapshot(bb):
    return [t.clone() for t in _state(bb)], bb.clock.clone(), (bb._t, bb._step_no)


def _restore(bb, snap):
    tensors, clock, (t, step_no) = snap
    for dst, src in zip(_state(bb), tensors):
        dst.copy_(src)
    # clock[2] (block parity of the engine's work-list counters) keeps running: only the
    # simulated time and the RNG step go back
    bb.clock[0:2].copy_(clock[0:2])
    bb._t, bb._step_no = t, step_no


def _time(bb, shape: dict, steps: int, reps: int, snap) -> float:
    """Best of ``reps`` windows replayed as the HIP graph that ``run`` uses (the graph's
    concurrent fly-half chains overlap kernels that eager launches serialise, so eager timings
    can rank the shapes wrongly), timed with HIP events on the engine's stream; inf for a shape
    the engine rejects (e.g. LDS over the limit)."""
    bb.set_tune(**shape)
    ms = steps * bb.p.dt
    try:
        _restore(bb, snap)
        bb.run(ms, copy=False)  # graph capture and first-launch costs stay out of the timing
    except ValueError:
        return float("inf")
    best = float("inf")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(reps):
        _restore(bb, snap)
        e0.record()
        bb.run(ms, copy=False)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def tune(bb, duration_ms: float = 1000.0 / 60.0988, reps: int = 2, use_cache: bool = True,
         verbose: bool = False) -> dict:
    """Pick and apply the fastest launch shape for ``bb``'s current workload.  Returns it."""
    key = _key(bb)
    if use_cache:
        hit = _mem.get(key)
        if hit is None and os.path.exists(CACHE):
            try:
                hit = json.load(open(CACHE)).get(key)
            except (OSError, ValueError):
                hit = None
        if hit is not None:
            bb.set_tune(**hit)
            return hit
    steps = bb.steps_of(duration_ms)
    n_cu = torch.cuda.get_device_properties(bb.device).multi_processor_count
    snap = _snapshot(bb)
    # warm-up window (first launches pay code-object loading)
    _time(bb, {}, steps, 1, snap)
    if bb.engine == "tile":
        # full grid (small): threads per workgroup (1, 2, 4 or 8 neurons per thread) x block length
        # x blocks in one or two parts x segment source (table lookups, with staging depth 1 or 4,
        # or buckets) x, with plastic synapses, plastic rows pushed or pulled; shapes over the LDS
        # budget are skipped
        best_t, cur = float("inf"), None
        pulls = (0, 1) if len(bb.plastic_pos) else (0,)
        for ut in (bb.tile // k for k in (8, 4, 2, 1)):
            if not 64 <= ut <= 1024:
                continue
            for ks in sorted({bb.n_delay, max(1, bb.n_delay // 2)}):
                for parts in (1, 2):
                    for src, stage in ((0, 1), (0, 4), (1, 1)):
                        for pull in pulls:
                            trial = {"upd_threads": ut, "block_steps": ks, "tile_stage": stage, "tile_source": src}
                            if parts == 2:
                                trial["block_parts"] = 2
                            if pull:
                                trial["plastic_pull"] = 1
                            t = _time(bb, trial, steps, reps, snap)
                            if verbose:
                                print(f"autotune {trial}: {t * 1000:.1f} us", flush=True)
                            if t < best_t:
                                best_t, cur = t, trial
        space = []
    else:
        cur = {"upd_threads": 256, "push_threads": 256, "push_chunk": 256, "push_grid": n_cu * 8}
        space = [("upd_threads", UPD_THREADS), ("push_chunk", PUSH_CHUNK), ("push_threads", PUSH_THREADS),
                 ("push_grid", tuple(n_cu * k for k in PUSH_GRID_PER_CU))]
        best_t = _time(bb, cur, steps, reps, snap)
    for name, values in space:
        for v in values:
            if v == cur.get(name):
                continue
            trial = dict(cur, **{name: v})
            t = _time(bb, trial, steps, reps, snap)
            if verbose:
                print(f"autotune {trial}: {t * 1000:.1f} us", flush=True)
            if t < best_t:
                best_t, cur = t, trial
    _restore(bb, snap)
    bb.set_tune(**cur)
    cur = dict(cur)
    _mem[key] = cur
    if use_cache:
        try:
            os.makedirs(os.path.dirname(CACHE), exist_ok=True)
            data = json.load(open(CACHE)) if os.path.exists(CACHE) else {}
            data[key] = cur
            json.dump(data, open(CACHE, "w"), indent=1)
        except (OSError, ValueError):
            pass
    if verbose:
        print(f"autotune picked {cur}: {best_t * 1000:.1f} us per window", flush=True)
    return cur


def current(bb) -> dict:
    return {f: getattr(bb.tune, f) for f, _ in _lib.FzTune._fields_}
