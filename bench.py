"""Benchmark: fly-sim-seconds per wall-second of the whole-brain LIF step on one MI355X.

One bench "step" is one game frame: 166 LIF steps (dt = 0.1 ms, 16.6 ms of brain time) for every
fly, with the frame's Poisson rates resampled from a smooth retinotopic field (the N=1 workload
is the single-fly 10,000-step run on the seeded FlyWire-shaped connectome: N = 138,639 neurons,
E = 15.1 M connections).  The frame is one HIP graph replay; the rates of every frame are staged
in HBM before the timed region and copied device-to-device into the input slots each frame.

Also reported:
* ``roofline``: the dominant kernel's algorithmic bytes per launch / its mean launch time (HIP
  events on the engine's stream), against the MI355X HBM peak;
* ``cpu_baseline``: the CPU oracle (a C restatement of the reference's time loop, one thread)
  on a bounded sample of the same workload on this host;
* ``batched``: the same metric for B = 256 flies (per-fly drive and visual gain) on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "can-the-fly-play-f-zero-_amd"))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
FRAME_MS = 1000.0 / 60.0988
METRIC = "fly-sim-seconds/wall-second at B=1 and B=256; surrogate frames/s at DN-rate err"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def frame_rates(fly, frames: int, gain=None, drive=None) -> np.ndarray:
    """[frames, B, n_inputs + 1] Hz (last column: the drive neuron)."""
    from flyhip.connectome import smooth_field

    vis = np.stack([smooth_field(fly.input_xy, f) for f in range(frames)])        # [F, n_in]
    gain = np.ones(1, np.float32) if gain is None else np.asarray(gain, np.float32)
    drive = np.full(len(gain), fly.drive_hz, np.float32) if drive is None else np.asarray(drive, np.float32)
    out = np.concatenate([vis[:, None, :] * gain[None, :, None],
                          np.broadcast_to(drive[None, :, None], (frames, len(gain), 1))], -1)
    return np.minimum(out, 1e4).astype(np.float32)


def run_workload(bb, fly, rates_dev, frames, warmup, torch):
    import flyhip  # noqa: F401

    idx = np.r_[fly.input_idx, fly.drive_idx]
    j = bb.slots(idx)
    dt_scale = np.float32(bb.p.dt / 1000.0)
    probs = rates_dev * float(dt_scale)                   # [F, B, n_in+1] per-step probability
    total = warmup + frames
    assert probs.shape[0] >= total

    def one(f):
        bb.p_ext[:, j] = probs[f]
        return bb.run(FRAME_MS, copy=False)

    for f in range(warmup):
        one(f)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(warmup, total):
        one(f)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60, help="timed game frames (166 LIF steps each)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=256, help="flies in the batched line (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample length")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the LIF path shards flies across GPUs with no exchange, and "
              "this build runs and reports the single-GPU share only (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from flyhip import BatchBrain, LIFParams, autotune
    from flyhip.connectome import flywire_like

    t = time.perf_counter()
    fly = flywire_like()
    log(f"connectome: n={fly.n} nnz={fly.weights.nnz} ({time.perf_counter() - t:.1f}s)")
    p = LIFParams()
    n, n_in = fly.n, len(fly.input_idx) + 1
    steps_per_frame = int(round(FRAME_MS / p.dt))
    total = args.warmup + args.steps

    # ---- B = 1 (headline) -------------------------------------------------
    bb = BatchBrain(fly.weights, p, batch=1, seed=0)
    rates = torch.from_numpy(frame_rates(fly, total)).cuda()
    j1 = bb.slots(np.r_[fly.input_idx, fly.drive_idx])
    bb.p_ext[:, j1] = rates[0] * float(np.float32(p.dt / 1000.0))
    shape1 = autotune.tune(bb, use_cache=False)
    log(f"B=1 launch shape: {shape1}")
    wall = run_workload(bb, fly, rates, args.steps, args.warmup, torch)
    sim_s = args.steps * steps_per_frame * p.dt / 1000.0
    value = sim_s / wall
    ms_per_frame = wall / args.steps * 1000.0
    log(f"B=1: {args.steps} frames in {wall:.3f}s -> {value:.2f} fly-sim-s/s ({ms_per_frame:.3f} ms/frame)")

    # per-kernel timing (HIP events on the engine stream) over a few eager frames
    deg = np.diff(fly.weights.indptr).astype(np.float64)
    tim = {"update_ms": 0.0, "update_launches": 0, "push_ms": 0.0, "push_launches": 0}
    push_bytes = 0.0
    n_spikes = 0.0
    kframes = 6
    idx = np.r_[fly.input_idx, fly.drive_idx]
    j = bb.slots(idx)
    for f in range(kframes):
        bb.p_ext[:, j] = rates[f % total] * float(np.float32(p.dt / 1000.0))
        r = bb.run_timed(steps_per_frame)
        c = bb.counts[0].cpu().numpy().astype(np.float64)
        push_bytes += float((c * (8.0 + 10.0 * deg)).sum())
        n_spikes += float(c.sum())
        for k in tim:
            tim[k] += r[k]
    upd_steps = kframes * steps_per_frame
    upd_bytes_total = upd_steps * (26.0 * n + 4.0 * n_in)
    upd_avg_ms = tim["update_ms"] / max(tim["update_launches"], 1)
    push_avg_ms = tim["push_ms"] / max(tim["push_launches"], 1)
    if tim["update_ms"] >= tim["push_ms"]:
        dom, bytes_launch, avg_ms = "lif_update", upd_bytes_total / tim["update_launches"], upd_avg_ms
    else:
        dom, bytes_launch, avg_ms = "lif_push", push_bytes / tim["push_launches"], push_avg_ms
    achieved = bytes_launch / (avg_ms / 1000.0) / 1e9
    traffic = None
    tpath = os.path.join(REPO, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(dom, {}).get("bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "avg_launch_us": round(avg_ms * 1000.0, 3), "bytes_per_launch": round(bytes_launch),
                "update_avg_us": round(upd_avg_ms * 1000.0, 3), "push_avg_us": round(push_avg_ms * 1000.0, 3),
                "update_launches": tim["update_launches"], "push_launches": tim["push_launches"],
                "spikes_per_step": round(n_spikes / max(upd_steps, 1), 1), "launch_shape": shape1}
    log(f"kernels: update {upd_avg_ms * 1000:.1f} us/launch, push {push_avg_ms * 1000:.1f} us/launch; "
        f"dominant {dom}: {achieved:.1f} GB/s")

    # ---- B = 256 (batched line) --------------------------------------------
    batched = None
    if args.batch > 0:
        B = args.batch
        rng = np.random.default_rng(5)
        drive = rng.uniform(20.0, 120.0, B).astype(np.float32)
        gain = rng.uniform(0.5, 2.0, B).astype(np.float32)
        nb = args.batch_steps + 2
        gbb = BatchBrain(fly.weights, p, batch=B, seed=1)
        rb = torch.from_numpy(frame_rates(fly, nb, gain, drive)).cuda()
        jb = gbb.slots(np.r_[fly.input_idx, fly.drive_idx])
        gbb.p_ext[:, jb] = rb[0] * float(np.float32(p.dt / 1000.0))
        shapeb = autotune.tune(gbb, use_cache=False, reps=1)
        log(f"B={B} launch shape: {shapeb}")
        wb = run_workload(gbb, fly, rb, args.batch_steps, 2, torch)
        vb = B * args.batch_steps * steps_per_frame * p.dt / 1000.0 / wb
        batched = {"B": B, "frames": args.batch_steps, "value": round(vb, 3), "unit": "fly-sim-seconds/wall-second",
                   "launch_shape": shapeb,
                   "ms_per_frame": round(wb / args.batch_steps * 1000.0, 3),
                   "vs_b1": round(vb / value, 2)}
        log(f"B={B}: {vb:.2f} fly-sim-s/s aggregate ({batched['ms_per_frame']:.2f} ms/frame)")
        del gbb, rb
        torch.cuda.empty_cache()

    # ---- CPU baseline (oracle, one thread) ----------------------------------
    cpu = None
    if not args.no_cpu:
        from oracle.oracle import OracleBrain

        ob = OracleBrain(fly.weights, p, batch=1, seed=0)
        rr = frame_rates(fly, 64)
        ob.set_input(idx, rr[0, 0])
        t0 = time.perf_counter()
        ob.run(FRAME_MS)
        one = time.perf_counter() - t0
        nf = max(1, min(60, int(args.cpu_seconds / max(one, 1e-3))))
        t0 = time.perf_counter()
        for f in range(nf):
            ob.set_input(idx, rr[(f + 1) % 64, 0])
            ob.run(FRAME_MS)
        cw = time.perf_counter() - t0
        cv = nf * steps_per_frame * p.dt / 1000.0 / cw
        cpu = {"value": round(cv, 5), "unit": "fly-sim-seconds/wall-second", "cores": 1, "kind": "port",
               "sample": f"{nf} game frames ({nf * steps_per_frame} LIF steps) of the same B=1 workload, "
                         f"C oracle -O2 single thread, {platform.processor() or platform.machine()}, "
                         f"{os.cpu_count()} logical CPUs on host",
               "gpu_over_cpu": round(value / cv, 1)}
        log(f"CPU oracle: {cv:.4f} fly-sim-s/s ({cw / nf * 1000:.1f} ms/frame)")

    line = {"metric": METRIC, "value": round(value, 4), "unit": "fly-sim-seconds/wall-second", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_frame, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded FlyWire-shaped connectome, smooth-field Poisson input)",
            "config": {"workload": "plain LIF, 1 fly, dt=0.1 ms, 166-step game frames (hipGraph), "
                                   "rates resampled per frame", "neurons": n, "connections": int(fly.weights.nnz),
                       "batch": 1, "steps_per_frame": steps_per_frame, "parallelism": "single GPU"},
            "roofline": roofline, "cpu_baseline": cpu, "batched": batched}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
