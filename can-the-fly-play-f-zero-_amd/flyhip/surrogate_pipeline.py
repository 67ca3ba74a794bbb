"""Data, training, fallback and judging for the learned DN-rate surrogate (``flyhip.surrogate``).

The workload is the batched simulator's own (B flies sharing the connectome): per-sequence visual
drive with a visual gain U(0.5, 2.0) and a drive-neuron rate U(20, 120) Hz, and the fleet's
steady 7.8 mV bias on the motor descending neurons (flyzero/fleet.py:61-63), which puts them in
the firing regime the fleet runs them in.

* ``VisualDrive``: per-sequence input on the GPU: drifting Gaussian blobs over the input
  neurons' retinotopic positions, gain and drive rate; rates for all sequences or any subset.
* ``generate``: runs the batched simulator frame by frame (one game frame = 166 steps) and records
  the pooled visual drive (64 x 32 grid of mean input rates, BF16) and the 8 smoothed DN rates
  (``DNReadout``, tau = 80 ms, the reference's readout smoothing) of every fly and frame.
  ``replicas`` flies share one sequence with independent spiking noise, so the simulation's
  expected rate is estimated by their mean.
* ``sim_frames_per_s``: the full simulation's throughput on the same setup (B = 256, graph
  replays of the brain alone, inputs already on the device): the denominator of the speedup.
* ``train_ensemble``: members trained by the package's own kernels on random windows of ``T``
  consecutive frames (BPTT through the GRU).
* ``FallbackSim`` and ``surrogate_loop``: the surrogate run frame by frame for S sequences; a
  frame whose ensemble disagreement exceeds the threshold is simulated instead.  The simulation
  has no state for a sequence the surrogate has been running, so the fallback re-warms one: a
  brain at rest runs the sequence's previous ``warm`` frames of drive and then the frame itself
  (the network's time constants are 5-20 ms, a frame is 16.6 ms), and the frame's simulated DN
  counts enter the readout's exponential average in place of the network's prediction.  The
  loop is timed with the fallback simulations in it.
* ``judge``: held-out sequences, each simulated by several replica flies; the replicas' mean is
  the simulation's expected rate.  Errors of the surrogate alone, of the loop with its real
  fallback, of a constant predictor (the method without the network) and of one simulated
  trial (the simulation's trial-to-trial noise).

``python -m flyhip.surrogate_pipeline`` runs all of it and prints one JSON report.
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

FRAME_MS = 1000.0 / 60.0988
BIAS_TYPES = ("DNa02", "DNg02*", "DNa01", "DNp09")   # the fleet's biased motor DNs


class VisualDrive:
    """S input sequences: visual rates [S, n_in] per frame and a drive-neuron rate per sequence."""

    def __init__(self, xy: np.ndarray, sequences: int, seed: int, n_blobs: int = 6, device="cuda",
                 peak_hz: float = 200.0, gain=(0.5, 2.0), drive_hz=(20.0, 120.0)):
        self.dev = torch.device(device)
        self.S = sequences
        r = np.random.default_rng(seed)
        t = lambda a: torch.from_numpy(np.asarray(a, np.float32)).to(self.dev)  # noqa: E731
        self.c0 = t(r.random((sequences, n_blobs, 2)))
        self.vel = t((r.random((sequences, n_blobs, 2)) - 0.5) * 0.03)
        self.width = t(0.06 + 0.14 * r.random((sequences, n_blobs)))
        self.amp = t(0.3 + 0.7 * r.random((sequences, n_blobs)))
        self.gain = t(r.uniform(gain[0], gain[1], sequences))
        self.drive_hz = t(r.uniform(drive_hz[0], drive_hz[1], sequences))
        self.xy = torch.from_numpy(xy.astype(np.float32)).to(self.dev)
        self.peak = float(peak_hz)

    def rates(self, frame: int, seqs: torch.Tensor | None = None) -> torch.Tensor:
        """Visual rates (Hz) of the input neurons at ``frame``: [S, n_in], or [len(seqs), n_in]."""
        pick = (lambda a: a) if seqs is None else (lambda a: a[seqs])  # noqa: E731
        c = torch.remainder(pick(self.c0) + pick(self.vel) * float(frame), 1.0)        # [s, nb, 2]
        d = torch.abs(self.xy[None, :, None, :] - c[:, None, :, :])                    # [s, n, nb, 2]
        d = torch.minimum(d, 1.0 - d)
        w = pick(self.width)
        f = (pick(self.amp)[:, None, :] * torch.exp(-(d * d).sum(-1) / (2 * w[:, None, :] ** 2))).sum(-1)
        return self.peak * torch.clamp(f, 0.0, 1.0) * pick(self.gain)[:, None]


class GridPool:
    """Mean rate of the input neurons in each cell of an nx x ny grid over [0, 1)^2."""

    def __init__(self, xy: np.ndarray, nx: int = 64, ny: int = 32, device="cuda"):
        cx = np.minimum((xy[:, 0] * nx).astype(np.int64), nx - 1)
        cy = np.minimum((xy[:, 1] * ny).astype(np.int64), ny - 1)
        cell = cy * nx + cx
        self.n = nx * ny
        self.cell = torch.from_numpy(cell).to(device)
        cnt = np.bincount(cell, minlength=self.n).astype(np.float32)
        self.inv = torch.from_numpy(np.where(cnt > 0, 1.0 / np.maximum(cnt, 1), 0.0).astype(np.float32)).to(device)

    def __call__(self, rates: torch.Tensor) -> torch.Tensor:
        out = torch.zeros((rates.shape[0], self.n), dtype=torch.float32, device=rates.device)
        out.index_add_(1, self.cell, rates)
        return out * self.inv


class _Driven:
    """A BatchBrain wired to a ``VisualDrive``: visual slots, the drive neuron and the DN bias."""

    def __init__(self, fly, batch: int, seed: int, dn_bias_mv: float, engine="tile", tile=1024, tune=True):
        from . import BatchBrain, LIFParams, autotune

        self.p = LIFParams()
        self.bb = bb = BatchBrain(fly.weights, self.p, batch=batch, seed=seed, engine=engine, tile=tile)
        self.j_in = bb.slots(fly.input_idx)
        self.j_dr = bb.slots([fly.drive_idx])
        if dn_bias_mv:
            bias_idx = np.concatenate([fly.find(t) for t in BIAS_TYPES])
            bb.set_bias(bias_idx, dn_bias_mv)
        self.dt_s = float(np.float32(self.p.dt / 1000.0))
        self.tuned = autotune.tune(bb, reps=1) if tune else None

    def set_rates(self, visual: torch.Tensor, drive: torch.Tensor, rows=None):
        if rows is None:
            self.bb.p_ext[:, self.j_in] = visual * self.dt_s
            self.bb.p_ext[:, self.j_dr] = drive[:, None] * self.dt_s
        else:
            self.bb.p_ext[rows[:, None], self.j_in[None, :]] = visual * self.dt_s
            self.bb.p_ext[rows, self.j_dr[0]] = drive * self.dt_s


def generate(fly, frames: int, batch: int, seed: int, replicas: int = 1, skip: int = 10, dn_bias_mv: float = 7.8,
             rate_norm: float = 200.0, log=None, sim: "_Driven | None" = None):
    """Simulate ``frames`` + ``skip`` frames of ``batch`` flies (``batch // replicas`` sequences);
    returns (X bf16 [frames, S, 2048] normalised pooled visual drive of each sequence, Y f32
    [frames, batch, 8] smoothed DN rates in Hz, the sequences' drive).  ``sim``: a brain of
    ``batch`` flies to reuse (reset to rest first)."""
    from .batch import DNReadout

    if batch % replicas:
        raise ValueError("batch must be a multiple of replicas")
    S = batch // replicas
    own = sim is None
    if own:
        sim = _Driven(fly, batch, seed, dn_bias_mv)
    else:
        sim.bb.reset()
    drive = VisualDrive(fly.input_xy, S, seed)
    seq_of = torch.arange(batch, device=sim.bb.device) // replicas
    pool = GridPool(fly.input_xy)
    readout = DNReadout(fly, batch)
    X = torch.empty((frames, S, pool.n), dtype=torch.bfloat16, device=sim.bb.device)
    Y = torch.empty((frames, batch, 8), dtype=torch.float32, device=sim.bb.device)
    for f in range(frames + skip):
        r = drive.rates(f)
        sim.set_rates(r[seq_of], drive.drive_hz[seq_of])
        dn = readout.update(sim.bb.run(FRAME_MS, copy=False), FRAME_MS)
        if f >= skip:
            X[f - skip] = (pool(r) / rate_norm).to(torch.bfloat16)
            Y[f - skip] = dn
        if log and (f + 1) % 1000 == 0:
            log(f"  generate: frame {f + 1}/{frames + skip}")
    if own:
        sim.bb.close()
    return X, Y, drive


def sim_frames_per_s(fly, batch: int = 256, frames: int = 20, dn_bias_mv: float = 7.8, seed: int = 5):
    """Fly-frames per second of the full simulation on the same setup: ``batch`` flies, their
    rates for every frame staged on the device before the timed region, one graph replay per
    frame (the brain alone: what a fallback or a surrogate replaces)."""
    sim = _Driven(fly, batch, seed, dn_bias_mv)
    drive = VisualDrive(fly.input_xy, batch, seed)
    staged = [drive.rates(f) * sim.dt_s for f in range(frames + 2)]
    sim.set_rates(staged[0] / sim.dt_s, drive.drive_hz)
    for f in range(2):
        sim.bb.p_ext[:, sim.j_in] = staged[f]
        sim.bb.run(FRAME_MS, copy=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(2, frames + 2):
        sim.bb.p_ext[:, sim.j_in] = staged[f]
        sim.bb.run(FRAME_MS, copy=False)
    torch.cuda.synchronize()
    fps = batch * frames / (time.perf_counter() - t0)
    sim.bb.close()
    return fps, sim.tuned


def train_ensemble(X, Y, members=3, T=16, seqs=256, steps=600, lr=1e-3, rate_scale=100.0, seed=0, log=None,
                   hidden=1024, gru=256):
    from .surrogate import RateSurrogate

    F, B, n_in = X.shape
    rng = np.random.default_rng(seed)
    out, curves = [], []
    for k in range(members):
        m = RateSurrogate(n_in, hidden, gru, 8, seed=seed + 1000 * k, device=X.device, rate_scale=rate_scale)
        h0 = torch.zeros((seqs, gru), dtype=torch.float32, device=X.device)
        curve = []
        for it in range(steps):
            f0 = torch.from_numpy(rng.integers(0, F - T + 1, seqs)).to(X.device)
            b = torch.from_numpy(rng.integers(0, B, seqs)).to(X.device)
            tt = torch.arange(T, device=X.device)[:, None]
            xb = X[f0[None, :] + tt, b[None, :]]                       # [T, seqs, n_in]
            yb = Y[f0[None, :] + tt, b[None, :]] / rate_scale
            m.forward(xb, h0, train=True)
            loss = m.loss_backward(xb, yb)
            lr_t = lr * 0.5 * (1 + math.cos(math.pi * it / steps))   # cosine decay
            m.adam(lr=lr_t)
            if it % 50 == 0 or it == steps - 1:
                curve.append((it, float(loss.item())))
                if log and (it % 1000 == 0 or it == steps - 1):
                    log(f"  member {k} step {it}: loss {curve[-1][1]:.5f}")
        out.append(m)
        curves.append(curve)
    return out, curves


class FallbackSim:
    """Simulates flagged (sequence, frame) pairs for ``surrogate_loop``: a brain at rest runs the
    sequence's last ``warm`` frames of drive and then the flagged frame, and returns that frame's
    instantaneous DN group rates (mean count of the group's neurons x 1000 / window).  Brains of
    several sizes (``sizes`` flies, smallest first) so that a few flagged sequences do not pay
    for a full batch; the smallest brain that holds them runs, a larger set runs in chunks of the
    largest.  Every fallback draws a fresh spiking trial."""

    def __init__(self, fly, drive: VisualDrive | None, warm: int = 3, sizes=(4, 32, 256), dn_bias_mv: float = 7.8,
                 seed: int = 900):
        from .batch import DNReadout

        self.drive, self.warm = drive, warm
        self.sims = [_Driven(fly, n, seed + k, dn_bias_mv) for k, n in enumerate(sorted(sizes))]
        # alpha = 1: the readout kernel returns the frame's instantaneous group rates
        self.inst = [DNReadout(fly, s.bb.B, tau_ms=1e-9) for s in self.sims]
        self.frames_simulated = 0

    def run(self, seqs: torch.Tensor, frame: int) -> torch.Tensor:
        outs = []
        for chunk in seqs.split(self.sims[-1].bb.B):
            n = chunk.numel()
            k = next(i for i, s in enumerate(self.sims) if s.bb.B >= n)
            sim, inst = self.sims[k], self.inst[k]
            bb = sim.bb
            bb.reset()
            bb.alive.zero_()
            bb.alive[:n] = 1
            rows = torch.arange(n, device=bb.device)
            counts = None
            for f in range(max(0, frame - self.warm), frame + 1):
                sim.set_rates(self.drive.rates(f, chunk), self.drive.drive_hz[chunk], rows)
                counts = bb.run(FRAME_MS, copy=False)
                self.frames_simulated += n
            inst.rates.zero_()
            outs.append(inst.update(counts, FRAME_MS)[:n].clone())
        return torch.cat(outs)


def surrogate_loop(ens, X: torch.Tensor, threshold: float, fallback: FallbackSim | None, tau_ms: float = 80.0):
    """Runs the ensemble over X [F, S, n_in] frame by frame; frames whose disagreement exceeds
    ``threshold`` are simulated by ``fallback`` (their simulated counts enter the readout's
    exponential average: r = r_prev + a (inst - r_prev)).  Returns (rates [F, S, 8] with fallback,
    the network's own rates [F, S, 8], fallback mask [F, S]); the last call's disagreements [F, S]
    stay in ``surrogate_loop.spread``."""
    F, S, _ = X.shape
    a = float(1.0 - math.exp(-FRAME_MS / tau_ms))
    ens.reset(S)
    out = torch.empty((F, S, 8), dtype=torch.float32, device=X.device)
    net = torch.empty_like(out)
    flags = torch.zeros((F, S), dtype=torch.bool, device=X.device)
    spreads = torch.empty((F, S), dtype=torch.float32, device=X.device)
    prev = None
    for f in range(F):
        pred, spread = ens.predict(X[f])
        net[f] = pred
        spreads[f] = spread
        y = pred
        if fallback is not None:
            flag = spread > threshold
            idx = torch.nonzero(flag).squeeze(1)           # the host needs the count to launch
            if idx.numel():
                base = prev[idx] if prev is not None else pred[idx]
                y = pred.clone()
                y[idx] = base + a * (fallback.run(idx, f) - base)
                flags[f] = flag
        out[f] = y
        prev = out[f]
    surrogate_loop.spread = spreads
    return out, net, flags


def generate_judged(fly, frames: int, sequences: int, replicas: int, batch: int = 256, seed: int = 77, log=None,
                    **kw):
    """Held-out data: ``sequences`` sequences, each simulated by ``replicas`` flies.  Returns X
    [F, S, 2048], the replicas' mean rate (the judge), one trial, the mean of the other trials,
    and each run's ``VisualDrive`` with the range of sequences it holds."""
    per = batch // replicas
    if per < 1 or batch % replicas:
        raise ValueError("batch must be a multiple of replicas")
    sim = _Driven(fly, batch, seed, kw.get("dn_bias_mv", 7.8))
    xs, mean, trial, others, drives = [], [], [], [], []
    done, k = 0, 0
    while done < sequences:
        X, Y, drive = generate(fly, frames, batch, seed=seed + 1000 * k, replicas=replicas, sim=sim, **kw)
        Ys = Y.view(frames, per, replicas, 8)
        xs.append(X)
        mean.append(Ys.mean(2))
        trial.append(Ys[:, :, 0].clone())
        others.append((Ys.sum(2) - Ys[:, :, 0]) / max(replicas - 1, 1))
        drives.append((done, drive))
        del Y, Ys
        done += per
        k += 1
        if log:
            log(f"  held-out: {done}/{sequences} sequences")
    sim.bb.close()
    cat = lambda a: torch.cat(a, 1)[:, :sequences].contiguous()  # noqa: E731
    return cat(xs), cat(mean), cat(trial), cat(others), drives


def judge(pred, ref, trial, others, replicas: int, loop=None, flags=None, const=None):
    """Errors are mean |a - ref| / mean |ref| over frames, sequences and the 8 groups; ref is the
    simulation's expected rate (replica mean)."""
    denom = ref.abs().mean()
    err = lambda a, b: float((a - b).abs().mean() / denom)  # noqa: E731
    noise = err(trial, others)
    F, S, _ = ref.shape
    res = {"frames": F * S, "sequences": S, "frames_per_sequence": F, "replicas": replicas,
           "mean_rate_hz": float(ref.mean()), "rate_hz_per_group": [float(ref[..., g].mean()) for g in range(8)],
           "err_surrogate": err(pred, ref), "err_single_sim_trial": err(trial, ref),
           "err_sim_trial_vs_other_trials": noise,
           # a trial's deviation from the mean of R-1 others is ~sqrt(R/(R-1)) x its deviation from
           # the expected rate; the judge (mean of R) deviates ~1/sqrt(R) of that
           "judge_noise_est": noise * math.sqrt((replicas - 1) / replicas) / math.sqrt(replicas),
           "err_per_group": [float((pred[..., g] - ref[..., g]).abs().mean() / ref[..., g].abs().mean().clamp_min(1e-6))
                             for g in range(ref.shape[-1])]}
    if loop is not None:
        res["err_with_fallback"] = err(loop, ref)
        res["fallback_frac"] = float(flags.float().mean())
    if const is not None:
        res["err_constant_predictor"] = err(const.expand_as(ref), ref)
    return res


def inference_fps(ens, batch: int = 4096, frames: int = 20):
    """Ensemble frames/s at ``batch`` sequences, timed with HIP events on the current stream."""
    n_in = ens.members[0].n_in
    x = torch.randn((batch, n_in), device=ens.members[0].dev).to(torch.bfloat16)
    ens.reset(batch)
    for _ in range(3):
        ens.predict(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(frames):
        ens.predict(x)
    e1.record()
    e1.synchronize()
    return batch * frames / (e0.elapsed_time(e1) / 1000.0)


def timed_loop(fly, ens, threshold: float, warm: int, sequences: int = 4096, frames: int = 30, seed: int = 5151,
               rate_norm: float = 200.0, fallback: FallbackSim | None = None):
    """Frames/s of the surrogate with its real fallback: ``sequences`` fresh sequences, their pooled
    inputs staged on the device first, then ``frames`` frames of ``surrogate_loop`` timed end to
    end (ensemble, disagreement, fallback simulations and the readout update)."""
    drive = VisualDrive(fly.input_xy, sequences, seed)
    pool = GridPool(fly.input_xy)
    X = torch.stack([(pool(drive.rates(f)) / rate_norm).to(torch.bfloat16) for f in range(frames)])
    fb = fallback if fallback is not None else FallbackSim(fly, drive, warm=warm)
    fb.drive, fb.warm = drive, warm
    surrogate_loop(ens, X[:2], threshold, fb)        # warm-up: graph capture of both fallback brains
    fb.frames_simulated = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, flags = surrogate_loop(ens, X, threshold, fb)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"sequences": sequences, "frames": frames, "frames_per_s": sequences * frames / wall,
            "fallback_frac": float(flags.float().mean()), "fallback_sim_fly_frames": fb.frames_simulated}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-frames", type=int, default=8000, help="frames per fly (x --batch flies)")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--train-replicas", type=int, default=4,
                    help="flies per training sequence; targets are their mean rate")
    ap.add_argument("--test-frames", type=int, default=200, help="frames per held-out sequence")
    ap.add_argument("--test-sequences", type=int, default=250)
    ap.add_argument("--replicas", type=int, default=32, help="simulated flies per held-out sequence")
    ap.add_argument("--members", type=int, default=3)
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--threshold", type=float, default=0.15)
    ap.add_argument("--warm", type=int, default=3, help="frames of drive a fallback re-warms its brain on")
    ap.add_argument("--sweep", default="0.03,0.02", help="further disagreement thresholds to run the loop at")
    ap.add_argument("--rate-scale", type=float, default=20.0, help="Hz per unit of the network's output")
    ap.add_argument("--dn-bias", type=float, default=7.8, help="mV on the motor DNs (the fleet's)")
    ap.add_argument("--n", type=int, default=0, help="neurons (0: the full 138,639)")
    ap.add_argument("--inference-only", action="store_true",
                    help="time ensemble inference at B = 4,096 with freshly initialised members (for profiling)")
    args = ap.parse_args(argv)
    if args.inference_only:
        from .surrogate import Ensemble, RateSurrogate

        ens = Ensemble([RateSurrogate(seed=k) for k in range(args.members)])
        fps = inference_fps(ens)
        print(json.dumps({"members": args.members, "batch": 4096, "surrogate_frames_per_s": round(fps)}), flush=True)
        return {"surrogate_frames_per_s": fps}

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    from .connectome import flywire_like
    from .surrogate import Ensemble

    fly = flywire_like() if not args.n else flywire_like(n=args.n, n_edges=args.n * 109, seed=11)
    sim_fps, sim_shape = sim_frames_per_s(fly, args.batch, dn_bias_mv=args.dn_bias)
    log(f"full simulation: {sim_fps:.0f} fly-frames/s at B = {args.batch} ({sim_shape})")
    t0 = time.perf_counter()
    X, Y, _ = generate(fly, args.train_frames, args.batch, seed=1, replicas=args.train_replicas,
                       dn_bias_mv=args.dn_bias, log=log)
    F, R = X.shape[0], args.train_replicas
    Y = Y.view(F, -1, R, 8).mean(2).contiguous()          # one input per sequence, the replicas' mean
    torch.cuda.empty_cache()
    log(f"training data: {tuple(X.shape)} from {F * args.batch} simulated frames in {time.perf_counter() - t0:.1f}s; "
        f"DN rate mean {float(Y.mean()):.2f} Hz, sd {float(Y.std()):.2f}")
    t0 = time.perf_counter()
    members, curves = train_ensemble(X, Y, members=args.members, steps=args.steps, lr=args.lr, log=log,
                                     rate_scale=args.rate_scale)
    train_s = time.perf_counter() - t0
    const = Y.mean((0, 1))
    del X, Y
    torch.cuda.empty_cache()
    ens = Ensemble(members)

    # held-out sequences, each simulated by R replica flies; the surrogate runs on them with its
    # real fallback (every flagged frame re-simulated)
    Xt, ref, trial, others, drives = generate_judged(fly, args.test_frames, args.test_sequences, args.replicas,
                                                     args.batch, log=log, dn_bias_mv=args.dn_bias)
    fb = FallbackSim(fly, None, warm=args.warm, dn_bias_mv=args.dn_bias)

    def judged_loop(threshold):
        loops, nets, flags, spreads = [], [], [], []
        for start, drive in drives:
            S = min(drive.S, args.test_sequences - start)
            fb.drive = drive
            lo, ne, fl = surrogate_loop(ens, Xt[:, start:start + S].contiguous(), threshold, fb)
            loops.append(lo)
            nets.append(ne)
            flags.append(fl)
            spreads.append(surrogate_loop.spread)
        fb.drive = None
        return torch.cat(loops, 1), torch.cat(nets, 1), torch.cat(flags, 1), torch.cat(spreads, 1)

    loop, net, flag, spread = judged_loop(args.threshold)
    res = judge(net, ref, trial, others, args.replicas, loop=loop, flags=flag, const=const)
    q = torch.quantile(spread.flatten()[:1_000_000], torch.tensor([0.5, 0.9, 0.99, 0.999, 1.0], device=spread.device))
    res["disagreement_quantiles"] = {k: round(float(v), 5) for k, v in zip(("p50", "p90", "p99", "p999", "max"), q)}
    fps = inference_fps(ens)
    tl = timed_loop(fly, ens, args.threshold, args.warm, fallback=fb)
    # the same loop with lower thresholds, so that the fallback runs on a measurable share of frames
    sweep = []
    for t in [float(x) for x in args.sweep.split(",") if x]:
        lo, _, fl, _ = judged_loop(t)
        tt = timed_loop(fly, ens, t, args.warm, fallback=fb)
        sweep.append({"threshold": t, "err_with_fallback": judge(net, ref, trial, others, args.replicas, loop=lo,
                                                                  flags=fl)["err_with_fallback"],
                      "fallback_frac": float(fl.float().mean()), "timed_fallback_frac": round(tt["fallback_frac"], 5),
                      "frames_per_s_with_fallback": round(tt["frames_per_s"]),
                      "speedup_with_fallback": round(tt["frames_per_s"] / sim_fps, 1),
                      "fallback_sim_fly_frames": tt["fallback_sim_fly_frames"]})
        log(f"  threshold {t}: {sweep[-1]}")
    res["threshold_sweep"] = sweep
    res.update({"train_sim_frames": args.train_frames * args.batch, "train_replicas": args.train_replicas,
                "train_steps": args.steps, "rate_scale": args.rate_scale, "dn_bias_mv": args.dn_bias,
                "threshold": args.threshold, "warm_frames": args.warm,
                "train_s": round(train_s, 1), "members": args.members,
                "loss_curves": [c[::24] + [c[-1]] for c in curves],
                "surrogate_frames_per_s": round(fps), "sim_frames_per_s": round(sim_fps),
                "speedup_vs_sim": round(fps / sim_fps, 1),
                "timed_loop": {k: (round(v, 5) if isinstance(v, float) else v) for k, v in tl.items()},
                "frames_per_s_with_fallback": round(tl["frames_per_s"]),
                "speedup_with_fallback": round(tl["frames_per_s"] / sim_fps, 1)})
    print(json.dumps(res), flush=True)
    return res


if __name__ == "__main__":
    main()
