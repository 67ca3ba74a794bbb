"""Data, training and judging for the learned DN-rate surrogate (``flyhip.surrogate``).

* ``VisualDrive``: per-fly visual input on the GPU: drifting Gaussian blobs over the input
  neurons' retinotopic positions (a different pattern and gain per fly), rates in Hz.
* ``generate``: runs the batched LIF simulator frame by frame (one game frame = 166 steps) and
  records the pooled drive (64 x 32 grid of mean input rates, BF16) and the 8 smoothed DN rates
  (``DNReadout``, tau = 80 ms, the reference's readout smoothing) of every fly and frame.
  ``replicas`` flies share one input sequence with independent spiking noise, so the
  simulation's own expected rate can be estimated by their mean.
* ``train_ensemble``: members trained by the package's own kernels on random windows of
  ``T`` consecutive frames (BPTT through the GRU).
* ``generate_judged`` / ``judge``: held-out sequences, each simulated by several replica flies;
  the replicas' mean is the simulation's expected rate (the judge).  Reports the surrogate's
  error with and without fallback, the error of a constant predictor (the method without the
  network) and of a single simulation trial (the simulation's own trial-to-trial noise), and
  frames/s with and without fallback.

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


class VisualDrive:
    def __init__(self, xy: np.ndarray, batch: int, seed: int, replicas: int = 1, n_blobs: int = 6,
                 device="cuda", peak_hz: float = 200.0):
        self.dev = torch.device(device)
        n_seq = batch // replicas
        r = np.random.default_rng(seed)
        f = lambda a: torch.from_numpy(np.repeat(a.astype(np.float32), replicas, 0)).to(self.dev)  # noqa: E731
        self.c0 = f(r.random((n_seq, n_blobs, 2)))
        self.vel = f((r.random((n_seq, n_blobs, 2)) - 0.5) * 0.03)
        self.width = f(0.06 + 0.14 * r.random((n_seq, n_blobs)))
        self.amp = f(0.3 + 0.7 * r.random((n_seq, n_blobs)))
        self.gain = f(0.5 + 1.5 * r.random(n_seq))
        self.xy = torch.from_numpy(xy.astype(np.float32)).to(self.dev)
        self.peak = float(peak_hz)

    def rates(self, frame: int) -> torch.Tensor:
        c = torch.remainder(self.c0 + self.vel * float(frame), 1.0)                 # [B, nb, 2]
        d = torch.abs(self.xy[None, :, None, :] - c[:, None, :, :])                # [B, n, nb, 2]
        d = torch.minimum(d, 1.0 - d)
        f = (self.amp[:, None, :] * torch.exp(-(d * d).sum(-1) / (2 * self.width[:, None, :] ** 2))).sum(-1)
        return self.peak * torch.clamp(f, 0.0, 1.0) * self.gain[:, None]


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


def generate(fly, frames: int, batch: int, seed: int, replicas: int = 1, engine: str = "tile", tile: int = 1024,
             skip: int = 10, drive_hz: float = 60.0, rate_norm: float = 200.0, log=None):
    """Simulate ``frames`` + ``skip`` frames of ``batch`` flies; returns (X bf16 [frames, B, 2048]
    normalised pooled drive, Y f32 [frames, B, 8] smoothed DN rates in Hz, sim frames/s)."""
    from . import BatchBrain, LIFParams, autotune
    from .batch import DNReadout

    p = LIFParams()
    bb = BatchBrain(fly.weights, p, batch=batch, seed=seed, engine=engine, tile=tile)
    j_in = bb.slots(fly.input_idx)
    j_dr = bb.slots([fly.drive_idx])
    dt_s = float(np.float32(p.dt / 1000.0))
    drive = VisualDrive(fly.input_xy, batch, seed, replicas)
    pool = GridPool(fly.input_xy)
    readout = DNReadout(fly, batch)
    bb.p_ext[:, j_dr] = drive_hz * dt_s
    bb.p_ext[:, j_in] = drive.rates(0) * dt_s
    autotune.tune(bb, reps=1)
    X = torch.empty((frames, batch, pool.n), dtype=torch.bfloat16, device=bb.device)
    Y = torch.empty((frames, batch, 8), dtype=torch.float32, device=bb.device)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(frames + skip):
        r = drive.rates(f)
        bb.p_ext[:, j_in] = r * dt_s
        counts = bb.run(FRAME_MS, copy=False)
        dn = readout.update(counts, FRAME_MS)
        if f >= skip:
            X[f - skip] = (pool(r) / rate_norm).to(torch.bfloat16)
            Y[f - skip] = dn
        if log and (f + 1) % 500 == 0:
            log(f"  generate: frame {f + 1}/{frames + skip}")
    torch.cuda.synchronize()
    fps = batch * (frames + skip) / (time.perf_counter() - t0)
    bb.close()
    return X, Y, fps


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
                if log:
                    log(f"  member {k} step {it}: loss {curve[-1][1]:.5f}")
        out.append(m)
        curves.append(curve)
    return out, curves


def generate_judged(fly, frames: int, sequences: int, replicas: int, batch: int = 256, seed: int = 77, log=None,
                    **kw):
    """Held-out data: ``sequences`` input sequences, each simulated by ``replicas`` flies with
    independent spiking noise (``batch // replicas`` sequences per batched run).  Returns the
    shared inputs X [F, S, 2048], the replicas' mean rate (the simulation's expected rate, the
    judge), one trial, the mean of the other trials, and sim frames/s."""
    per = batch // replicas
    if per < 1 or batch % replicas:
        raise ValueError("batch must be a multiple of replicas")
    xs, mean, trial, others, fps = [], [], [], [], []
    done = 0
    k = 0
    while done < sequences:
        X, Y, f = generate(fly, frames, batch, seed=seed + 1000 * k, replicas=replicas, **kw)
        Ys = Y.view(frames, per, replicas, 8)
        xs.append(X.view(frames, per, replicas, -1)[:, :, 0].clone())
        mean.append(Ys.mean(2))
        trial.append(Ys[:, :, 0].clone())
        others.append((Ys.sum(2) - Ys[:, :, 0]) / max(replicas - 1, 1))
        fps.append(f)
        del X, Y, Ys
        done += per
        k += 1
        if log:
            log(f"  held-out: {done}/{sequences} sequences")
    cat = lambda a: torch.cat(a, 1)[:, :sequences].contiguous()  # noqa: E731
    return cat(xs), cat(mean), cat(trial), cat(others), float(np.mean(fps))


def judge(ens, X, ref, trial, others, replicas: int, threshold: float = 0.15, const=None):
    """X [F, S, n_in] inputs, ref [F, S, 8] the simulation's expected rate (replica mean),
    trial one simulation trial, others the mean of the remaining replicas.  Errors are
    mean |a - ref| / mean |ref| over frames, sequences and the 8 groups."""
    F, S, n_in = X.shape
    ens.reset(S)
    pred = torch.empty_like(ref)
    spread = torch.empty((F, S), dtype=torch.float32, device=X.device)
    for f in range(F):
        pred[f], spread[f] = ens.predict(X[f].contiguous())
    fb = spread > threshold
    mixed = torch.where(fb[..., None], trial, pred)
    denom = ref.abs().mean()
    err = lambda a, b: float((a - b).abs().mean() / denom)  # noqa: E731
    noise = err(trial, others)
    res = {"frames": F * S, "sequences": S, "frames_per_sequence": F, "replicas": replicas,
           "mean_rate_hz": float(ref.mean()), "err_surrogate": err(pred, ref), "err_with_fallback": err(mixed, ref),
           "fallback_frac": float(fb.float().mean()), "err_single_sim_trial": err(trial, ref),
           "err_sim_trial_vs_other_trials": noise,
           # a trial's deviation from the mean of R-1 others is ~sqrt(R/(R-1)) x its deviation from
           # the expected rate; the judge (mean of R) deviates ~1/sqrt(R) of that
           "judge_noise_est": noise * math.sqrt((replicas - 1) / replicas) / math.sqrt(replicas),
           "err_per_group": [float((pred[..., g] - ref[..., g]).abs().mean() / ref[..., g].abs().mean().clamp_min(1e-6))
                             for g in range(ref.shape[-1])],
           "threshold": threshold}
    if const is not None:
        res["err_constant_predictor"] = err(const.expand_as(ref), ref)
    return res


def inference_fps(ens, batch: int = 4096, frames: int = 20):
    n_in = ens.members[0].n_in
    x = torch.randn((batch, n_in), device=ens.members[0].dev).to(torch.bfloat16)
    ens.reset(batch)
    for _ in range(3):
        ens.predict(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        ens.predict(x)
    torch.cuda.synchronize()
    return batch * frames / (time.perf_counter() - t0)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-frames", type=int, default=23400, help="frames per fly (x --batch flies)")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--test-frames", type=int, default=150, help="frames per held-out sequence")
    ap.add_argument("--test-sequences", type=int, default=128)
    ap.add_argument("--replicas", type=int, default=128, help="simulated flies per held-out sequence")
    ap.add_argument("--train-replicas", type=int, default=8,
                    help="flies per training sequence; targets are their mean rate (same simulated frames)")
    ap.add_argument("--members", type=int, default=3)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--threshold", type=float, default=0.15)
    ap.add_argument("--rate-scale", type=float, default=2.0, help="Hz per unit of the network's output")
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
    t0 = time.perf_counter()
    X, Y, sim_fps = generate(fly, args.train_frames, args.batch, seed=1, replicas=args.train_replicas, log=log)
    if args.train_replicas > 1:  # one input per sequence, the replicas' mean rate as its target
        R, F = args.train_replicas, X.shape[0]
        X = X.view(F, -1, R, X.shape[-1])[:, :, 0].contiguous()
        Y = Y.view(F, -1, R, 8).mean(2).contiguous()
        torch.cuda.empty_cache()
    log(f"training data: {tuple(X.shape)} in {time.perf_counter() - t0:.1f}s ({sim_fps:.0f} sim frames/s); "
        f"DN rate mean {float(Y.mean()):.2f} Hz, sd {float(Y.std()):.2f}")
    t0 = time.perf_counter()
    members, curves = train_ensemble(X, Y, members=args.members, steps=args.steps, lr=args.lr, log=log,
                                     rate_scale=args.rate_scale)
    train_s = time.perf_counter() - t0
    const = Y.mean((0, 1))
    del X, Y
    torch.cuda.empty_cache()
    Xt, ref, trial, others, _ = generate_judged(fly, args.test_frames, args.test_sequences, args.replicas,
                                                args.batch, log=log)
    ens = Ensemble(members)
    res = judge(ens, Xt, ref, trial, others, args.replicas, args.threshold, const)
    fps = inference_fps(ens)
    f = res["fallback_frac"]
    # every frame runs the ensemble; a fallback frame also costs one simulated frame
    eff = 1.0 / (1.0 / fps + f / sim_fps)
    res.update({"train_frames": args.train_frames * args.batch, "train_replicas": args.train_replicas,
                "train_steps": args.steps, "rate_scale": args.rate_scale,
                "train_s": round(train_s, 1), "members": args.members,
                "loss_curves": [c[::6] + [c[-1]] for c in curves], "surrogate_frames_per_s": round(fps),
                "sim_frames_per_s": round(sim_fps), "speedup_vs_sim": round(fps / sim_fps, 1),
                "frames_per_s_with_fallback": round(eff), "speedup_with_fallback": round(eff / sim_fps, 1)})
    print(json.dumps(res), flush=True)
    return res


if __name__ == "__main__":
    main()
