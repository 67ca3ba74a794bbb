"""A fleet: B flies on the GPU, each at the controls of its own game (flyzero/fleet.py:22-127).

Glue between ``BatchBrain`` (brains), a game pool (games), ``BatchMotionEye`` (eyes),
``BatchMotor`` (buttons) and the batched learning rules.  The fly's interface is the reference
fleet's: the steady bias on the motor neurons, dt = 0.25 ms, one game frame per 16.64 ms of brain
time, the eye's rates as the brain's Poisson input.

Where the reference keeps one ``MotionEye`` per emulator subprocess and ships its rates back to
the main process (flyzero/pool.py:41-44), here the pool hands over frames: every fly's frame
goes through one eye kernel that writes the brain's input probabilities on the device.  A pool
is any object with

    load(states, flips=None, which=None) -> (frames, infos)
    step(buttons, which=None) -> (frames, infos)

where ``frames`` is uint8 (B, h, w, 3), host (numpy) or device (torch), and ``infos`` one dict per
game with ``segment``, ``lap``, ``energy``, ``done`` and optionally ``finished`` and ``rank``
(flyzero/pool.py:133-203, the F-Zero telemetry).  Host frames are staged through pinned memory
into one device buffer.  ``SyntheticRace`` is such a pool: a deterministic road game with the
same telemetry, for tests and measurement where no emulator exists.  ``GamePool`` runs one game
per subprocess (an emulator allows one per process) and gathers their frames in shared memory.
"""

from __future__ import annotations

import multiprocessing as mp
import time
from multiprocessing import shared_memory

import numpy as np
import torch

BIAS_TYPES = ("DNa02", "DNg02*", "DNa01", "DNp09")
WINDOW_MS = 1000.0 / 60.0988
FRAME_SHAPE = (224, 256)
# the fleet's eye: motion, medulla and colour channels (flyzero/train.py:21, gain = 10 ** 2.7)
VISION = dict(log_gain=2.7, y_min=0.2, transient_gain=4.0, sustained_gain=4.0, chromatic_gain=1.0)


class Progress:
    """Unwrapped track progress from the per-frame segment number (flyzero/tune.py:57-71): the
    segment step is taken modulo the lap, in [-segments / 2, segments / 2)."""

    def __init__(self, segments: int = 59):
        self.segments = segments
        self.last = None
        self.total = 0

    def update(self, seg: int, frame: int) -> int:
        if self.last is not None:
            n = self.segments
            self.total += (seg - self.last + n // 2) % n - n // 2
        self.last = seg
        return self.total


def summarize(results: list[dict]) -> dict:
    """flyzero/fleet.py:129-136."""
    prog = [r["progress"] for r in results]
    return {"mean_progress": round(float(np.mean(prog)), 1), "best": int(max(prog)), "worst": int(min(prog)),
            "laps_best": max(r["lap"] for r in results), "finished": sum(r["finished"] for r in results),
            "laps5": sum(r.get("laps5", False) for r in results), "n": len(results),
            "mean_frames": int(np.mean([r["frames"] for r in results]))}


def buttons_to_mask(pressed: dict) -> np.ndarray:
    """flyzero/record.py:87-88."""
    from .batch import BUTTONS

    return np.array([pressed.get(b, False) for b in BUTTONS], np.uint8)


def vision_params(vision: dict | None = None):
    """``MotionParams`` from a ``VISION``-style dict (``log_gain`` -> ``gain``)."""
    from .eye import MotionParams

    v = dict(VISION if vision is None else vision)
    if "log_gain" in v:
        v["gain"] = 10 ** v.pop("log_gain")
    return MotionParams(**v)


class _Part:
    """The flies [b0, b0 + nb) of a fleet with their own brain, eye and readout.  A fleet of
    several parts lets the host serve one part (buttons, games) while the device runs another.
    The brain numbers its flies from b0 in the Poisson draw (``fly_base``), so every fly draws
    the same spiking noise as in one brain of all the flies."""

    def __init__(self, fleet, conn, b0: int, nb: int, seed: int, vision, motor_params, engine, tile, device):
        from . import BatchBrain, LIFParams
        from .batch import BatchMotor
        from .eye import BatchMotionEye

        self.b0, self.nb = b0, nb
        self.brain = BatchBrain(conn.weights, LIFParams(dt=0.25), batch=nb, seed=seed, plastic_pos=fleet.pos,
                                engine=engine, tile=tile, device=device)
        self.brain.fly_base = b0
        dev = self.brain.device
        if dev.type == "cuda" and dev.index is None:   # "cuda" != "cuda:0" in comparisons
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.eye = BatchMotionEye.for_fly(conn, FRAME_SHAPE, nb, vision, device=dev)
        self.motor = BatchMotor(conn, nb, motor_params, device=dev)
        self.brain.set_bias(fleet.bias_idx, fleet.bias_mv)
        self.eye_slots = self.brain.slots(self.eye.idx)
        self.counts = None
        self._stage = None
        self._rates = torch.zeros((nb, self.motor.rates.shape[1]), dtype=torch.float32, pin_memory=True)
        self._ready = torch.cuda.Event()

    @property
    def slots(self):
        return range(self.b0, self.b0 + self.nb)

    def frames_on_device(self, frames) -> torch.Tensor:
        """uint8 (nb, h, w, 3) frames on the device; host frames go through one pinned buffer and
        one asynchronous copy on the current stream (the buffer is refilled only once the
        previous copy has finished)."""
        if isinstance(frames, torch.Tensor) and frames.device == self.device:
            return frames
        shape = (self.nb, *FRAME_SHAPE, 3)
        if self._stage is None:
            self._stage = (torch.empty(shape, dtype=torch.uint8, pin_memory=True),
                           torch.empty(shape, dtype=torch.uint8, device=self.device), torch.cuda.Event())
        else:
            self._stage[2].synchronize()
        host, dev, copied = self._stage
        src = frames if isinstance(frames, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(frames))
        if tuple(src.shape) != shape or src.dtype != torch.uint8:
            raise ValueError(f"frames must be uint8 {shape}, got {src.dtype} {tuple(src.shape)}")
        host.copy_(src)
        dev.copy_(host, non_blocking=True)
        copied.record()
        return dev

    def see(self, frames):
        """``p_ext`` of this frame: zero, then the eye cells' probabilities."""
        self.brain.p_ext.zero_()
        self.eye.see(self.frames_on_device(frames), brain=self.brain)

    def launch(self, frames, window: float, copy: bool):
        """Queue eye, one window of the brain, the readout and its copy to the host."""
        self.see(frames)
        self.counts = self.brain.run(window, copy=copy)
        self.motor.readout.update(self.counts, window)
        self._rates.copy_(self.motor.readout.rates, non_blocking=True)
        self._ready.record()

    def collect(self) -> list:
        """Wait for this part's readout; one button dict per fly (``BatchMotor.update``'s)."""
        from .batch import BUTTONS

        self._ready.synchronize()
        self.motor.rates = self._rates.numpy().copy()
        return [dict(zip(BUTTONS, row)) for row in self.motor.masks(self.motor.rates).tolist()]


class Fleet:
    def __init__(self, pool, batch: int, seed: int = 0, plastic_types=None, bias_mv: float = 7.8, conn=None,
                 deep: bool | int = False, taps: bool = False, steer_span: float = 150.0, lean_span: float = 70.0,
                 steer_threshold: float = 10.0, lean_threshold: float = 10.0, readout_tau: float = 80.0,
                 vision: dict | None = None, segments: int = 59, engine: str = "tile", tile: int = 1024,
                 tune: bool = True, parts: int = 1, device=None):
        """flyzero/fleet.py:22-68 with ``pool`` (a game pool, see the module docstring) in place of
        the ROM, ``conn`` (a connectome; default the synthetic FlyWire-sized fly) and ``segments``
        (per lap, for ``Progress``).  ``deep``: also make the synapses onto the DNs' presynaptic
        partners plastic (2: one layer further).  ``tune``: pick the brains' launch shapes by
        timing them (cached per workload).  ``parts``: split the flies into that many brains
        that ``exam`` runs in turn, so that the host serves one part's games while the device
        simulates another (``brain``, ``eye``, ``motor`` and ``think`` need ``parts=1``; the
        parts are in ``self.parts``)."""
        from .batch import MotorParams
        from .instruct import PLASTIC_TYPES, deep2_positions, deep_positions, plastic_positions

        if conn is None:
            from .connectome import flywire_like

            conn = flywire_like()
        if not 1 <= parts <= batch:
            raise ValueError(f"parts must be in [1, batch], got {parts}")
        self.conn = conn
        self.B = batch
        self.pos = plastic_positions(conn, plastic_types or PLASTIC_TYPES)
        if deep:
            dn = self.pos
            self.pos = np.union1d(dn, deep_positions(conn, dn))
            if int(deep) >= 2:
                self.pos = np.union1d(self.pos, deep2_positions(conn, dn))
        self.bias_idx = np.concatenate([conn.find(t) for t in BIAS_TYPES])
        self.bias_mv = bias_mv
        mp = MotorParams(taps=taps, steer_span=steer_span, lean_span=lean_span, steer_threshold=steer_threshold,
                         lean_threshold=lean_threshold, tau_ms=readout_tau)
        vis = vision_params(vision)
        edges = np.linspace(0, batch, parts + 1).round().astype(int)
        self.parts = [_Part(self, conn, int(b0), int(b1 - b0), seed, vis, mp, engine, tile, device)
                      for b0, b1 in zip(edges[:-1], edges[1:])]
        one = self.parts[0] if parts == 1 else None
        self.brain = one.brain if one else None
        self.eye = one.eye if one else None
        self.motor = one.motor if one else None
        self.eye_slots = one.eye_slots if one else None
        self.device = self.parts[0].device
        self.segments = segments
        self.pool = pool
        self.window = WINDOW_MS
        self.extra_idx = np.zeros(0, np.int64)    # optional extra Poisson inputs (exploration, heat)
        if tune:
            from . import autotune

            for part in self.parts:
                # tuned on a driven workload: the eye sees a coarse random scene and the same scene
                # moved by 3 pixels (motion everywhere, as in a race); then its filters restart
                # and p_ext is cleared, so the fleet's first frame is untouched
                rng = np.random.default_rng(seed + part.b0)
                h, w = FRAME_SHAPE
                scene = rng.integers(0, 256, (part.nb, -(-h // 8), -(-w // 8), 3), dtype=np.uint8)
                scene = np.repeat(np.repeat(scene, 8, axis=1), 8, axis=2)[:, :h, :w]
                for shift in (0, 3):
                    part.see(np.ascontiguousarray(np.roll(scene, shift, axis=2)))
                autotune.tune(part.brain)
                part.brain.p_ext.zero_()
                part.eye.reset()

    def close(self):
        close = getattr(self.pool, "close", None)
        if close is not None:
            close()

    def _one(self) -> _Part:
        if len(self.parts) != 1:
            raise RuntimeError("this call needs a fleet of one part (Fleet(parts=1))")
        return self.parts[0]

    def frames_on_device(self, frames) -> torch.Tensor:
        """uint8 (B, h, w, 3) frames on the device (one pinned buffer, one asynchronous copy)."""
        return self._one().frames_on_device(frames)

    def think(self, frames, extra: np.ndarray | None = None, copy: bool = True) -> torch.Tensor:
        """One frame of brain time for every slot (flyzero/fleet.py:73-81): the eye turns
        ``frames`` into the brain's input probabilities, ``extra`` (B, len(extra_idx)) Hz adds
        to the extra inputs.  Returns the window's (B, n) spike counts on the device
        (``copy=False``: the engine's own buffer, overwritten by the next window)."""
        part = self._one()
        brain = part.brain
        if extra is not None and len(self.extra_idx):
            j = brain.slots(self.extra_idx)
        part.see(frames)
        if extra is not None and len(self.extra_idx):
            brain.p_ext[:, j] += (torch.as_tensor(np.asarray(extra, np.float32), device=self.device) *
                                  np.float32(brain.p.dt / 1000.0))
        return brain.run(self.window, copy=copy)

    def reset_slots(self, slots):
        """Brain, readout and eye filters of the given slots back to their start."""
        slots = np.atleast_1d(slots)
        for part in self.parts:
            mine = slots[(slots >= part.b0) & (slots < part.b0 + part.nb)] - part.b0
            if len(mine):
                part.brain.reset(mine)
                part.motor.reset(mine)
                part.eye.reset(mine)

    def exam(self, state, frames: int, on_frame=None, log_every: int = 0, record: bool = False,
             retire: bool = True) -> list[dict]:
        """flyzero/fleet.py:87-127: every slot drives alone from ``state`` (no learning), with its
        own spiking noise.  ``record``: each result also holds the drive's button masks and DN
        rates.  ``retire``: a slot whose race has ended stops costing brain time from the next
        window but one (its results are final; the other slots' spikes are unchanged by it).
        Each part's next window is queued as soon as its games have stepped, so with several
        parts the device simulates one part while the host serves the next.  ``on_frame(i,
        counts, buttons, infos)``: counts are the (B, n) device counts with one part, None with
        several."""
        B = self.B
        for part in self.parts:
            part.brain.reset()
            part.motor.reset()
            part.eye.reset()
        frames_now, infos = self.pool.load([state] * B)
        progs = [Progress(self.segments) for _ in range(B)]
        done = np.zeros(B, bool)
        res = [None] * B
        masks, dn = [[] for _ in range(B)], [[] for _ in range(B)]
        copy = on_frame is not None
        for part in self.parts:
            part.launch(frames_now[part.b0:part.b0 + part.nb], self.window, copy)
        t0 = time.time()
        try:
            for i in range(frames):
                buttons, infos = [None] * B, [None] * B
                counts = None
                for part in self.parts:
                    bt = part.collect()
                    sl = slice(part.b0, part.b0 + part.nb)
                    buttons[sl] = bt
                    if record:
                        for k in part.slots:
                            if not done[k]:
                                masks[k].append(buttons_to_mask(bt[k - part.b0]))
                                dn[k].append(part.motor.rates[k - part.b0].astype(np.float16))
                    counts = part.counts
                    if len(self.parts) == 1:    # the reference's pool signature
                        fr, infos[sl] = self.pool.step(bt)
                    else:
                        fr, infos[sl] = self.pool.step(bt, which=part.slots)
                    if i + 1 < frames:
                        part.launch(fr, self.window, copy)
                newly = []
                for k, info in enumerate(infos):
                    if done[k]:
                        continue
                    progs[k].update(info["segment"], i)
                    if info["done"] or i == frames - 1:
                        done[k] = True
                        newly.append(k)
                        res[k] = {"progress": progs[k].total, "lap": info["lap"], "frames": i + 1,
                                  "energy": info["energy"], "finished": bool(info.get("finished")),
                                  "laps5": info["lap"] >= 5, "rank": info.get("rank")}
                        if record:
                            res[k]["masks"], res[k]["rates"] = np.array(masks[k]), np.array(dn[k])
                if retire and newly:
                    for part in self.parts:
                        mine = [k - part.b0 for k in newly if part.b0 <= k < part.b0 + part.nb]
                        if mine:   # a pinned, asynchronous index upload: no wait for queued work
                            ix = torch.tensor(mine, dtype=torch.long).pin_memory().to(part.device, non_blocking=True)
                            part.brain.alive[ix] = 0
                if on_frame:
                    on_frame(i, counts if len(self.parts) == 1 else None, buttons, infos)
                if log_every and (i + 1) % log_every == 0:
                    print(f"  frame {i + 1}: {(i + 1) * B / (time.time() - t0):.0f} fly-frames/s, "
                          f"progress {[p.total for p in progs]}", flush=True)
                if done.all():
                    break
        finally:
            for part in self.parts:
                part.brain.alive.fill_(1)
        return res


class SyntheticRace:
    """A deterministic road game for ``n`` players with F-Zero's telemetry, standing in for the
    emulator pool where no emulator exists.

    The track is a closed loop of ``segments`` segments with curvature
    kappa(s) = 0.8 sin(4 pi s / S) + 0.4 sin(10 pi s / S + 1).  Per frame a player's speed v
    (segments per frame) rises by 0.004 with B (up to 0.25), falls by 0.01 with Y and by 0.001
    otherwise; its lateral offset x (road half-widths, + right) drifts outwards with
    0.6 v kappa(s) and is steered by 0.05 (0.2 + 4 v) (LEFT - RIGHT + 0.5 (L - R)); off the road
    (|x| > 1) it loses 0.5 energy and 3 % of its speed; x stays in [-2, 2]; s advances by v.
    The race ends with no energy, after 5 laps (finished) or at ``max_frames``.

    The view is 224 x 256 RGB: sky above row 84, below it per row y (p = (y - 83) / 140) a
    road of half-width 0.45 W p + 2 pixels centred at W / 2 - x hw + 60 kappa(s) (1 - p)^2,
    white edges of max(1, hw / 20) pixels, and grass / road shades alternating with
    floor(4 s + 8 / p) mod 2, so the road moves under the eye.  Row tables are float64 on the
    host; the picture is integer selection, drawn with torch on ``device`` (identical bits on
    every device).  States are bytes, like emulator save states; ``flips`` mirror a player's
    view only, as the emulator pool does for its eye (flyzero/pool.py:37-41)."""

    SKY = 84
    PALETTE = np.array([[90, 140, 220], [30, 120, 30], [40, 160, 40], [100, 100, 100], [120, 120, 120],
                        [250, 250, 250]], np.uint8)

    def __init__(self, n: int, segments: int = 59, max_frames: int | None = None, device="cpu",
                 frame_shape=FRAME_SHAPE):
        self.n, self.S, self.max_frames = n, segments, max_frames
        self.h, self.w = frame_shape
        self.device = torch.device(device)
        self.st = np.zeros((n, 5))      # s, x, v, energy, frame_no
        self.flip = np.zeros(n, bool)
        self._pal = torch.from_numpy(self.PALETTE).to(self.device)
        self._cols = torch.arange(self.w, device=self.device, dtype=torch.int32)
        # per-row constants of the perspective: p, half-width, bend weight, shade phase
        p = (np.arange(self.SKY, self.h, dtype=np.float64) - (self.SKY - 1)) / (self.h - self.SKY)
        self._hw = (0.45 * self.w * p + 2.0)[None, :]
        self._bend = (60.0 * (1.0 - p) ** 2)[None, :]
        self._phase = (8.0 / p)[None, :]
        self._edge = np.maximum(1, np.floor(self._hw / 20.0)).astype(np.int64)

    @staticmethod
    def start_state(s: float = 0.0, x: float = 0.0, v: float = 0.0, energy: float = 100.0) -> bytes:
        return np.array([s, x, v, energy, 0.0]).tobytes()

    def kappa(self, s):
        a = 2 * np.pi * np.asarray(s, np.float64) / self.S
        return 0.8 * np.sin(2 * a) + 0.4 * np.sin(5 * a + 1.0)

    def _done(self, which):
        s, e, f = self.st[which, 0], self.st[which, 3], self.st[which, 4]
        finished = s // self.S >= 5
        done = (e <= 0) | finished
        if self.max_frames is not None:
            done |= f >= self.max_frames
        return done, finished

    def _infos(self, which):
        which = np.asarray(list(which), np.int64)
        done, finished = self._done(which)
        s = self.st[which, 0]
        cols = zip((s.astype(np.int64) % self.S).tolist(), (s // self.S).astype(np.int64).tolist(),
                   self.st[which, 3].tolist(), done.tolist(), finished.tolist(), self.st[which, 4].astype(np.int64).tolist())
        return [{"segment": seg, "lap": lap, "energy": e, "done": d, "finished": fin, "rank": None, "frame": f}
                for seg, lap, e, d, fin, f in cols]

    def rows(self, which=None):
        """Per-row integer tables (left, right, edge, shade) of each player's view, [len(which), rows]."""
        which = np.arange(self.n) if which is None else np.asarray(which)
        s, x = self.st[which, 0:1], self.st[which, 1:2]
        hw = self._hw
        cx = self.w / 2 - x * hw + self.kappa(s) * self._bend
        left = np.floor(cx - hw).astype(np.int64)
        right = np.floor(cx + hw).astype(np.int64)
        edge = np.broadcast_to(self._edge, left.shape)
        shade = np.floor(4.0 * s + self._phase).astype(np.int64) % 2
        return left, right, edge, shade

    def render(self, which=None) -> torch.Tensor:
        """The views of players ``which`` (default all), uint8 (m, h, w, 3) on ``device``.  The
        row tables and flip flags go up in one pinned, asynchronous copy, so drawing does not
        wait for work already queued on the device."""
        idx = np.arange(self.n) if which is None else np.asarray(which)
        tabs = np.stack(self.rows(idx)).astype(np.int32)            # (4, m, rows)
        flips = self.flip[idx]
        packed = torch.from_numpy(np.concatenate([tabs.ravel(), flips.astype(np.int32)]))
        if self.device.type == "cuda":
            packed = packed.pin_memory().to(self.device, non_blocking=True)
        m, R = tabs.shape[1], tabs.shape[2]
        left, right, edge, shade = (packed[k * m * R:(k + 1) * m * R].view(m, R)[..., None] for k in range(4))
        c = self._cols
        road = (c >= left) & (c < right)
        rim = road & ((c < left + edge) | (c >= right - edge))
        col = torch.where(rim, 5, torch.where(road, 3 + shade, 1 + shade))
        ground = self._pal[col.long()]
        sky = self._pal[0].expand(m, self.SKY, self.w, 3)
        img = torch.cat([sky, ground], 1)
        if flips.any():
            f = packed[4 * m * R:].bool()
            img = torch.where(f[:, None, None, None], img.flip(2), img)
        return img.contiguous()

    def load(self, states, flips=None, which=None):
        which = list(range(self.n)) if which is None else list(which)
        flips = flips or [False] * len(which)
        for k, st, f in zip(which, states, flips):
            self.st[k] = np.frombuffer(st, np.float64, count=5)
            self.st[k, 4] = 0.0
            self.flip[k] = f
        return self.render(which), self._infos(which)

    def save(self, which=None):
        which = range(self.n) if which is None else which
        return [self.st[k].tobytes() for k in which]

    KEYS = ("LEFT", "RIGHT", "L", "R", "B", "Y")

    def step(self, buttons, which=None):
        """Advance the games ``which`` (default all) by one frame with one button dict each."""
        which = np.arange(self.n) if which is None else np.asarray(list(which), np.int64)
        if len(buttons) != len(which):
            raise ValueError(f"one button dict per game ({len(which)}), got {len(buttons)}")
        P = np.array([[b.get(k, False) for k in self.KEYS] for b in buttons], np.float64).reshape(len(which), 6)
        steer = (P[:, 0] - P[:, 1]) + 0.5 * (P[:, 2] - P[:, 3])
        s, x, v, e, f = (self.st[which, i].copy() for i in range(5))
        live = ~self._done(which)[0]
        v = np.where(P[:, 4] > 0, np.minimum(v + 0.004, 0.25), np.where(P[:, 5] > 0, v - 0.01, v - 0.001))
        v = np.maximum(v, 0.0)
        x = x + 0.6 * v * self.kappa(s) - 0.05 * (0.2 + 4.0 * v) * steer
        off = np.abs(x) > 1.0
        e = np.where(off, e - 0.5, e)
        v = np.where(off, 0.97 * v, v)
        x = np.clip(x, -2.0, 2.0)
        s = s + v
        new = np.stack([s, x, v, e, f + 1], 1)
        self.st[which[live]] = new[live]
        return self.render(which if len(which) < self.n else None), self._infos(which)

    def close(self):
        pass


def _game_worker(conn, factory, shm_name, n, k, frame_shape):
    """One game in its own process (flyzero/pool.py:22-130): ``factory()`` makes a one-player
    pool; its frames go into row ``k`` of the shared (n, h, w, 3) block, its telemetry and saved
    states through the pipe."""
    torch.set_num_threads(1)
    game = factory()
    shm = shared_memory.SharedMemory(name=shm_name)
    frames = np.ndarray((n, *frame_shape, 3), np.uint8, buffer=shm.buf)

    def put(out):
        fr, infos = out
        frames[k] = fr[0].cpu().numpy() if isinstance(fr, torch.Tensor) else np.asarray(fr)[0]
        return infos[0]

    try:
        while True:
            msg = conn.recv()
            cmd = msg[0]
            if cmd == "load":
                conn.send(put(game.load([msg[1]], flips=[msg[2]])))
            elif cmd == "step":
                conn.send(put(game.step([msg[1]])))
            elif cmd == "save":
                conn.send(game.save([0])[0])
            elif cmd == "close":
                return
    finally:
        del frames
        shm.close()
        conn.close()


class GamePool:
    """``EmulatorPool`` (flyzero/pool.py:133-203) for any game: ``factory()`` (picklable) makes a
    one-player pool (``load`` / ``step`` / ``save`` of ``SyntheticRace``'s form, e.g. an emulator
    wrapper); each of the ``n`` games runs in its own spawned process.  The frames come back in
    one shared (n, h, w, 3) uint8 block, whose view ``load`` and ``step`` return (valid until the
    next call), so the fleet copies them to the device in one piece."""

    def __init__(self, factory, n: int, frame_shape=FRAME_SHAPE):
        ctx = mp.get_context("spawn")
        self.n, self.frame_shape = n, tuple(frame_shape)
        size = n * self.frame_shape[0] * self.frame_shape[1] * 3
        self._shm = shared_memory.SharedMemory(create=True, size=max(1, size))
        self.frames = np.ndarray((n, *self.frame_shape, 3), np.uint8, buffer=self._shm.buf)
        self.pipes, self.procs = [], []
        for k in range(n):
            a, b = ctx.Pipe()
            p = ctx.Process(target=_game_worker, args=(b, factory, self._shm.name, n, k, self.frame_shape),
                            daemon=True)
            p.start()
            b.close()
            self.pipes.append(a)
            self.procs.append(p)

    def _all(self, msgs, which=None):
        which = list(range(self.n)) if which is None else list(which)
        msgs = list(msgs)
        if len(msgs) != len(which):
            raise ValueError("one message per worker")
        for k, m in zip(which, msgs):
            self.pipes[k].send(m)
        return [self.pipes[k].recv() for k in which]

    def load(self, states, flips=None, which=None):
        which = list(range(self.n)) if which is None else list(which)
        flips = flips or [False] * len(which)
        infos = self._all([("load", s, f) for s, f in zip(states, flips)], which)
        return self._view(which), infos

    def step(self, buttons, which=None):
        """One frame of the games ``which`` (default all); returns their frames (a view of the
        shared block for a contiguous range) and telemetry."""
        which = list(range(self.n)) if which is None else list(which)
        infos = self._all([("step", b) for b in buttons], which)
        return self._view(which), infos

    def _view(self, which):
        if which and which == list(range(which[0], which[-1] + 1)):
            return self.frames[which[0]:which[-1] + 1]
        return self.frames[which]

    def save(self, which=None):
        which = list(range(self.n)) if which is None else list(which)
        return self._all([("save",)] * len(which), which)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        """Stop the workers and release the shared frame block (idempotent)."""
        if self._shm is None:
            return
        for p in self.pipes:
            try:
                p.send(("close",))
            except (BrokenPipeError, OSError):
                pass
        for p in self.procs:
            p.join(timeout=10)
            if p.is_alive():
                p.terminate()
                p.join(timeout=5)
        for c in self.pipes:
            c.close()
        self.frames = None
        try:
            self._shm.close()
        except BufferError:       # a caller still holds a frames view; the mapping goes with it
            pass
        self._shm.unlink()
        self._shm = None
