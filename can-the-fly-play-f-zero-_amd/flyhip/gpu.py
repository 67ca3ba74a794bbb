"""``BatchBrain``: B flies sharing one connectome, on the MI355X engine.

Drop-in for flyzero/gpu.py:131-254 (``BatchBrain``): same constructor, methods (``reset``,
``slots``, ``set_bias``, ``set_input``, ``run``) and attributes (``B, n, p, indptr, indices,
data, plastic_pos, pl_pre, pl_post, w0, pw, p_ext, bias_ext, ext, seed, step_no, t, v, g, ref``)
with torch tensors on the GPU where the reference holds CuPy arrays.  Differences, all
deliberate:

* the membrane state is ``u = v - v_rest`` (float32); ``v`` is an mV view over it
  (``flyhip.state.MembraneView``) that supports slice reads and writes;
* weights are integer synapse counts accumulated in an int32 ring (``flyhip.params``); the
  ring is ``ring`` [B, n_delay, n] int32, and ``pending`` is a read-only mV view of it;
* ``run`` replays one HIP graph per window length (kernels read inputs from device memory, so
  in-place edits of ``p_ext``/``bias_ext``/``pw`` between frames are seen);
* ``alive`` (uint8 [B]) retires flies: a retired fly is skipped (counts 0) until reset and
  revived, which is how flies with uneven episode lengths share one batch;
* no ``B * n < 2**24`` limit (flyzero/gpu.py:168): the spike work list is 64-bit addressed.
"""

from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from .params import LIFParams, derived, frac_bits_for, plastic_frac_bits, quantize, scales
from .state import MembraneView


def tile_tables(indptr: np.ndarray, col: np.ndarray, n: int, tile: int, pl_post: np.ndarray):
    """Tile engine tables.  tb[i, t] = first edge of row i (fixed CSR, columns sorted within rows)
    whose post neuron is >= t * tile, for t = 0 .. n_tiles; and the plastic post neurons of each
    tile as a CSR (ptg_ptr, ptg_idx)."""
    nt = (n + tile - 1) // tile
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
    key = rows * (nt + 1) + np.asarray(col, np.int64) // tile
    tb = np.empty((n, nt + 1), np.int32)
    step = max(1, (1 << 22) // (nt + 1))
    for r0 in range(0, n, step):
        r1 = min(n, r0 + step)
        q = (np.arange(r0, r1, dtype=np.int64)[:, None] * (nt + 1) + np.arange(nt + 1)[None, :]).ravel()
        tb[r0:r1] = np.searchsorted(key, q, side="left").reshape(r1 - r0, nt + 1)
    upost = np.unique(np.asarray(pl_post, np.int64))
    ptg_ptr = np.concatenate([[0], np.cumsum(np.bincount(upost // tile, minlength=nt))]).astype(np.int32)
    return tb, ptg_ptr, upost.astype(np.int32)


class BatchBrain:
    """``batch`` flies with shared weights and private plastic synapses.

    ``plastic_pos``: indices into the CSR ``data`` (``Brain.data`` layout) of synapses that may
    differ between flies.  ``self.pw`` (batch, n_plastic) holds their weights in mV.
    """

    def __init__(self, weights: sp.csr_matrix, params: LIFParams | None = None, batch: int = 1,
                 seed: int = 0, plastic_pos: np.ndarray | None = None, frac_bits: int | None = None,
                 device=None, engine: str = "push", tile: int = 1024):
        """``engine``: ``"push"`` (spikes scattered into an int32 delay ring with global atomics) or
        ``"tile"`` (arrivals pulled into LDS per ``tile`` neurons from per-step spike lists; fixed
        weights are then immutable, which the reference's ``BatchBrain`` also assumes)."""
        if engine not in ("push", "tile"):
            raise ValueError(f"engine must be 'push' or 'tile', got {engine!r}")
        self.engine = engine
        if not torch.cuda.is_available():
            raise RuntimeError("flyhip.BatchBrain needs a HIP device (there is no CPU fallback)")
        self._lib = lib = _lib.load()
        self.device = torch.device(device if device is not None else "cuda")
        self.p = p = params or LIFParams()
        if batch < 1:
            raise ValueError("batch must be >= 1")
        self.B = batch
        w = sp.csr_matrix(weights, dtype=np.float32)
        w.sort_indices()
        self.n = n = w.shape[0]
        if w.shape[1] != n:
            raise ValueError("weights must be square (pre x post)")
        # host copies with the reference's layout (plasticity code indexes into them)
        self.indptr = w.indptr.astype(np.int64)
        self.indices = w.indices.astype(np.int32)
        self.data = (w.data * p.w_syn).astype(np.float32)
        count = w.data.astype(np.float32)
        pos = np.zeros(0, np.int64) if plastic_pos is None else np.sort(np.asarray(plastic_pos, np.int64))
        if len(pos) and (pos[0] < 0 or pos[-1] >= len(count) or np.any(np.diff(pos) == 0)):
            raise ValueError("plastic_pos must be distinct CSR positions")
        self.plastic_pos = pos
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(self.indptr))
        self.pl_pre = rows[pos]
        self.pl_post = self.indices[pos]
        self.w0 = self.data[pos].copy()
        self.w0_syn = count[pos].copy()          # the same weights in synapses

        c = derived(p)
        self.n_delay, self.n_ref = c["n_delay"], c["n_ref"]
        self.decay_m, self.decay_syn, self.kick = c["decay_m"], c["decay_syn"], c["kick"]
        if not 0 <= self.n_ref <= 127:
            raise ValueError("refractory counter is int8 on the GPU: need round(t_ref/dt) <= 127")
        if self.n_delay > 256:
            raise ValueError("delay ring longer than 256 steps is not supported")
        if frac_bits is None:
            frac_bits = frac_bits_for(count, self.indices, n)
            if frac_bits == 0 and len(pos):   # sub-synapse resolution for the plastic weights
                frac_bits = plastic_frac_bits(count, self.indices, n, int(tile) if engine == "tile" else None)
        self.frac_bits = int(frac_bits)
        self.in_scale, self.pw_qscale = scales(p.w_syn, self.frac_bits)

        # fixed (non-plastic) CSR, integer weights
        keep = np.ones(len(count), bool)
        keep[pos] = False
        q = quantize(count[keep], self.frac_bits)
        fixed_indptr = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=n))])
        if q.size and np.abs(q).max() >= 2 ** 31:
            raise ValueError("weights do not fit the int32 fixed-point format")
        self.w_bits = 16 if (self.frac_bits == 0 and (q.size == 0 or np.abs(q).max() < 2 ** 15)) else 32
        dev = self.device
        self.d_indptr = torch.from_numpy(fixed_indptr.astype(np.int32)).to(dev)
        self.d_col = torch.from_numpy(self.indices[keep].astype(np.int32)).to(dev)
        self.d_w = torch.from_numpy(q.astype(np.int16 if self.w_bits == 16 else np.int32)).to(dev)
        pl_indptr = np.searchsorted(self.pl_pre, np.arange(n + 1)).astype(np.int32)
        self.d_pl_indptr = torch.from_numpy(pl_indptr).to(dev)
        self.d_pl_post = torch.from_numpy(self.pl_post.astype(np.int32)).to(dev)
        self.pw = torch.from_numpy(np.tile(self.w0, (batch, 1)).astype(np.float32)).to(dev)

        self.seed = np.uint32(seed)
        # the Poisson draw numbers fly b as b + fly_base: a brain holding flies [fly_base,
        # fly_base + batch) of a larger batch draws the same noise as that batch does for them
        self.fly_base = 0
        self._step_no = 0
        self._t = 0
        # neurons that take outside input (Poisson rates and/or a bias): "ext" slots
        self.ext = np.zeros(0, np.int64)
        self._slot = np.full(n, -1, np.int32)
        self.d_ext_idx = torch.from_numpy(self._slot.copy()).to(dev)
        self.p_ext = torch.zeros((batch, 0), dtype=torch.float32, device=dev)     # per-step probability
        self.bias_ext = torch.zeros((batch, 0), dtype=torch.float32, device=dev)  # mV
        self.alive = torch.ones(batch, dtype=torch.uint8, device=dev)

        # caller-owned state (the engine only holds pointers); u = v - v_rest
        self.u = torch.zeros((batch, n), dtype=torch.float32, device=dev)
        self.g = torch.zeros((batch, n), dtype=torch.float32, device=dev)
        self.ref = torch.zeros((batch, n), dtype=torch.int8, device=dev)
        self.ring_slots = 2 * self.n_delay if engine == "tile" else self.n_delay
        self.ring = torch.zeros((batch, self.ring_slots, n), dtype=torch.int32, device=dev)
        self.spk_idx = self.spk_cnt = None
        if engine == "tile":   # in-flight spikes: per fly, per step (3 n_delay slots), neuron ids
            self.spk_idx = torch.zeros((batch, 3 * self.n_delay, n), dtype=torch.int32, device=dev)
            self.spk_cnt = torch.zeros((batch, 3 * self.n_delay), dtype=torch.int32, device=dev)
        self.counts = torch.zeros((batch, n), dtype=torch.int32, device=dev)
        self.clock = torch.zeros(4, dtype=torch.int64, device=dev)

        self.u_th = np.float32(np.float32(p.v_th) - np.float32(p.v_rest))
        self.u_reset = np.float32(np.float32(p.v_reset) - np.float32(p.v_rest))
        prm = _lib.FzParams(self.decay_syn, self.decay_m, self.u_th, self.u_reset, self.kick, self.in_scale,
                            self.pw_qscale, self.n_ref, self.n_delay)
        h = ctypes.c_void_p()
        torch.cuda.synchronize(dev)
        _lib.check(lib.fz_create(n, int(self.d_col.numel()), _lib.ptr(self.d_indptr), _lib.ptr(self.d_col),
                                 _lib.ptr(self.d_w), self.w_bits, self.frac_bits, ctypes.byref(prm),
                                 ctypes.byref(h)), "fz_create")
        self._h = h
        _lib.check(lib.fz_set_plastic(h, len(pos), _lib.ptr(self.d_pl_indptr) if len(pos) else None,
                                      _lib.ptr(self.d_pl_post), _lib.stream_handle(dev)), "fz_set_plastic")
        if engine == "tile":
            self.tile = int(tile)
            tb, ptg_ptr, ptg_idx = tile_tables(fixed_indptr, self.indices[keep], n, self.tile, self.pl_post)
            self.d_tb = torch.from_numpy(tb).to(dev)
            self.d_ptg_ptr = torch.from_numpy(ptg_ptr).to(dev)
            self.d_ptg_idx = torch.from_numpy(ptg_idx).to(dev)
            torch.cuda.synchronize(dev)
            _lib.check(lib.fz_set_tiles(h, self.tile, tb.shape[1] - 1, _lib.ptr(self.d_tb),
                                        _lib.ptr(self.d_ptg_ptr) if len(pos) else None,
                                        _lib.ptr(self.d_ptg_idx) if len(pos) else None), "fz_set_tiles")
        self._st = _lib.FzState(batch, self.u.data_ptr(), self.g.data_ptr(), self.ref.data_ptr(),
                                self.ring.data_ptr(), self.counts.data_ptr(), self.clock.data_ptr(),
                                self.ring_slots, _lib.ptr(self.spk_idx), _lib.ptr(self.spk_cnt))
        _lib.check(lib.fz_bind(h, ctypes.byref(self._st)), "fz_bind")
        self.tune = _lib.FzTune()  # all zero: engine defaults
        self._graphs: dict = {}

    # ------------------------------------------------------------------ clock
    @property
    def t(self) -> int:
        return self._t

    @t.setter
    def t(self, value: int):
        self._t = int(value)
        self.clock[0] = self._t

    @property
    def step_no(self) -> int:
        return self._step_no

    @step_no.setter
    def step_no(self, value: int):
        self._step_no = int(value)
        self.clock[1] = self._step_no & 0xFFFFFFFF

    @property
    def v(self) -> MembraneView:
        """Membrane potential in mV (a view over ``u``; slice reads and writes)."""
        return MembraneView(self.u, self.p.v_rest)

    @property
    def pending(self) -> torch.Tensor:
        """Delay ring in mV, (n_delay, B, n) like the reference; a read-only copy (push engine)."""
        if self.engine != "push":
            raise AttributeError("the tile engine keeps in-flight spikes as lists (spk_idx/spk_cnt), not a ring")
        return (self.ring.to(torch.float32) * float(self.in_scale)).transpose(0, 1).contiguous()

    # ------------------------------------------------------------------ state
    def reset(self, flies=None):
        """Back to rest (all flies, or the given ones). Plastic weights are kept."""
        if flies is None:
            self.u.zero_()
            self.g.zero_()
            self.ref.zero_()
            self.ring.zero_()
            if self.spk_cnt is not None:
                self.spk_cnt.zero_()
            self.t = 0
            return
        if isinstance(flies, torch.Tensor):  # a device index tensor: no host round trip
            f = flies.to(self.device, torch.long)
        else:
            f = torch.as_tensor(np.atleast_1d(flies), dtype=torch.long, device=self.device)
        self.u[f] = 0.0
        self.g[f] = 0.0
        self.ref[f] = 0
        self.ring[f] = 0
        if self.spk_cnt is not None:
            self.spk_cnt[f] = 0

    def slots(self, idx) -> torch.Tensor:
        """Ext slots of neurons ``idx`` (adding any that don't have one yet)."""
        idx = np.asarray(idx, np.int64).ravel()
        new = np.unique(idx[self._slot[idx] < 0])
        if len(new):
            k = len(self.ext)
            self.ext = np.r_[self.ext, new]
            self._slot[new] = np.arange(k, k + len(new), dtype=np.int32)
            self.d_ext_idx = torch.from_numpy(self._slot.copy()).to(self.device)
            pad = torch.zeros((self.B, len(new)), dtype=torch.float32, device=self.device)
            self.p_ext = torch.cat([self.p_ext, pad], 1)
            self.bias_ext = torch.cat([self.bias_ext, pad], 1)
        return torch.from_numpy(self._slot[idx].astype(np.int64)).to(self.device)

    def set_bias(self, idx, mv, flies=None):
        """Steady depolarisation (mV) for neurons ``idx``, as ``Brain.set_bias`` (replaces it)."""
        j = self.slots(idx)
        if flies is None:
            self.bias_ext.zero_()
            self.bias_ext[:, j] = torch.as_tensor(mv, dtype=torch.float32, device=self.device)
        else:
            f = torch.as_tensor(np.atleast_1d(flies), dtype=torch.long, device=self.device)
            self.bias_ext[f] = 0.0
            self.bias_ext[f[:, None], j[None, :]] = torch.as_tensor(mv, dtype=torch.float32, device=self.device)

    def set_input(self, idx, rate_hz):
        """Poisson drive (Hz): ``rate_hz`` (batch, len(idx)) or broadcastable, replacing earlier input."""
        j = self.slots(idx)
        self.p_ext.zero_()
        r = torch.as_tensor(np.asarray(rate_hz, np.float32), device=self.device)
        self.p_ext[:, j] = r * np.float32(self.p.dt / 1000.0)

    # ------------------------------------------------------------------ stepping
    def _inputs(self) -> _lib.FzInputs:
        n_ext = len(self.ext)
        if n_ext:
            for name in ("p_ext", "bias_ext"):
                t = getattr(self, name)
                if t.shape != (self.B, n_ext) or t.dtype != torch.float32 or not t.is_contiguous():
                    setattr(self, name, t.to(torch.float32).reshape(self.B, n_ext).contiguous())
        if len(self.plastic_pos):
            if self.pw.shape != (self.B, len(self.plastic_pos)) or self.pw.dtype != torch.float32 \
                    or not self.pw.is_contiguous():
                self.pw = self.pw.to(torch.float32).reshape(self.B, -1).contiguous()
        if self.alive.dtype != torch.uint8 or not self.alive.is_contiguous():
            self.alive = self.alive.to(torch.uint8).contiguous()
        return _lib.FzInputs(n_ext, _lib.ptr(self.d_ext_idx) if n_ext else None,
                             _lib.ptr(self.p_ext) if n_ext else None, _lib.ptr(self.bias_ext) if n_ext else None,
                             _lib.ptr(self.pw) if len(self.plastic_pos) else None, _lib.ptr(self.alive),
                             int(self.seed), int(self.fly_base) & 0xFFFFFFFF)

    def _advance(self, steps: int):
        self._t += steps
        self._step_no = (self._step_no + steps) & 0xFFFFFFFF

    def steps_of(self, duration_ms: float) -> int:
        return max(1, int(round(duration_ms / self.p.dt)))

    def run(self, duration_ms: float, graph: bool = True, copy: bool = True) -> torch.Tensor:
        """Advance every fly; returns spike counts (batch, n) int32 on the GPU for this window.

        ``copy=False`` returns the engine's own counts buffer (overwritten by the next run)."""
        steps = self.steps_of(duration_ms)
        inp = self._inputs()
        stream = _lib.stream_handle(self.device)
        if graph:
            gr = self._graph(steps, inp)
            _lib.check(self._lib.fz_graph_launch(gr, stream), "fz_graph_launch")
        else:
            _lib.check(self._lib.fz_run(self._h, ctypes.byref(inp), steps, None, None, ctypes.byref(self.tune),
                                        stream), "fz_run")
        self._advance(steps)
        return self.counts.clone() if copy else self.counts

    def run_steps(self, steps: int, hits: torch.Tensor | None = None, raster: bool = False):
        """Eager run of ``steps`` steps, optionally with precomputed Poisson hits
        (int32 [steps, B, ceil(n_ext/32)], bit j of a word = ext slot j kicked) and returning the
        raster (int32 [steps, B, ceil(n/32)], bit i%32 of word i/32 = neuron i spiked)."""
        inp = self._inputs()
        hw = (len(self.ext) + 31) // 32
        if hits is not None:
            if tuple(hits.shape) != (steps, self.B, hw) or hits.dtype != torch.int32:
                raise ValueError(f"hits must be int32 {(steps, self.B, hw)}, got {hits.dtype} {tuple(hits.shape)}")
            hits = hits.to(self.device).contiguous()
        ras = None
        if raster:
            ras = torch.zeros((steps, self.B, (self.n + 31) // 32), dtype=torch.int32, device=self.device)
        _lib.check(self._lib.fz_run(self._h, ctypes.byref(inp), steps, _lib.ptr(hits), _lib.ptr(ras),
                                    ctypes.byref(self.tune), _lib.stream_handle(self.device)), "fz_run")
        self._advance(steps)
        return (self.counts.clone(), ras) if raster else self.counts.clone()

    def run_timed(self, steps: int) -> dict:
        """Eager window with every kernel bracketed by HIP events (waits for the stream)."""
        inp = self._inputs()
        um, pm = ctypes.c_float(), ctypes.c_float()
        un, pn = ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.fz_run_timed(self._h, ctypes.byref(inp), steps, ctypes.byref(self.tune),
                                          _lib.stream_handle(self.device), ctypes.byref(um), ctypes.byref(un),
                                          ctypes.byref(pm), ctypes.byref(pn)), "fz_run_timed")
        self._advance(steps)
        return {"update_ms": um.value, "update_launches": un.value, "push_ms": pm.value, "push_launches": pn.value}

    def check(self):
        """Raise if a device-side capacity check fired (waits for the stream)."""
        _lib.check(self._lib.fz_check(self._h, _lib.stream_handle(self.device)), "fz_check")

    def set_tune(self, upd_threads=0, push_threads=0, push_chunk=0, push_grid=0, block_steps=0, tile_stage=0,
                 tile_source=0, plastic_pull=0, block_parts=0, acc_bits=0):
        self.tune = _lib.FzTune(upd_threads, push_threads, push_chunk, push_grid, block_steps, tile_stage,
                                tile_source, plastic_pull, block_parts, acc_bits)
        self._drop_graphs()

    def _graph(self, steps: int, inp: _lib.FzInputs):
        key = (steps, inp.n_ext, inp.ext_idx, inp.p_ext, inp.bias_ext, inp.pw, inp.alive, inp.seed, inp.fly_base,
               tuple(getattr(self.tune, f) for f, _ in _lib.FzTune._fields_))
        gr = self._graphs.get(key)
        if gr is None:
            if len(self._graphs) > 16:
                self._drop_graphs()
            h = ctypes.c_void_p()
            _lib.check(self._lib.fz_graph_create(self._h, ctypes.byref(inp), steps, ctypes.byref(self.tune),
                                                 ctypes.byref(h)), "fz_graph_create")
            gr = self._graphs[key] = h
        return gr

    def _drop_graphs(self):
        if self._graphs:
            torch.cuda.synchronize(self.device)
        for h in self._graphs.values():
            self._lib.fz_graph_destroy(h)
        self._graphs = {}

    def close(self):
        if getattr(self, "_h", None) is not None:
            torch.cuda.synchronize(self.device)
            self._drop_graphs()
            self._lib.fz_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
