"""Learned descending-neuron rate surrogate, trained and run with this package's own kernels.

Maps the pooled visual drive of a frame (the 12,246 input neurons' rates averaged onto a 64 x 32
retinotopic grid = 2,048 values) to the 8 DN group rates (flyzero/instruct.py:43-48) that the
full LIF simulation produces for that frame.  Network:

    h1 = gelu(x W1^T + b1)       2048 -> 1024
    h2 = gelu(h1 W2^T + b2)      1024 -> 1024
    s' = GRU(h2, s)              1024 -> 256, carried across frames
    y  = [h2, s'] Wo^T + bo      1280 -> 8 rates (scaled by ``rate_scale``)

Every GEMM is ``fz_sg_gemm`` (BF16 operands on MFMA, FP32 accumulation, fused bias + GELU);
back-propagation through time, the MSE loss and Adam are the ``fz_sg_*`` kernels.  FP32 master
weights live in one flat buffer (one Adam launch per step) with a BF16 copy for the GEMMs.  The
output layer is padded to 32 rows so every GEMM's K is a multiple of 32; padded rows stay zero.

The full simulation is the fallback and the judge: ``Ensemble.predict`` reports, per frame, the
disagreement of independently trained members; frames above a threshold are handed back to
the simulation.  There is no counterpart in the reference (its nearest ancestor is the PyTorch
MLP of flyzero/implant.py:136-154).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib

N_OUT_PAD = 32
STORE32, ACC32, STORE16, GELU, GELU16 = 0, 1, 2, 3, 4


def _p(t):
    return None if t is None else t.data_ptr()


class RateSurrogate:
    def __init__(self, n_in=2048, hidden=1024, gru=256, n_out=8, seed=0, device="cuda", rate_scale=100.0):
        for d in (n_in, hidden, gru):
            if d % 32:
                raise ValueError("n_in, hidden and gru must be multiples of 32")
        if n_out > N_OUT_PAD:
            raise ValueError(f"n_out <= {N_OUT_PAD}")
        self.n_in, self.hid, self.H, self.n_out = n_in, hidden, gru, n_out
        self.rate_scale = float(rate_scale)
        self.dev = torch.device(device)
        self.lib = _lib.load()
        H3 = 3 * gru
        self.layout = [("W1", (hidden, n_in)), ("b1", (hidden,)), ("W2", (hidden, hidden)), ("b2", (hidden,)),
                       ("Wih", (H3, hidden)), ("bih", (H3,)), ("Whh", (H3, gru)), ("bhh", (H3,)),
                       ("Wo", (N_OUT_PAD, hidden + gru)), ("bo", (N_OUT_PAD,))]
        self.off, o = {}, 0
        for name, shp in self.layout:
            self.off[name] = (o, shp)
            o += int(np.prod(shp))
        self.size = o
        # PyTorch-style init: U(-1/sqrt(fan_in), 1/sqrt(fan_in)); GRU U(-1/sqrt(H), 1/sqrt(H))
        rng = np.random.default_rng(seed)
        w = np.zeros(o, np.float32)
        fan = {"W1": n_in, "b1": n_in, "W2": hidden, "b2": hidden, "Wih": gru, "bih": gru, "Whh": gru, "bhh": gru,
               "Wo": hidden + gru, "bo": hidden + gru}
        for name, shp in self.layout:
            a, _ = self.off[name]
            k = 1.0 / math.sqrt(fan[name])
            v = rng.uniform(-k, k, shp).astype(np.float32)
            if name in ("Wo", "bo"):
                v[n_out:] = 0.0
            w[a:a + v.size] = v.ravel()
        self.w = torch.from_numpy(w).to(self.dev)
        self.w16 = torch.empty(o, dtype=torch.bfloat16, device=self.dev)
        self.grad = torch.zeros(o, dtype=torch.float32, device=self.dev)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.step_count = 0
        self._sync16()
        self._bufs = {}
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.dev)

    # views ---------------------------------------------------------------------------------
    def view(self, name, half=False):
        a, shp = self.off[name]
        t = self.w16 if half else self.w
        return t[a:a + int(np.prod(shp))].view(shp)

    def gview(self, name):
        a, shp = self.off[name]
        return self.grad[a:a + int(np.prod(shp))].view(shp)

    def params(self) -> dict:
        return {name: self.view(name).detach().cpu().numpy().astype(np.float64) for name, _ in self.layout}

    def grads(self) -> dict:
        return {name: self.gview(name).detach().cpu().numpy().astype(np.float64) for name, _ in self.layout}

    def _sync16(self):
        _lib.check(self.lib.fz_sg_cast(self.size, _p(self.w), _p(self.w16), self._stream()), "fz_sg_cast")

    def _stream(self):
        return _lib.stream_handle(self.dev)

    # kernels -------------------------------------------------------------------------------
    def _gemm(self, M, N, K, A, lda, ta, B, ldb, tb, mode, C32=None, C16=None, ldc=None, bias=None, a_off=0,
              b_off=0, c_off=0):
        """Operands as tensors plus element offsets (views into flat buffers)."""
        ap = A.data_ptr() + a_off * A.element_size()
        bp = B.data_ptr() + b_off * B.element_size()
        c32 = None if C32 is None else C32.data_ptr() + c_off * 4
        c16 = None if C16 is None else C16.data_ptr() + c_off * 2
        _lib.check(self.lib.fz_sg_gemm(M, N, K, ap, lda, ta, bp, ldb, tb, mode, c32, c16, ldc if ldc else N,
                                       _p(bias), self._stream()), "fz_sg_gemm")

    def _buffers(self, T, B, train):
        key = (T, B, train)
        if key not in self._bufs:
            f32, b16, d = torch.float32, torch.bfloat16, self.dev
            TB, hid, H = T * B, self.hid, self.H
            z = lambda *s, dt=f32: torch.zeros(s, dtype=dt, device=d)  # noqa: E731
            bf = {"pre1": z(TB, hid), "h1": z(TB, hid, dt=b16), "pre2": z(TB, hid), "h2": z(TB, hid, dt=b16),
                  "gx": z(TB, 3 * H), "gh": z(B, 3 * H), "S": z(T + 1, B, H), "S16": z(T + 1, B, H, dt=b16),
                  "y": z(TB, N_OUT_PAD)}
            if train:
                bf.update({"gates": z(T, B, 4 * H), "dy": z(TB, N_OUT_PAD), "dy16": z(TB, N_OUT_PAD, dt=b16),
                           "dh2": z(TB, hid), "dS": z(T, B, H), "dh0": z(B, H), "dgx": z(TB, 3 * H),
                           "dgh": z(TB, 3 * H), "dgx16": z(TB, 3 * H, dt=b16), "dgh16": z(TB, 3 * H, dt=b16),
                           "dp2": z(TB, hid), "dp2_16": z(TB, hid, dt=b16), "dh1": z(TB, hid),
                           "dp1": z(TB, hid), "dp1_16": z(TB, hid, dt=b16)})
            self._bufs[key] = bf
        return self._bufs[key]

    # forward / backward --------------------------------------------------------------------
    def forward(self, X16: torch.Tensor, h0: torch.Tensor, train: bool = False):
        """X16: bf16 [T, B, n_in] (time-major); h0: f32 [B, H].  Returns y [T, B, n_out]
        (scaled rates) and the final state (a view into a reused buffer)."""
        T, B, n_in = X16.shape
        if n_in != self.n_in or (train and (T * B) % 32):
            raise ValueError("bad input shape (training needs T*B a multiple of 32)")
        X16 = X16.contiguous()
        bf = self._buffers(T, B, train)
        TB, hid, H = T * B, self.hid, self.H
        w16 = self.w16
        o = {k: v[0] for k, v in self.off.items()}
        self._gemm(TB, hid, n_in, X16, n_in, 0, w16, n_in, 1, GELU, bf["pre1"], bf["h1"], hid, self.view("b1"),
                   b_off=o["W1"])
        self._gemm(TB, hid, hid, bf["h1"], hid, 0, w16, hid, 1, GELU, bf["pre2"], bf["h2"], hid, self.view("b2"),
                   b_off=o["W2"])
        self._gemm(TB, 3 * H, hid, bf["h2"], hid, 0, w16, hid, 1, STORE32, bf["gx"], None, 3 * H, self.view("bih"),
                   b_off=o["Wih"])
        bf["S"][0].copy_(h0)
        bf["S16"][0].copy_(h0.to(torch.bfloat16))
        for t in range(T):
            self._gemm(B, 3 * H, H, bf["S16"], H, 0, w16, H, 1, STORE32, bf["gh"], None, 3 * H, self.view("bhh"),
                       a_off=t * B * H, b_off=o["Whh"])
            gates = _p(bf["gates"]) + t * B * 4 * H * 4 if train else None
            _lib.check(self.lib.fz_sg_gru_fwd(B, H, bf["gx"].data_ptr() + t * B * 3 * H * 4, _p(bf["gh"]),
                                              bf["S"].data_ptr() + t * B * H * 4,
                                              bf["S"].data_ptr() + (t + 1) * B * H * 4,
                                              bf["S16"].data_ptr() + (t + 1) * B * H * 2, gates, self._stream()),
                       "fz_sg_gru_fwd")
        ld = hid + H
        self._gemm(TB, N_OUT_PAD, hid, bf["h2"], hid, 0, w16, ld, 1, STORE32, bf["y"], None, N_OUT_PAD,
                   self.view("bo"), b_off=o["Wo"])
        self._gemm(TB, N_OUT_PAD, H, bf["S16"], H, 0, w16, ld, 1, ACC32, bf["y"], None, N_OUT_PAD,
                   a_off=B * H, b_off=o["Wo"] + hid)
        return bf["y"].view(T, B, N_OUT_PAD)[..., :self.n_out], bf["S"][T]

    def loss_backward(self, X16: torch.Tensor, Y: torch.Tensor):
        """After ``forward(X16, h0, train=True)``: MSE against Y [T, B, n_out] (scaled rates),
        gradients into ``self.grad``.  Returns the loss (device scalar)."""
        T, B, _ = X16.shape
        bf = self._buffers(T, B, True)
        TB, hid, H, n_in = T * B, self.hid, self.H, self.n_in
        w16 = self.w16
        o = {k: v[0] for k, v in self.off.items()}
        tgt = torch.zeros((TB, N_OUT_PAD), dtype=torch.float32, device=self.dev)
        tgt[:, :self.n_out] = Y.reshape(TB, self.n_out)
        self.loss.zero_()
        L = self.lib
        s = self._stream()
        _lib.check(L.fz_sg_mse(TB * N_OUT_PAD, _p(bf["y"]), _p(tgt), _p(bf["dy"]), 1.0 / (TB * self.n_out),
                               _p(self.loss), s), "fz_sg_mse")
        _lib.check(L.fz_sg_cast(TB * N_OUT_PAD, _p(bf["dy"]), _p(bf["dy16"]), s), "fz_sg_cast")
        g = self.grad
        ld = hid + H
        # output layer
        self._gemm(N_OUT_PAD, hid, TB, bf["dy16"], N_OUT_PAD, 1, bf["h2"], hid, 0, STORE32, g, None, ld,
                   c_off=o["Wo"])
        self._gemm(N_OUT_PAD, H, TB, bf["dy16"], N_OUT_PAD, 1, bf["S16"], H, 0, STORE32, g, None, ld,
                   b_off=B * H, c_off=o["Wo"] + hid)
        _lib.check(L.fz_sg_colsum(TB, N_OUT_PAD, _p(bf["dy"]), _p(self.gview("bo")), 0, s), "fz_sg_colsum")
        self._gemm(TB, hid, N_OUT_PAD, bf["dy16"], N_OUT_PAD, 0, w16, ld, 0, STORE32, bf["dh2"], None, hid,
                   b_off=o["Wo"])
        self._gemm(TB, H, N_OUT_PAD, bf["dy16"], N_OUT_PAD, 0, w16, ld, 0, STORE32, bf["dS"], None, H,
                   b_off=o["Wo"] + hid)
        # GRU, back through time
        bf["dh0"].zero_()
        for t in range(T - 1, -1, -1):
            dprev = bf["dS"].data_ptr() + (t - 1) * B * H * 4 if t > 0 else _p(bf["dh0"])
            _lib.check(L.fz_sg_gru_bwd(B, H, bf["dS"].data_ptr() + t * B * H * 4, bf["S"].data_ptr() + t * B * H * 4,
                                       bf["gates"].data_ptr() + t * B * 4 * H * 4,
                                       bf["dgx"].data_ptr() + t * B * 3 * H * 4, bf["dgh"].data_ptr() + t * B * 3 * H * 4,
                                       bf["dgx16"].data_ptr() + t * B * 3 * H * 2,
                                       bf["dgh16"].data_ptr() + t * B * 3 * H * 2, dprev, s), "fz_sg_gru_bwd")
            _lib.check(L.fz_sg_gemm(B, H, 3 * H, bf["dgh16"].data_ptr() + t * B * 3 * H * 2, 3 * H, 0,
                                    w16.data_ptr() + o["Whh"] * 2, H, 0, ACC32, dprev, None, H, None, s),
                       "fz_sg_gemm")
        self._gemm(3 * H, H, TB, bf["dgh16"], 3 * H, 1, bf["S16"], H, 0, STORE32, g, None, H, c_off=o["Whh"])
        _lib.check(L.fz_sg_colsum(TB, 3 * H, _p(bf["dgh"]), _p(self.gview("bhh")), 0, s), "fz_sg_colsum")
        self._gemm(3 * H, hid, TB, bf["dgx16"], 3 * H, 1, bf["h2"], hid, 0, STORE32, g, None, hid, c_off=o["Wih"])
        _lib.check(L.fz_sg_colsum(TB, 3 * H, _p(bf["dgx"]), _p(self.gview("bih")), 0, s), "fz_sg_colsum")
        self._gemm(TB, hid, 3 * H, bf["dgx16"], 3 * H, 0, w16, hid, 0, ACC32, bf["dh2"], None, hid, b_off=o["Wih"])
        # MLP
        _lib.check(L.fz_sg_gelu_bwd(TB * hid, _p(bf["dh2"]), _p(bf["pre2"]), _p(bf["dp2"]), _p(bf["dp2_16"]), s),
                   "fz_sg_gelu_bwd")
        self._gemm(hid, hid, TB, bf["dp2_16"], hid, 1, bf["h1"], hid, 0, STORE32, g, None, hid, c_off=o["W2"])
        _lib.check(L.fz_sg_colsum(TB, hid, _p(bf["dp2"]), _p(self.gview("b2")), 0, s), "fz_sg_colsum")
        self._gemm(TB, hid, hid, bf["dp2_16"], hid, 0, w16, hid, 0, STORE32, bf["dh1"], None, hid, b_off=o["W2"])
        _lib.check(L.fz_sg_gelu_bwd(TB * hid, _p(bf["dh1"]), _p(bf["pre1"]), _p(bf["dp1"]), _p(bf["dp1_16"]), s),
                   "fz_sg_gelu_bwd")
        self._gemm(hid, n_in, TB, bf["dp1_16"], hid, 1, X16.reshape(TB, n_in), n_in, 0, STORE32, g, None, n_in,
                   c_off=o["W1"])
        _lib.check(L.fz_sg_colsum(TB, hid, _p(bf["dp1"]), _p(self.gview("b1")), 0, s), "fz_sg_colsum")
        return self.loss

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step_count += 1
        c1, c2 = 1 - b1 ** self.step_count, 1 - b2 ** self.step_count
        _lib.check(self.lib.fz_sg_adam(self.size, _p(self.w), _p(self.grad), _p(self.m), _p(self.v), _p(self.w16),
                                       lr, b1, b2, eps, c1, c2, wd, self._stream()), "fz_sg_adam")


class Ensemble:
    """Independently initialised members; the spread of their predictions flags frames for the
    full simulation.

    ``predict`` runs every member in the same launches: the members' weights are stacked
    contiguously (``pack``; call it again after training a member further), each layer is one
    member-batched GEMM (the input shared, stride 0), the GRU one launch over all members' rows,
    and one kernel computes every member's output layer, the ensemble mean and the disagreement.
    ``predict_members`` is the same computation member by member."""

    def __init__(self, members: list[RateSurrogate]):
        self.members = members
        self.state = None
        self._W16 = self._W32 = None
        m0 = members[0]
        for m in members:
            if (m.n_in, m.hid, m.H, m.n_out, m.rate_scale) != (m0.n_in, m0.hid, m0.H, m0.n_out, m0.rate_scale):
                raise ValueError("ensemble members must share their shape and rate scale")

    def pack(self):
        self._W16 = torch.stack([m.w16 for m in self.members]).contiguous()
        self._W32 = torch.stack([m.w for m in self.members]).contiguous()

    def reset(self, B):
        m0 = self.members[0]
        E, H, hid, d = len(self.members), m0.H, m0.hid, m0.dev
        self.state = [torch.zeros((B, H), dtype=torch.float32, device=m.dev) for m in self.members]
        self._S = torch.zeros((2, E, B, H), dtype=torch.float32, device=d)
        self._S16 = torch.zeros((2, E, B, H), dtype=torch.bfloat16, device=d)
        self._cur = 0
        self._h1 = torch.empty((E, B, hid), dtype=torch.bfloat16, device=d)
        self._h2 = torch.empty((E, B, hid), dtype=torch.bfloat16, device=d)
        self._gx = torch.empty((E, B, 3 * H), dtype=torch.float32, device=d)
        self._gh = torch.empty((E, B, 3 * H), dtype=torch.float32, device=d)
        self._mean = torch.empty((B, m0.n_out), dtype=torch.float32, device=d)
        self._spread = torch.empty(B, dtype=torch.float32, device=d)

    def predict(self, X16: torch.Tensor):
        """One frame for B sequences: X16 [B, n_in] bf16.  Returns (mean rates [B, n_out] in Hz,
        relative disagreement [B]: the members' standard deviation over the rate vector relative
        to the mean rate vector, ||std|| / (||mean|| + 1 Hz)); views into buffers that the next
        call overwrites."""
        if self._W16 is None:
            self.pack()
        m0 = self.members[0]
        E, (B, n_in), hid, H = len(self.members), X16.shape, m0.hid, m0.H
        if n_in != m0.n_in or self._mean.shape[0] != B:
            raise ValueError("predict: input does not match the ensemble or the batch of reset()")
        X16 = X16.contiguous()
        lib, st, sz = m0.lib, m0._stream(), m0.size
        o = {k: v[0] for k, v in m0.off.items()}
        W16, W32 = self._W16.data_ptr(), self._W32.data_ptr()
        w = lambda name: W16 + 2 * o[name]          # noqa: E731  (member 0's block; members at + sz)
        b = lambda name: W32 + 4 * o[name]          # noqa: E731
        gb = lambda *a: _lib.check(lib.fz_sg_gemm_batched(*a, st), "fz_sg_gemm_batched")  # noqa: E731
        cur, nxt = self._cur, 1 - self._cur
        gb(E, B, hid, n_in, _p(X16), n_in, 0, 0, w("W1"), n_in, sz, 1, GELU16, None, _p(self._h1), hid, B * hid,
           b("b1"), sz)
        gb(E, B, hid, hid, _p(self._h1), hid, B * hid, 0, w("W2"), hid, sz, 1, GELU16, None, _p(self._h2), hid,
           B * hid, b("b2"), sz)
        gb(E, B, 3 * H, hid, _p(self._h2), hid, B * hid, 0, w("Wih"), hid, sz, 1, STORE32, _p(self._gx), None,
           3 * H, B * 3 * H, b("bih"), sz)
        gb(E, B, 3 * H, H, _p(self._S16[cur]), H, B * H, 0, w("Whh"), H, sz, 1, STORE32, _p(self._gh), None,
           3 * H, B * 3 * H, b("bhh"), sz)
        _lib.check(lib.fz_sg_gru_fwd(E * B, H, _p(self._gx), _p(self._gh), _p(self._S[cur]), _p(self._S[nxt]),
                                     _p(self._S16[nxt]), None, st), "fz_sg_gru_fwd")
        self._cur = nxt
        _lib.check(lib.fz_sg_ens_head(B, E, hid, H, m0.n_out, _p(self._h2), _p(self._S16[nxt]), W16, W32, sz,
                                      o["Wo"], o["bo"], m0.rate_scale, _p(self._mean), _p(self._spread), st),
                   "fz_sg_ens_head")
        return self._mean, self._spread

    def predict_members(self, X16: torch.Tensor):
        """``predict`` member by member through each member's own forward (for comparison)."""
        ys = []
        for k, m in enumerate(self.members):
            y, s = m.forward(X16.unsqueeze(0), self.state[k])
            self.state[k].copy_(s)
            ys.append(y[0] * m.rate_scale)
        Y = torch.stack(ys)
        mean = Y.mean(0)
        std = Y.std(0, unbiased=False) if len(ys) > 1 else torch.zeros_like(mean)
        spread = std.norm(dim=1) / (mean.norm(dim=1) + 1.0)
        return mean, spread
