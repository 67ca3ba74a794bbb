"""FP64 oracle of the learned rate surrogate's forward and backward passes -- TEST
INFRASTRUCTURE ONLY (the checker for flyhip.surrogate; the product never imports it).

Network (the same as can-the-fly-play-f-zero-_amd/csrc/surrogate.hip):
    h1 = gelu(x W1^T + b1),  h2 = gelu(h1 W2^T + b2)       (tanh-approximated GELU)
    s_t = GRU(h2_t, s_{t-1})                                 (gate order r, z, n)
    y_t = [h2_t, s_t] Wo^T + bo,  loss = sum (y - target)^2 / (T B n_out)

With ``bf16=True`` every tensor the GPU stores in BF16 before a GEMM (inputs, weights,
activations h1 h2 s, and the back-propagated dy, dgx, dgh, dpre) is rounded to BF16 at the
same point, so the only remaining difference from the GPU is FP32 versus FP64 accumulation.
The backward pass itself is checked against finite differences (tests/test_surrogate.py).
"""

from __future__ import annotations

import numpy as np

C = 0.7978845608028654


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * C * (1.0 + 3.0 * 0.044715 * x * x)


def sigm(x):
    return 1.0 / (1.0 + np.exp(-x))


def forward_backward(P, X, h0, Y, n_out, bf16_ops=False):
    """P: dict of float64 parameters (W1 b1 W2 b2 Wih bih Whh bhh Wo bo; Wo/bo padded rows are
    zero); X [T, B, n_in]; h0 [B, H]; Y [T, B, n_out_padded].  Returns (loss, grads, y)."""
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    T, B, _ = X.shape
    H = P["Whh"].shape[1]
    hid = P["W1"].shape[0]
    W = {k: (r(v) if k[0] == "W" else np.asarray(v, np.float64)) for k, v in P.items()}
    x = r(X).reshape(T * B, -1)
    pre1 = x @ W["W1"].T + W["b1"]
    h1 = r(gelu(pre1))
    pre2 = h1 @ W["W2"].T + W["b2"]
    h2 = r(gelu(pre2))
    gx = (h2 @ W["Wih"].T + W["bih"]).reshape(T, B, 3 * H)
    s = [np.asarray(h0, np.float64)]
    s16 = [r(h0)]
    gates = []
    for t in range(T):
        gh = s16[t] @ W["Whh"].T + W["bhh"]
        a = gx[t]
        rr = sigm(a[:, :H] + gh[:, :H])
        z = sigm(a[:, H:2 * H] + gh[:, H:2 * H])
        nn = np.tanh(a[:, 2 * H:] + rr * gh[:, 2 * H:])
        hn = (1 - z) * nn + z * s[t]
        gates.append((rr, z, nn, gh[:, 2 * H:]))
        s.append(hn)
        s16.append(r(hn))
    S = np.stack(s16[1:]).reshape(T * B, H)
    y = h2 @ W["Wo"][:, :hid].T + S @ W["Wo"][:, hid:].T + W["bo"]
    tgt = np.asarray(Y, np.float64).reshape(T * B, -1)
    scale = 1.0 / (T * B * n_out)
    e = y - tgt
    loss = float((e * e).sum() * scale)
    dy = r(2.0 * e * scale)
    g = {}
    g["Wo"] = np.concatenate([dy.T @ h2, dy.T @ S], 1)
    g["bo"] = dy.sum(0) if not bf16_ops else (2.0 * e * scale).sum(0)
    dh2 = dy @ W["Wo"][:, :hid]
    dS = (dy @ W["Wo"][:, hid:]).reshape(T, B, H)
    dgx = np.zeros((T, B, 3 * H))
    dgh = np.zeros((T, B, 3 * H))
    dgx16 = np.zeros_like(dgx)
    dgh16 = np.zeros_like(dgh)
    carry = np.zeros((B, H))
    for t in range(T - 1, -1, -1):
        rr, z, nn, ghn = gates[t]
        d = dS[t] + carry
        dn = d * (1 - z) * (1 - nn * nn)
        dz = d * (s[t] - nn) * z * (1 - z)
        dr = dn * ghn * rr * (1 - rr)
        dgx[t] = np.concatenate([dr, dz, dn], 1)
        dgh[t] = np.concatenate([dr, dz, dn * rr], 1)
        dgx16[t], dgh16[t] = r(dgx[t]), r(dgh[t])
        carry = d * z + dgh16[t] @ W["Whh"]
    DGX, DGH = dgx16.reshape(T * B, -1), dgh16.reshape(T * B, -1)
    Sprev = np.stack(s16[:-1]).reshape(T * B, H)
    g["Whh"] = DGH.T @ Sprev
    g["bhh"] = dgh.reshape(T * B, -1).sum(0)
    g["Wih"] = DGX.T @ h2
    g["bih"] = dgx.reshape(T * B, -1).sum(0)
    dh2 = dh2 + DGX @ W["Wih"]
    dp2 = dh2 * gelu_grad(pre2)
    dp2_16 = r(dp2)
    g["W2"] = dp2_16.T @ h1
    g["b2"] = dp2.sum(0)
    dh1 = dp2_16 @ W["W2"]
    dp1 = dh1 * gelu_grad(pre1)
    g["W1"] = r(dp1).T @ x
    g["b1"] = dp1.sum(0)
    return loss, g, y.reshape(T, B, -1)
