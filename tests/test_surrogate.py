"""Learned rate surrogate: the FP64 oracle is pinned by finite differences; the GPU GEMM by
numpy; the GPU forward/backward (BF16 MFMA, FP32 accumulation) by the oracle with the same
BF16 rounding points."""

import numpy as np
import pytest

from oracle.surrogate_oracle import bf16, forward_backward


def _params(n_in, hid, H, n_out, pad, rng, scale=0.3):
    P = {"W1": rng.normal(0, scale, (hid, n_in)), "b1": rng.normal(0, 0.1, hid),
         "W2": rng.normal(0, scale, (hid, hid)), "b2": rng.normal(0, 0.1, hid),
         "Wih": rng.normal(0, scale, (3 * H, hid)), "bih": rng.normal(0, 0.1, 3 * H),
         "Whh": rng.normal(0, scale, (3 * H, H)), "bhh": rng.normal(0, 0.1, 3 * H),
         "Wo": np.zeros((pad, hid + H)), "bo": np.zeros(pad)}
    P["Wo"][:n_out] = rng.normal(0, scale, (n_out, hid + H))
    P["bo"][:n_out] = rng.normal(0, 0.1, n_out)
    return P


def test_oracle_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, hid, H, n_out, T, B = 6, 5, 4, 3, 3, 2
    P = _params(n_in, hid, H, n_out, n_out, rng)
    X = rng.normal(0, 1, (T, B, n_in))
    h0 = rng.normal(0, 0.5, (B, H))
    Y = rng.normal(0, 1, (T, B, n_out))
    loss, g, _ = forward_backward(P, X, h0, Y, n_out)
    eps = 1e-6
    for name in P:
        flat = P[name].ravel()
        for idx in rng.choice(flat.size, min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp = forward_backward(P, X, h0, Y, n_out)[0]
            flat[idx] = old - eps
            lm = forward_backward(P, X, h0, Y, n_out)[0]
            flat[idx] = old
            num = (lp - lm) / (2 * eps)
            assert abs(num - g[name].ravel()[idx]) < 1e-6 + 1e-5 * abs(num), (name, idx, num, g[name].ravel()[idx])


def test_bf16_rounding_matches_torch():
    import torch

    x = np.random.default_rng(1).normal(0, 3, 10000).astype(np.float32)
    x[:4] = [1.00390625, 1.01171875, -2.0, 0.0]          # ties round to even
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(bf16(x).astype(np.float32), want)


# ------------------------------------------------------------------------------------- GPU


def _bf(t):
    import torch

    return torch.from_numpy(np.asarray(t, np.float32)).cuda().to(torch.bfloat16)


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(70, 45, 96), (200, 136, 96), (256, 384, 64), (136, 130, 32)])
def test_mfma_gemm_matches_numpy(ta, tb, M, N, K):
    # 64 x 64 tiles for small or unaligned operands, 128 x 128 tiles from M, N >= 128 with
    # leading dimensions a multiple of 8 (N = 130 is neither for the [K][N] layouts)
    import torch

    from flyhip import _lib

    lib = _lib.load()
    rng = np.random.default_rng(2)
    A = bf16(rng.normal(0, 1, (M, K)))
    Bm = bf16(rng.normal(0, 1, (K, N)))
    bias = rng.normal(0, 1, N).astype(np.float32)
    As = _bf(A.T.copy() if ta else A)
    Bs = _bf(Bm.T.copy() if tb else Bm)
    lda = M if ta else K
    ldb = K if tb else N
    want = A @ Bm + bias
    c32 = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    c16 = torch.zeros((M, N), dtype=torch.bfloat16, device="cuda")
    bt = torch.from_numpy(bias).cuda()
    s = _lib.stream_handle()
    _lib.check(lib.fz_sg_gemm(M, N, K, As.data_ptr(), lda, ta, Bs.data_ptr(), ldb, tb, 0, c32.data_ptr(), None, N,
                              bt.data_ptr(), s), "gemm")
    np.testing.assert_allclose(c32.cpu().numpy(), want, rtol=1e-5, atol=1e-4)
    _lib.check(lib.fz_sg_gemm(M, N, K, As.data_ptr(), lda, ta, Bs.data_ptr(), ldb, tb, 1, c32.data_ptr(), None, N,
                              None, s), "gemm acc")
    np.testing.assert_allclose(c32.cpu().numpy(), want + (want - bias), rtol=1e-5, atol=2e-4)
    _lib.check(lib.fz_sg_gemm(M, N, K, As.data_ptr(), lda, ta, Bs.data_ptr(), ldb, tb, 3, c32.data_ptr(),
                              c16.data_ptr(), N, bt.data_ptr(), s), "gemm gelu")
    from oracle.surrogate_oracle import gelu

    np.testing.assert_allclose(c16.float().cpu().numpy(), bf16(gelu(c32.cpu().numpy())), rtol=1e-2, atol=1e-2)


@pytest.mark.gpu
def test_forward_backward_matches_oracle_with_bf16_points():
    import torch

    from flyhip.surrogate import N_OUT_PAD, RateSurrogate

    n_in, hid, H, n_out, T, B = 64, 64, 32, 8, 4, 8
    m = RateSurrogate(n_in, hid, H, n_out, seed=3)
    P = m.params()
    rng = np.random.default_rng(4)
    X = bf16(rng.normal(0, 1, (T, B, n_in)))
    h0 = rng.normal(0, 0.3, (B, H)).astype(np.float32)
    Y = rng.normal(0, 1, (T, B, n_out)).astype(np.float32)
    y, _ = m.forward(_bf(X), torch.from_numpy(h0).cuda(), train=True)
    loss = float(m.loss_backward(_bf(X), torch.from_numpy(Y).cuda()).item())
    Yp = np.zeros((T, B, N_OUT_PAD))
    Yp[..., :n_out] = Y
    want_loss, g, want_y = forward_backward(P, X, h0.astype(np.float64), Yp, n_out, bf16_ops=True)
    np.testing.assert_allclose(y.cpu().numpy(), want_y[..., :n_out], rtol=1e-3, atol=1e-3)
    assert abs(loss - want_loss) < 1e-4 * abs(want_loss) + 1e-6
    G = m.grads()
    for name in P:
        a, b = G[name], g[name]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel < 2e-3, (name, rel)


@pytest.mark.gpu
@pytest.mark.parametrize("hid,H,B", [(64, 32, 40), (1024, 256, 4096)])
def test_member_batched_inference_equals_member_by_member(hid, H, B):
    # Ensemble.predict (member-batched GEMMs, one GRU launch, the fused output layer and
    # statistics) against each member's own forward, over 4 frames of a recurrent state
    import torch

    from flyhip.surrogate import Ensemble, RateSurrogate

    n_in = 2048 if hid == 1024 else 64
    members = [RateSurrogate(n_in, hid, H, 8, seed=40 + k, rate_scale=20.0) for k in range(3)]
    fast, slow = Ensemble(members), Ensemble(members)
    fast.reset(B)
    slow.reset(B)
    rng = np.random.default_rng(5)
    for f in range(4):
        x = _bf(bf16(rng.random((B, n_in)) * 2.0))
        m1, s1 = fast.predict(x)
        m2, s2 = slow.predict_members(x)
        np.testing.assert_allclose(m1.cpu().numpy(), m2.cpu().numpy(), rtol=1e-4, atol=1e-3)
        np.testing.assert_allclose(s1.cpu().numpy(), s2.cpu().numpy(), rtol=1e-3, atol=1e-5)
        assert float(s1.max()) > 0


@pytest.mark.gpu
def test_training_reduces_loss_on_a_recurrent_target():
    import torch

    from flyhip.surrogate import RateSurrogate

    n_in, T, B = 64, 8, 32
    m = RateSurrogate(n_in, 128, 64, 8, seed=5)
    rng = np.random.default_rng(6)
    Wt = rng.normal(0, 0.3, (n_in, 8))
    losses = []
    for it in range(400):
        X = rng.normal(0, 1, (T, B, n_in))
        lin = X @ Wt
        Y = np.tanh(lin + 0.5 * np.concatenate([np.zeros((1, B, 8)), lin[:-1]], 0))   # depends on the past frame
        m.forward(_bf(X), torch.zeros((B, 64), device="cuda"), train=True)
        losses.append(float(m.loss_backward(_bf(X), torch.from_numpy(Y.astype(np.float32)).cuda()).item()))
        m.adam(lr=2e-3)
    assert np.mean(losses[-20:]) < 0.15 * np.mean(losses[:10]), (losses[:3], losses[-3:])


def test_grid_pool_is_the_cell_mean():
    import torch

    from flyhip.surrogate_pipeline import GridPool

    rng = np.random.default_rng(7)
    xy = rng.random((500, 2))
    rates = rng.random((3, 500)).astype(np.float32)
    got = GridPool(xy, nx=8, ny=4, device="cpu")(torch.from_numpy(rates)).numpy()
    cell = np.minimum((xy[:, 1] * 4).astype(int), 3) * 8 + np.minimum((xy[:, 0] * 8).astype(int), 7)
    for c in range(32):
        m = cell == c
        want = rates[:, m].mean(1) if m.any() else np.zeros(3)
        np.testing.assert_allclose(got[:, c], want, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_surrogate_pipeline_end_to_end_small():
    import torch

    from flyhip.connectome import flywire_like
    from flyhip.surrogate import Ensemble
    from flyhip.surrogate_pipeline import (FallbackSim, generate, generate_judged, judge, surrogate_loop, timed_loop,
                                           train_ensemble)

    fly = flywire_like(n=4000, n_edges=4000 * 60, seed=11)
    X, Y, drive = generate(fly, frames=40, batch=16, seed=1, skip=2)
    assert X.shape == (40, 16, 2048) and Y.shape == (40, 16, 8) and drive.S == 16
    assert np.isfinite(Y.cpu().numpy()).all() and float(Y.sum()) > 0
    members, curves = train_ensemble(X, Y, members=2, T=8, seqs=16, steps=60, lr=2e-3, rate_scale=20.0)
    assert curves[0][-1][1] < curves[0][0][1]
    Xt, ref, trial, others, drives = generate_judged(fly, frames=20, sequences=6, replicas=4, batch=16, skip=2)
    assert Xt.shape == (20, 6, 2048) and ref.shape == (20, 6, 8)
    ens = Ensemble(members)
    fb = FallbackSim(fly, None, warm=2, sizes=(4, 16))
    loops, nets, flags = [], [], []
    for start, d in drives:
        S = min(d.S, 6 - start)
        fb.drive = d
        lo, ne, fl = surrogate_loop(ens, Xt[:, start:start + S].contiguous(), 0.0, fb)   # every frame falls back
        loops.append(lo)
        nets.append(ne)
        flags.append(fl)
    loop, net, flag = torch.cat(loops, 1), torch.cat(nets, 1), torch.cat(flags, 1)
    res = judge(net, ref, trial, others, 4, loop=loop, flags=flag, const=Y.mean((0, 1)))
    for k in ("err_surrogate", "err_with_fallback", "fallback_frac", "err_single_sim_trial", "err_constant_predictor"):
        assert np.isfinite(res[k]), (k, res)
    assert res["fallback_frac"] == 1.0 and not torch.equal(loop, net)
    tl = timed_loop(fly, ens, 0.15, 2, sequences=32, frames=3, fallback=fb)
    assert tl["frames_per_s"] > 0 and 0.0 <= tl["fallback_frac"] <= 1.0


@pytest.mark.gpu
def test_fallback_is_a_brain_at_rest_warmed_on_the_sequence():
    # FallbackSim.run(seqs, f): each flagged sequence's DN group rates of frame f from a brain at
    # rest that ran the sequence's frames f - warm .. f; restated here with a brain of its own
    import torch

    from flyhip.batch import DNReadout
    from flyhip.connectome import flywire_like
    from flyhip.surrogate_pipeline import FRAME_MS, FallbackSim, VisualDrive, _Driven

    fly = flywire_like(n=4000, n_edges=4000 * 60, seed=11)
    drive = VisualDrive(fly.input_xy, 10, seed=3)
    fb = FallbackSim(fly, drive, warm=3, sizes=(4, 8), seed=900)
    seqs = torch.tensor([7, 2, 5], device="cuda")
    got = fb.run(seqs, 6)                                   # 3 flagged: the 4-fly brain (seed 900)
    ref = _Driven(fly, 4, 900, 7.8, tune=False)
    rd = DNReadout(fly, 4, tau_ms=1e-9)
    rows = torch.arange(3, device="cuda")
    ref.bb.alive[3:] = 0
    for f in range(3, 7):
        ref.set_rates(drive.rates(f, seqs), drive.drive_hz[seqs], rows)
        counts = ref.bb.run(FRAME_MS, copy=False)
    want = rd.update(counts, FRAME_MS)[:3]
    np.testing.assert_array_equal(got.cpu().numpy(), want.cpu().numpy())
    assert float(want.sum()) > 0
    many = fb.run(torch.arange(10, device="cuda").repeat(2), 6)    # 20 > 8: chunks on the large brain
    assert many.shape == (20, 8) and np.isfinite(many.cpu().numpy()).all()


@pytest.mark.gpu
def test_trained_surrogate_beats_the_method_without_it():
    # what the network gains: on held-out sequences, its DN-rate error against the simulation's
    # expected rate (replica mean) is well below that of the best constant prediction (the
    # training mean rate), the method without the network
    from flyhip.connectome import flywire_like
    from flyhip.surrogate import Ensemble
    from flyhip.surrogate_pipeline import generate, generate_judged, judge, surrogate_loop, train_ensemble

    fly = flywire_like(n=8000, n_edges=8000 * 109, seed=11)
    X, Y, _ = generate(fly, frames=400, batch=64, seed=1, replicas=4, skip=5)
    Y = Y.view(400, 16, 4, 8).mean(2).contiguous()
    members, _ = train_ensemble(X, Y, members=2, T=16, seqs=64, steps=500, lr=1e-3, rate_scale=20.0)
    Xt, ref, trial, others, _ = generate_judged(fly, frames=60, sequences=8, replicas=8, batch=64, seed=9, skip=5)
    _, net, _ = surrogate_loop(Ensemble(members), Xt, 0.15, None)
    res = judge(net, ref, trial, others, 8, const=Y.mean((0, 1)))
    print(res)
    assert res["mean_rate_hz"] > 0
    assert res["err_surrogate"] < 0.5 * res["err_constant_predictor"], res
