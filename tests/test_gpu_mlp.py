"""The surrogate network's kernels on the GPU against numpy: the BF16 MFMA GEMM in all four transpose
forms and every epilogue, the loss, Adam, and the full backward pass against the FP64 oracle
(oracle/mlp_oracle.py) on the same BF16-valued weights and inputs."""
import sys

import numpy as np
import pytest

from chaos_amd.surrogate import to_bf16_bits, from_bf16_bits
from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def dev_bf16(a):
    import torch
    return torch.from_numpy(to_bf16_bits(a).view(np.int16)).cuda()


def host_bf16(t):
    return from_bf16_bits(t.cpu().numpy().view(np.uint16)).astype(np.float64)


@pytest.mark.parametrize("tA,tB", [(0, 0), (1, 0), (0, 1), (1, 1)])
# (256, 768, 4096), (2, 256, 4096) and (70, 100, 1000) are split over K (the last with a partial final split)
@pytest.mark.parametrize("shape", [(64, 64, 32), (70, 100, 50), (2, 256, 256), (256, 4096, 768), (768, 300, 96),
                                   (256, 768, 4096), (2, 256, 4096), (70, 100, 1000)])
def test_gemm_against_numpy(tA, tB, shape):
    import torch
    from chaos_amd.surrogate import SurrogateMLP
    M, N, K = shape
    say("gemm", tA, tB, shape)
    net = SurrogateMLP(sizes=(4, 2))  # only for its context and gemm helper
    rng = np.random.default_rng(M * 7 + N * 3 + K + tA * 11 + tB * 13)
    Am = rng.standard_normal((M, K)).astype(np.float32)
    Bm = rng.standard_normal((K, N)).astype(np.float32)
    As = Am.T.copy() if tA else Am
    Bs = Bm.T.copy() if tB else Bm
    Ad, Bd = dev_bf16(As), dev_bf16(Bs)
    Aq = from_bf16_bits(to_bf16_bits(Am)).astype(np.float64)
    Bq = from_bf16_bits(to_bf16_bits(Bm)).astype(np.float64)
    ref = Aq @ Bq
    bias = rng.standard_normal(M).astype(np.float32)
    bd = torch.from_numpy(bias).cuda()
    Hm = rng.standard_normal((M, N)).astype(np.float32)
    Hd = dev_bf16(Hm)
    Hq = from_bf16_bits(to_bf16_bits(Hm))
    for epi in (0, 1, 2, 3):
        want = ref.copy()
        if epi in (1, 2):
            want = want + bias[:, None]
        if epi == 1:
            want = np.maximum(want, 0)
        if epi == 3:
            want = np.where(Hq > 0, want, 0)
        C = torch.empty((M, N), dtype=torch.float32, device="cuda")
        net.gemm(tA, tB, M, N, K, Ad.data_ptr(), As.shape[1], Bd.data_ptr(), Bs.shape[1], C.data_ptr(), N, True,
                 bd.data_ptr() if epi in (1, 2) else None, epi, Hd.data_ptr() if epi == 3 else None, N)
        got = C.cpu().numpy().astype(np.float64)
        tol = 1e-5 * np.sqrt(K) * (np.abs(Aq) @ np.abs(Bq) + 1)
        assert np.all(np.abs(got - want) <= tol), (epi, np.max(np.abs(got - want) / tol))
        Cb = torch.empty((M, N), dtype=torch.int16, device="cuda")
        net.gemm(tA, tB, M, N, K, Ad.data_ptr(), As.shape[1], Bd.data_ptr(), Bs.shape[1], Cb.data_ptr(), N, False,
                 bd.data_ptr() if epi in (1, 2) else None, epi, Hd.data_ptr() if epi == 3 else None, N)
        gb = host_bf16(Cb)
        assert np.all(np.abs(gb - got) <= np.abs(got) * 2.0 ** -8 + 1e-30)


def test_loss_and_adam_against_numpy():
    import torch
    from chaos_amd.surrogate import SurrogateMLP
    say("loss, adam")
    net = SurrogateMLP(sizes=(4, 2))
    rng = np.random.default_rng(1)
    B = 1000
    out = rng.standard_normal((2, B)).astype(np.float32) * 3
    y = (rng.random(B) > 0.4).astype(np.float32)
    t = rng.standard_normal(B).astype(np.float32)
    w = rng.random(B).astype(np.float32)
    dout, part = net.loss(torch.from_numpy(out).cuda(), torch.from_numpy(y).cuda(), torch.from_numpy(t).cuda(),
                          torch.from_numpy(w).cuda(), 0.3)
    lb, lr, d = MO.loss_and_dout(out.astype(np.float64), y, t, w, 0.3)
    p = part.cpu().numpy().reshape(-1, 2).sum(0) / B
    assert abs(p[0] - lb) < 1e-5 * max(1, lb) and abs(p[1] - lr) < 1e-5 * max(1, lr)
    gd = host_bf16(dout)
    assert np.all(np.abs(gd - d) <= np.abs(d) * 2.0 ** -8 + 1e-12)
    # Adam against numpy, three steps
    n = 5000
    p0 = rng.standard_normal(n).astype(np.float32)
    pd = torch.from_numpy(p0.copy()).cuda()
    m = torch.zeros(n, device="cuda")
    v = torch.zeros(n, device="cuda")
    pb = torch.empty(n, dtype=torch.int16, device="cuda")
    pn, mn, vn = p0.astype(np.float64), np.zeros(n), np.zeros(n)
    for s in range(1, 4):
        g = rng.standard_normal(n).astype(np.float32)
        rc = net.lib.cl_mlp_adam(net.ctx, n, pd.data_ptr(), pb.data_ptr(), torch.from_numpy(g).cuda().data_ptr(),
                                 m.data_ptr(), v.data_ptr(), s, 1e-2, 0.9, 0.999, 1e-8, 0.0, None)
        assert rc == 0
        mn = 0.9 * mn + 0.1 * g
        vn = 0.999 * vn + 0.001 * g * g
        pn = pn - 1e-2 * (mn / (1 - 0.9 ** s)) / (np.sqrt(vn / (1 - 0.999 ** s)) + 1e-8)
    torch.cuda.synchronize()
    assert np.max(np.abs(pd.cpu().numpy() - pn)) < 1e-5
    assert np.array_equal(pb.cpu().numpy().view(np.uint16), to_bf16_bits(pd.cpu().numpy()))


def test_backward_against_fp64_oracle():
    """Full network gradients on the GPU against the FP64 oracle on the same BF16 weights and inputs. The GPU
    stores activations and back-propagated errors in BF16, so its error against exact arithmetic is bounded by
    that of the same network with BF16 rounding at those points (the oracle run with a BF16 rounding), which
    grows with depth through cancellation in the first layer's gradient (about 5% there, 0.02% at the last)."""
    import torch
    from chaos_amd.surrogate import SurrogateMLP
    say("backward vs oracle")
    net = SurrogateMLP(seed=3)
    rng = np.random.default_rng(2)
    B = 2048
    X = rng.standard_normal((768, B)).astype(np.float32)
    y = (rng.random(B) > 0.5).astype(np.float32)
    t = rng.standard_normal(B).astype(np.float32)
    w = np.ones(B, dtype=np.float32)
    Xd = dev_bf16(X)
    acts, out = net.forward(Xd, B)
    dout, _ = net.loss(out, torch.from_numpy(y).cuda(), torch.from_numpy(t).cuda(), torch.from_numpy(w).cuda(), 0.5)
    gW, gb = net.backward(Xd, B, acts, dout)
    Ws = [host_bf16(W) for W in net.Wb]
    bs = [b.cpu().numpy().astype(np.float64) for b in net.b]
    Xq = from_bf16_bits(to_bf16_bits(X)).astype(np.float64)
    ref_acts = MO.forward(Ws, bs, Xq)
    o = out.cpu().numpy()
    assert np.max(np.abs(o - ref_acts[-1])) < 0.05 * (np.max(np.abs(ref_acts[-1])) + 1)
    _, _, d = MO.loss_and_dout(ref_acts[-1], y, t, w, 0.5)
    rW, rb = MO.backward(Ws, ref_acts, d)

    def q(a):
        return from_bf16_bits(to_bf16_bits(a)).astype(np.float64)
    em_acts = MO.forward(Ws, bs, Xq, rnd=q)
    _, _, de = MO.loss_and_dout(em_acts[-1].astype(np.float32).astype(np.float64), y, t, w, 0.5)
    eW, eb = MO.backward(Ws, em_acts, de, rnd=q)
    for l in range(4):
        n = np.linalg.norm(rW[l])
        budget = np.linalg.norm(eW[l] - rW[l]) / n
        rel = np.linalg.norm(gW[l].cpu().numpy() - rW[l]) / n
        assert rel < 1.5 * budget + 5e-3, (l, rel, budget)
        nb = np.linalg.norm(rb[l])
        budget_b = np.linalg.norm(eb[l] - rb[l]) / nb
        relb = np.linalg.norm(gb[l].cpu().numpy() - rb[l]) / nb
        assert relb < 1.5 * budget_b + 5e-3, (l, relb, budget_b)


def test_training_reduces_loss_on_separable_data():
    import torch
    from chaos_amd.surrogate import SurrogateMLP
    say("training")
    rng = np.random.default_rng(4)
    n = 32768
    X = rng.standard_normal((768, n)).astype(np.float32)
    wtrue = rng.standard_normal(768)
    s = wtrue @ X
    y = (s > 0).astype(np.float32)
    t = (s / 30).astype(np.float32)
    net = SurrogateMLP(seed=5)
    hist = net.train(dev_bf16(X), torch.from_numpy(y).cuda(), torch.from_numpy(t).cuda(),
                     torch.ones(n, device="cuda"), epochs=3, batch=2048, lr=1e-3)
    assert hist[-1][0] < 0.5 * hist[0][0]


def test_native_step_equals_one_call_per_kernel():
    """cl_mlp_train_step leaves every parameter, BF16 copy and Adam moment bit for bit where the separate calls do,
    on batches taken as column ranges of a wider matrix (row stride > batch), and its loss sums match theirs."""
    import torch
    from chaos_amd.surrogate import SurrogateMLP
    say("native training step")
    rng = np.random.default_rng(11)
    n = 3 * 2048 + 512
    X = rng.standard_normal((768, n)).astype(np.float32)
    s = rng.standard_normal(768) @ X
    y = torch.from_numpy((s > 0).astype(np.float32)).cuda()
    t = torch.from_numpy((s / 30).astype(np.float32)).cuda()
    w = torch.from_numpy(rng.random(n).astype(np.float32) + 0.5).cuda()
    Xd = dev_bf16(X)
    nets = [SurrogateMLP(seed=9), SurrogateMLP(seed=9)]
    hist = [net.train(Xd, y, t, w, epochs=2, batch=2048, lr=2e-3, native=nat) for net, nat in zip(nets, (True, False))]
    a, b = nets
    for l in range(a.nlayers):
        for nm in ("W", "Wb", "b", "mW", "vW", "mb", "vb"):
            x, z = getattr(a, nm)[l].cpu().numpy(), getattr(b, nm)[l].cpu().numpy()
            assert x.tobytes() == z.tobytes(), (nm, l)
    assert a.t == b.t == 6
    np.testing.assert_allclose(np.array(hist[0]), np.array(hist[1]), rtol=1e-6)
