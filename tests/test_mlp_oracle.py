"""Pin the numpy network oracle with central differences (FP64, a tiny network)."""
import numpy as np

from oracle import mlp_oracle as M


def total_loss(Ws, bs, X, y, t, w, wreg):
    acts = M.forward(Ws, bs, X)
    lb, lr, _ = M.loss_and_dout(acts[-1], y, t, w, wreg)
    return lb + lr


def test_gradients_match_central_differences():
    rng = np.random.default_rng(0)
    sizes = (7, 5, 4, 2)
    Ws = [rng.standard_normal((sizes[i + 1], sizes[i])) for i in range(3)]
    bs = [rng.standard_normal(sizes[i + 1]) * 0.1 for i in range(3)]
    X = rng.standard_normal((7, 11))
    y = (rng.random(11) > 0.5).astype(float)
    t = rng.standard_normal(11)
    w = rng.random(11)
    acts = M.forward(Ws, bs, X)
    _, _, d = M.loss_and_dout(acts[-1], y, t, w, 0.7)
    gW, gb = M.backward(Ws, acts, d)
    h = 1e-6
    for l in range(3):
        for idx in [(0, 0), (1, 2), (sizes[l + 1] - 1, sizes[l] - 1)]:
            Wp = [W.copy() for W in Ws]
            Wm = [W.copy() for W in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            num = (total_loss(Wp, bs, X, y, t, w, 0.7) - total_loss(Wm, bs, X, y, t, w, 0.7)) / (2 * h)
            assert abs(num - gW[l][idx]) < 1e-7 * max(1.0, abs(num))
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][0] += h
        bm[l][0] -= h
        num = (total_loss(Ws, bp, X, y, t, w, 0.7) - total_loss(Ws, bm, X, y, t, w, 0.7)) / (2 * h)
        assert abs(num - gb[l][0]) < 1e-7 * max(1.0, abs(num))
