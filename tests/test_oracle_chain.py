"""The coupled standard-map chain in the oracle, pinned by what the mathematics fixes
(the reference's coupled-maps suite, test/coupled-maps/cm_ac07_12.jl, cm_ac13_16.jl, cm_ac17_19.jl)."""
import ctypes

import mpmath as mp
import numpy as np
import pytest

from chaos_amd import StandardMapChain, StandardMap, QRSpectrum, HouseholderQR
from chaos_amd.tangent import make_chain, decode_states, finals_of
from tests import oracle_lib as O

QR = make_chain((QRSpectrum(HouseholderQR()),))
NONE = make_chain(())


def lambdas(res, k):
    chain, allr = QR
    d = res["x"].shape[1]
    return finals_of(decode_states(res["states"], allr, d, k), allr, d, k)[0]["lambda"]


def origin_exponents(N, K, eps, ring):
    """Lyapunov spectrum of the origin: G(0) = K I + eps Lap (the ring or path Laplacian, eigenvalues
    2 - 2 cos(2 pi j / N) or 2 - 2 cos(pi j / N)); each mode is the block [[1+g, 1], [g, 1]] with eigenvalues
    mu, 1/mu, mu + 1/mu = 2 + g (cm_ac07_12.jl:294-340 gives the eigenvalues K + eps l_j)."""
    out = []
    for j in range(N):
        lj = 2 - 2 * np.cos((2 if ring else 1) * np.pi * j / N)
        g = K + eps * lj
        tr = 2 + g
        if abs(tr) > 2:
            mu = (abs(tr) + np.sqrt(tr * tr - 4)) / 2
            out += [np.log(mu), -np.log(mu)]
        else:
            out += [0.0, 0.0]
    return np.sort(np.array(out))[::-1]


@pytest.mark.parametrize("N,ring", [(8, True), (16, True), (8, False), (32, False)])
def test_origin_is_fixed_and_spectrum_is_exact(N, ring):
    """At the origin the state stays exactly zero and the QR spectrum converges to +-ln mu_j of the modes."""
    k = 16
    K, eps = 1.5, 0.1
    sysd = StandardMapChain(N, K, eps, "ring" if ring else "open").descriptor()
    m = 2
    x0 = np.zeros((m, 2 * N))
    W0 = O.frames(3, 1, m, 2 * N, k)
    n = 4000
    for mode in (O.RED_ROWS, O.RED_SITES, O.RED_BLOCKED):
        r = O.run_chain(sysd, QR[0], x0, W0, n, 10, k, mode=mode)
        assert np.all(r["x"] == 0) and np.all(r["code"] == 0)
        lam = lambdas(r, k)
        want = origin_exponents(N, K, eps, ring)[:k]
        assert np.max(np.abs(lam - want)) < 10.0 / n, (mode, lam[0], want[0])


def test_zero_coupling_is_independent_standard_maps():
    """eps = 0: every site steps as StandardMap(K) bit for bit (standard_map_chain.jl:31-32)."""
    N, K = 8, 1.5
    rng = np.random.default_rng(5)
    x0 = rng.uniform(0, 2 * np.pi, (3, 2 * N))
    W0 = O.frames(5, 1, 3, 2 * N, 1)
    r = O.run_chain(StandardMapChain(N, K, 0.0, "ring").descriptor(), NONE[0], x0, W0, 700, 1, 1)
    for i in range(3):
        xs = np.stack([x0[i, :N], x0[i, N:]], axis=1)
        s = O.run(StandardMap(K).descriptor(), NONE[0], xs, O.frames(5, 1, N, 2, 1), 700, 1, 1)
        assert np.array_equal(r["x"][i, :N], s["x"][:, 0]) and np.array_equal(r["x"][i, N:], s["x"][:, 1])


def dense_tangent(q, N, K, eps, ring):
    """M = [[I + G, I], [G, I]] with G at q, in 256-bit arithmetic (standard_map_chain.jl:91-104)."""
    mp.mp.prec = 256
    G = mp.zeros(N, N)
    for j in range(N):
        nb = [j - 1 if j > 0 else (N - 1 if ring else None), j + 1 if j < N - 1 else (0 if ring else None)]
        G[j, j] = mp.mpf(K) * mp.cos(mp.mpf(q[j]))
        for kk in nb:
            if kk is not None:
                c = mp.mpf(eps) * mp.cos(mp.mpf(q[j]) - mp.mpf(q[kk]))
                G[j, j] += c
                G[j, kk] = -c
    M = mp.zeros(2 * N, 2 * N)
    for a in range(N):
        for b in range(N):
            e = 1 if a == b else 0
            M[a, b] = G[a, b] + e
            M[a, N + b] = e
            M[N + a, b] = G[a, b]
            M[N + a, N + b] = e
    return M


@pytest.mark.parametrize("ring", [True, False])
def test_banded_step_equals_dense_product(ring):
    """One step of the banded tangent against the dense M W in 256-bit arithmetic, within 16 eps (1 + a_max) per
    entry (cm_ac17_19.jl:4-31), and the state step against the 256-bit map within a few ulps."""
    N, K, eps, k = 10, 3.0, 0.6102, 4
    rng = np.random.default_rng(17)
    for trial in range(3):
        x0 = rng.uniform(-4 * np.pi, 4 * np.pi, (1, 2 * N))
        W0 = rng.uniform(-1, 1, (1, 2 * N, k))
        r = O.run_chain(StandardMapChain(N, K, eps, "ring" if ring else "open").descriptor(), NONE[0], x0, W0, 1, 1,
                        k, mode=O.RED_ROWS)
        M = dense_tangent(x0[0, :N], N, K, eps, ring)
        Wm = mp.matrix(W0[0].tolist())
        ref = M * Wm
        tol = 16 * 2.0 ** -52 * (1 + abs(K) + 4 * abs(eps))
        err = max(abs(float(ref[i, j]) - r["W"][0, i, j]) for i in range(2 * N) for j in range(k))
        assert err <= tol * max(1.0, float(max(abs(v) for v in ref)))


def test_modes_agree_and_symplectic_sum():
    """Row-order and site-butterfly reductions give the same spectrum to rounding; with k = 2N the exponents of a
    symplectic product sum to zero (det M = 1) and pair as lambda_i + lambda_{2N+1-i} -> 0 (cm_ac13_16.jl:3-45)."""
    N, k, n = 8, 16, 3000
    sysd = StandardMapChain(N, 1.5, 0.1, "ring").descriptor()
    rng = np.random.default_rng(7)
    x0 = rng.uniform(0, 2 * np.pi, (4, 2 * N))
    W0 = O.frames(7, 1, 4, 2 * N, k)
    a = lambdas(O.run_chain(sysd, QR[0], x0, W0, n, 10, k, mode=O.RED_ROWS), k)
    b = lambdas(O.run_chain(sysd, QR[0], x0, W0, n, 10, k, mode=O.RED_SITES), k)
    c = lambdas(O.run_chain(sysd, QR[0], x0, W0, n, 10, k, mode=O.RED_BLOCKED), k)
    # rounding level: the last columns span directions contracted by about exp(-(lambda_1 - lambda_16) tau) = 3e-4
    # per handoff relative to the first, which amplifies a rounding difference in the factorisation by the inverse
    assert np.max(np.abs(a - b)) < 1e-11
    assert np.max(np.abs(a - c)) < 1e-11
    assert np.max(np.abs(c.sum(axis=1))) < 1e-10
    # det M = 1 exactly; the sum collects the rounding of 300 handoffs, each amplified by cond(M^tau) ~ 5e3 in the
    # contracted columns. A wrong index or sign in the band or the reflectors moves it by O(0.1).
    assert np.max(np.abs(a.sum(axis=1))) < 1e-10
    assert np.max(np.abs(a + a[:, ::-1])) < 0.05
    assert np.all(a[:, 0] > 0.1)  # chaotic at K = 1.5


def test_ragged_step_counts():
    """Per-orbit step counts: each orbit equals its own uniform run."""
    N, k = 16, 16
    sysd = StandardMapChain(N, 1.5, 0.1, "ring").descriptor()
    rng = np.random.default_rng(9)
    x0 = rng.uniform(0, 2 * np.pi, (3, 2 * N))
    W0 = O.frames(9, 1, 3, 2 * N, k)
    Ns = np.array([200, 50, 130])
    r = O.run_chain(sysd, QR[0], x0, W0, 0, 10, k, Ns=Ns)
    for i in range(3):
        s = O.run_chain(sysd, QR[0], x0[i:i + 1], W0[i:i + 1], int(Ns[i]), 10, k)
        assert np.array_equal(r["states"][i], s["states"][0]) and np.array_equal(r["W"][i], s["W"][0])
        assert r["hstep"][i] == Ns[i]


@pytest.mark.parametrize("N", [16, 128])
def test_blocked_order_against_row_order_and_planted_defects(N):
    """The blocked factorisation (two panels, compact WY trailing update, Q = E - U T U_top') gives the row order's
    frame to rounding after one handoff and its spectrum to rounding over 100 handoffs (N = 128: the Gram sums in two
    128-row chains). Negative controls: T instead of T' in the trailing update, or the wrong sign of T's off-diagonal
    entries, moves the frame by O(1) or fails the orthogonality check."""
    k = 16
    sysd = StandardMapChain(N, 1.5, 0.1, "ring").descriptor()
    rng = np.random.default_rng(21)
    x0 = rng.uniform(0, 2 * np.pi, (3, 2 * N))
    W0 = O.frames(21, 1, 3, 2 * N, k)
    r1 = O.run_chain(sysd, QR[0], x0, W0, 10, 10, k, mode=O.RED_ROWS)
    b1 = O.run_chain(sysd, QR[0], x0, W0, 10, 10, k, mode=O.RED_BLOCKED)
    assert np.all(b1["code"] == 0) and np.max(np.abs(b1["W"] - r1["W"])) < 1e-12
    a = lambdas(O.run_chain(sysd, QR[0], x0, W0, 1000, 10, k, mode=O.RED_ROWS), k)
    c = lambdas(O.run_chain(sysd, QR[0], x0, W0, 1000, 10, k, mode=O.RED_BLOCKED), k)
    assert np.max(np.abs(a - c)) < 1e-9
    for defect in (O.planted.BLK_TNOTRANS, O.planted.BLK_TSIGN):
        with O.planted(defect):
            d1 = O.run_chain(sysd, QR[0], x0, W0, 10, 10, k, mode=O.RED_BLOCKED)
        assert np.any(d1["code"] != 0) or np.max(np.abs(d1["W"] - r1["W"])) > 1e-3, defect


@pytest.mark.parametrize("N", [8, 32, 128])
def test_f32_tangent_policy(N):
    """Float32 tangents (elt 2): the orbit is the Float64 orbit bit for bit, the frames handed back hold Float32
    values, the leading exponents stay within 1e-6 of the Float64 run on the same orbit and the whole spectrum within
    1e-2 (the trailing columns carry the rounding of the frame amplified by the alignment over tau steps); at the
    origin, with tau inside the reference's tau_max(Float32) = floor(ln(1/eps32) / (2 (lambda_1 - lambda_k)))
    (tangent-core-spec.md:45-48), the spectrum is +-ln mu_j within 10/n as for Float64."""
    k = 16
    sysd = StandardMapChain(N, 1.5, 0.1, "ring").descriptor()
    rng = np.random.default_rng(7)
    x0 = rng.uniform(0, 2 * np.pi, (3, 2 * N))
    W0 = O.frames(7, 1, 3, 2 * N, k)
    for mode in (O.RED_BLOCKED, O.RED_SITES):
        a = O.run_chain(sysd, QR[0], x0, W0, 3000, 10, k, mode=mode)
        b = O.run_chain(sysd, QR[0], x0, W0, 3000, 10, k, mode=mode, elt=2)
        assert np.array_equal(a["x"], b["x"]) and np.all(b["code"] == 0)
        assert np.array_equal(b["W"], b["W"].astype(np.float32).astype(np.float64))
        la, lb = lambdas(a, k), lambdas(b, k)
        assert not np.array_equal(la, lb)  # the frames are rounded
        assert np.max(np.abs(la[:, :2] - lb[:, :2])) < 1e-6
        assert np.max(np.abs(la - lb)) < 1e-2
    n, tau = 4000, 2 if N == 8 else 10  # N = 8, k = 2N: lambda_1 - lambda_16 = 2.3 at the origin, tau_max = 3
    r = O.run_chain(sysd, QR[0], np.zeros((2, 2 * N)), O.frames(3, 1, 2, 2 * N, k), n, tau, k, mode=O.RED_BLOCKED,
                    elt=2)
    assert np.all(r["x"] == 0) and np.all(r["code"] == 0)
    assert np.max(np.abs(lambdas(r, k) - origin_exponents(N, 1.5, 0.1, True)[:k])) < 10.0 / n
