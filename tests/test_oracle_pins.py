"""Pin GALI_k (k >= 3), OFLI and MEGNO's time averages in the CPU oracle against the reference suites' closed forms.

Each check restates a reference acceptance criterion and evaluates the exact answer in 200-256-bit arithmetic
(mpmath). Each has a negative control: the oracle is run again with one planted defect (oracle.cpp `g_defect`,
switched by `O.planted`), and the same check must then fail. Runs on the CPU.

- GALI: prescribed Gram matrices (test/gali/ga_ac01_04.jl:56-107) on the oracle's linear map with A = I, which hands
  the frame to the reducer unchanged; exact power-of-two column scalings (ga_ac01_04.jl AC3e).
- OFLI: the free particle in d = 4 and d = 2 (test/ofli/of_ac01_05.jl:1-143), as the linear map [[I, hI], [0, I]]
  whose displacement field (h p, 0) has the flow's direction bit for bit; the basis-free sandwich
  F* - ln sqrt(4) <= OFLI_B <= F* on Henon-Heiles orbits (test/ofli/of_ac06_10.jl:89-111).
- MEGNO: the M5 sums (test/megno/mg_helpers.jl:117-138) on the exact norms of the K = 0 shear and of the Y5 map,
  all three series Y, Ybar, Ybar20 and both columns (test/megno/mg_ac01_05.jl:141-286), within 6 B_S(n).

What values cannot show: a dropped second Gram-Schmidt pass in GALI. The values are products of the diagonal of R,
which modified Gram-Schmidt forms backward stably after one pass; a numpy restatement with one pass stays below 6% of
the bound on every set here (equal-cosine sets down to 1 - 1e-9, k up to 8). The second pass is the reference's
order of operations and is held by the bit-for-bit device comparison, not by a closed form."""
import math

import mpmath as mp
import numpy as np
import pytest

from chaos_amd import HenonHeilesGauss
from chaos_amd import _lib as L
from chaos_amd.inputs import hh_section_starts
from chaos_amd.reducers import GALI, OFLI, MEGNO, FLI
from chaos_amd.tangent import make_chain, decode_states, finals_of
from tests import oracle_lib as O

EPS = 2.0 ** -52
mp.mp.prec = 256


def _finals(out, reducers, d, k):
    _, allr = make_chain(reducers)
    return finals_of(decode_states(out["states"], allr, d, k), allr, d, k)


# ------------------------------------------------------------------------------------------------------ GALI
def _randorth(rng, d):
    q, r = np.linalg.qr(rng.standard_normal((d, d)))
    return q * np.sign(np.diag(r))


def _gram_set(G, d, Q):
    """Vectors with Gram matrix G (k x k, mpmath): columns of the 256-bit Cholesky factor's transpose, placed in d
    dimensions by the orthogonal map Q, rounded to float64 (ga_gram_vectors, test/gali/ga_helpers.jl)."""
    k = G.rows
    Lc = mp.cholesky(G)
    V = mp.matrix(d, k)
    for i in range(d):
        for j in range(k):
            V[i, j] = mp.fsum(mp.mpf(Q[i, r]) * Lc[j, r] for r in range(k))
    return np.array([[float(V[i, j]) for j in range(k)] for i in range(d)])


def _exact_gali(W):
    """sqrt(det Gram) of the unit columns of the float64 set W, orders 2..k, 256 bits (ga_exact)."""
    d, k = W.shape
    U = [[mp.mpf(W[i, j]) for i in range(d)] for j in range(k)]
    U = [[x / mp.sqrt(mp.fsum(y * y for y in u)) for x in u] for u in U]
    out = []
    for kk in range(2, k + 1):
        G = mp.matrix(kk, kk)
        for a in range(kk):
            for b in range(kk):
                G[a, b] = mp.fsum(U[a][i] * U[b][i] for i in range(d))
        out.append(mp.sqrt(max(mp.det(G), mp.mpf(0))))
    return out


def _gali_values(sets):
    """GALI_2..GALI_k of each set through the oracle's reducer: one step of A = I, chain (GALI,)."""
    out = []
    for W in sets:
        d, k = W.shape
        ch, _ = make_chain((GALI(),))
        r = O.run_linear_d(np.eye(d), ch, np.zeros((1, d)), W[None], 1, 1, k)
        assert r["code"][0] == 0
        out.append(_finals(r, (GALI(),), d, k)[0]["gali"][0])
    return out


def _gali_cases():
    """(set, closed forms of orders 2..k) of ga_ac01_04.jl AC2 (a)-(d), with closed forms only at order k."""
    rng = np.random.default_rng(20261012 + 40100)
    cases = []
    for m in (1, 4, 8, 14):
        for ang in (mp.mpf(10) ** -m, mp.pi / 2 - mp.mpf(10) ** -m):
            c = mp.cos(ang)
            for d in (2, 4, 8):
                W = _gram_set(mp.matrix([[1, c], [c, 1]]), d, _randorth(rng, d))
                cases.append((W, {2: mp.sin(ang)}))
    eq = lambda k, c: mp.matrix([[1 if i == j else c for j in range(k)] for i in range(k)])
    for c in [mp.mpf(x) for x in (0.0, 0.5, -0.4, 0.9)] + [1 - mp.mpf(10) ** -m for m in (1, 3, 5, 7)]:
        for d in (3, 4, 8):
            W = _gram_set(eq(3, c), d, _randorth(rng, d))
            cases.append((W, {3: mp.sqrt(max((1 + 2 * c) * (1 - c) ** 2, 0))}))
    for c in (0.0, 0.3, -0.3):
        c = mp.mpf(c)
        for d in (4, 8):
            W = _gram_set(eq(4, c), d, _randorth(rng, d))
            cases.append((W, {4: mp.sqrt((1 + 3 * c) * (1 - c) ** 3)}))
    for d in (4, 6, 8):  # (d) k = d random sets: |det| of the unit-column matrix
        for _ in range(3):
            W = rng.standard_normal((d, d))
            U = mp.matrix([[mp.mpf(W[i, j]) for j in range(d)] for i in range(d)])
            for j in range(d):
                nrm = mp.sqrt(mp.fsum(U[i, j] ** 2 for i in range(d)))
                for i in range(d):
                    U[i, j] /= nrm
            cases.append((W, {d: abs(mp.det(U))}))
    return cases


def _gali_worst(cases, values):
    """Largest |G_k - closed form| / (3 B_meas(k)), B_meas = 2 k eps (ga_meas), over every order with a closed form
    and every order against the exact evaluation of the rounded set."""
    worst = 0.0
    for (W, closed), g in zip(cases, values):
        ex = _exact_gali(W)
        for kk in range(2, W.shape[1] + 1):
            B = 3 * 2 * kk * EPS
            worst = max(worst, float(abs(mp.mpf(g[kk - 2]) - ex[kk - 2])) / B)
            if kk in closed:
                worst = max(worst, float(abs(mp.mpf(g[kk - 2]) - closed[kk])) / B)
    return worst


def test_gali_prescribed_gram_closed_forms():
    """ga_ac01_04.jl:56-107 AC2: GALI_2 = sin theta, GALI_3 = sqrt((1 + 2c)(1 - c)^2), GALI_4 = sqrt((1 + 3c)(1 - c)^3),
    |det| at k = d, within 3 B_meas(T, k); every order also against the 256-bit evaluation of the rounded set."""
    cases = _gali_cases()
    assert _gali_worst(cases, _gali_values([W for W, _ in cases])) <= 1.0


def test_gali_scaled_columns_and_planted_defects():
    """ga_ac01_04.jl AC3e: columns scaled by powers of two (exact) give the unscaled values within 2 B_meas. Negative
    controls: raw (unnormalised) columns and a product missing its last factor fail the closed forms."""
    cases = [c for c in _gali_cases() if c[0].shape[1] >= 3]
    rng = np.random.default_rng(7)
    scaled = []
    for W, closed in cases:
        s = 2.0 ** rng.integers(-40, 41, size=W.shape[1]).astype(np.float64)
        scaled.append((W * s, closed))
    base = _gali_values([W for W, _ in cases])
    sv = _gali_values([W for W, _ in scaled])
    for (W, _), a, b in zip(cases, base, sv):
        for kk in range(2, W.shape[1] + 1):
            assert abs(a[kk - 2] - b[kk - 2]) <= 2 * 2 * kk * EPS
    assert _gali_worst(scaled, sv) <= 1.0
    with O.planted(O.planted.GALI_RAW):
        assert _gali_worst(scaled, _gali_values([W for W, _ in scaled])) > 1e3
    with O.planted(O.planted.GALI_LASTFACTOR):
        assert _gali_worst(cases, _gali_values([W for W, _ in cases])) > 1e3
    assert _gali_worst(cases, _gali_values([W for W, _ in cases])) <= 1.0


# ------------------------------------------------------------------------------------------------------ OFLI
def _free_g(i, p0, t, ndof):
    """of_free_g (test/ofli/of_ac01_05.jl:8-18): ln ||P w(t)|| of column e_i of the free particle, w = (w_q + t w_p, w_p),
    field (p0, 0); -inf for a zero component."""
    d = 2 * ndof
    w = [mp.mpf(0)] * d
    w[i] = mp.mpf(1)
    for j in range(ndof):
        w[j] += t * w[ndof + j]
    f = [mp.mpf(x) for x in p0] + [mp.mpf(0)] * ndof
    c = mp.fsum(a * b for a, b in zip(f, w)) / mp.fsum(a * a for a in f)
    pw = [a - c * b for a, b in zip(w, f)]
    b = mp.sqrt(mp.fsum(x * x for x in pw))
    return mp.log(b) if b > 0 else -mp.inf


def _free_A(ndof, h):
    d = 2 * ndof
    A = np.eye(d)
    A[:ndof, ndof:] = h * np.eye(ndof)
    return A


def _ofli_free_ratio(p0, cols, N, tau, h=1 / 16, k1=True):
    """Worst |max(g0, OFLI_1) - max(g_cf(0), g_cf(N h))| / (32 eps max(1, |g_cf|)) over the single columns `cols`
    (of_ac1_ratio, of_ac01_05.jl:19-40; g is nondecreasing on the free particle, so the running maximum is g(N))."""
    ndof = len(p0)
    d = 2 * ndof
    A = _free_A(ndof, h)
    z0 = np.array([0.0] * ndof + list(p0))[None]
    worst = 0.0
    for i in cols:
        W0 = np.zeros((1, d, 1))
        W0[0, i, 0] = 1.0
        ch, _ = make_chain((OFLI(),))
        r = O.run_linear_d(A, ch, z0, W0, N, tau, 1)
        f = _finals(r, (OFLI(),), d, 1)[0]
        g0 = _free_g(i, p0, mp.mpf(0), ndof)  # the s = 0 term (ofli_initial)
        gc = max(g0, mp.mpf(f["ofli1"][0]))
        gf = max(g0, _free_g(i, p0, N * mp.mpf(h), ndof))
        if gf == -mp.inf:
            worst = max(worst, 0.0 if gc == gf else math.inf)
        else:
            worst = max(worst, float(abs(gc - gf)) / (32 * EPS * max(1.0, abs(float(gf)))))
    return worst


@pytest.mark.parametrize("tau", [1, 10])
def test_ofli_free_particle_d4(tau):
    """of_ac01_05.jl:42-61 AC1: h = 2^-4, p0 = (0.6, 0.8), columns e_1..e_4, and p0 = (1/2, 0) columns 3, 4, to
    n = 1e4, within 32 eps max(1, |g|); the k = 4 canonical frame's OFLI_B is the maximum of the four closed forms."""
    for N in (10, 1000, 10000):
        assert _ofli_free_ratio((0.6, 0.8), range(4), N, tau) <= 1
        assert _ofli_free_ratio((0.5, 0.0), (2, 3), N, tau) <= 1
    A = _free_A(2, 1 / 16)
    ch, _ = make_chain((OFLI(),))
    z0 = np.array([[0.0, 0.0, 0.6, 0.8]])
    for N in (10, 1000):
        r = O.run_linear_d(A, ch, z0, np.eye(4)[None], N, 10, 4)
        f = _finals(r, (OFLI(),), 4, 4)[0]
        gmax = max(_free_g(i, (0.6, 0.8), N * mp.mpf(1) / 16, 2) for i in range(4))
        gB0 = max(_free_g(i, (0.6, 0.8), mp.mpf(0), 2) for i in range(4))
        ref = max(gB0, gmax)
        assert abs(max(gB0, mp.mpf(f["ofli"][0])) - ref) <= 16 * EPS * max(1.0, abs(float(ref)))


def test_ofli_free_particle_d2_and_parallel_start():
    """of_ac01_05.jl:75-111 AC2: w = (0, 1) gives OFLI_1 = ln|w_p| = 0 at every handoff and FLI_1 = ln sqrt(t^2 + 1);
    w = (0.6, 0.8) gives ln 0.8; a start along the flow, w = (1, 0), gives -inf at every handoff, all counted."""
    A = _free_A(1, 1 / 16)
    z0 = np.array([[0.0, 0.5]])
    red = (OFLI(),)
    ch, _ = make_chain(red)
    for N, tau in ((10, 1), (10000, 1), (10000, 10)):
        f = _finals(O.run_linear_d(A, ch, z0, np.array([[[0.0], [1.0]]]), N, tau, 1), red, 2, 1)[0]
        assert abs(f["ofli1"][0]) <= 16 * EPS
        t = N / 16
        Fcf = 0.5 * math.log(t * t + 1)
        assert abs(f["fli1"][0] - Fcf) <= max(16 * EPS * Fcf, 8 * EPS * (N + 2))
        f = _finals(O.run_linear_d(A, ch, z0, np.array([[[0.6], [0.8]]]), N, tau, 1), red, 2, 1)[0]
        assert abs(max(math.log(0.8), f["ofli1"][0]) - math.log(0.8)) <= 8 * EPS * (N / tau + 2)
    f = _finals(O.run_linear_d(A, ch, z0, np.array([[[1.0], [0.0]]]), 100, 10, 1), red, 2, 1)[0]
    assert f["ofli1"][0] == -math.inf and f["nminf"][0] == 10


def test_ofli_free_particle_planted_no_projection():
    """Negative control (of_ac01_05.jl:63-73): an OFLI that does not project out the field fails AC1."""
    with O.planted(O.planted.OFLI_NOPROJ):
        assert _ofli_free_ratio((0.6, 0.8), range(4), 1000, 1) > 1e6
    assert _ofli_free_ratio((0.6, 0.8), range(4), 1000, 1) <= 1


HH = HenonHeilesGauss(h=0.01, tol=1e-15)


def _hh_field(z):
    """The Henon-Heiles vector field (dq = p, dp = -grad V), V = (q1^2 + q2^2)/2 + q1^2 q2 - q2^3/3
    (src/Systems/flow_models.jl:101-102)."""
    q1, q2, p1, p2 = z
    return np.array([p1, p2, -(q1 + 2 * q1 * q2), -(q2 + q1 * q1 - q2 * q2)])


def _sandwich(z0, N, tau):
    """(G, F*) at the last handoff: G = max(max g0, OFLI_B(N)) of OFLI(flow) on the canonical frame, and F* the running
    maximum over the handoffs (and the s = 0 term 0) of ln sigma_1(P_x Y_n), Y_n the cumulative tangent of the same orbit
    (of_ac06_10.jl:89-111)."""
    d = 4
    red = (OFLI(),)
    ch, _ = make_chain(red)
    r = O.run(HH.descriptor(), ch, z0[None], np.eye(d)[None], N, tau, d, nthreads=1)
    f = _finals(r, red, d, d)[0]
    u0 = _hh_field(z0)
    u0 = u0 / np.linalg.norm(u0)
    P0 = np.eye(d) - np.outer(u0, u0)
    g0 = max(math.log(np.linalg.norm(P0[:, i])) for i in range(d))
    Fstar = 0.0
    ch0, _ = make_chain(())
    for n in range(tau, N + 1, tau):
        y = O.run(HH.descriptor(), ch0, z0[None], np.eye(d)[None], n, tau, d, nthreads=1)
        x, Y = y["x"][0], y["W"][0]
        u = _hh_field(x)
        u = u / np.linalg.norm(u)
        P = np.eye(d) - np.outer(u, u)
        Fstar = max(Fstar, math.log(np.linalg.svd(P @ Y, compute_uv=False)[0]))
    return max(g0, float(f["ofli"][0])), Fstar


def test_ofli_basis_free_sandwich_henon_heiles():
    """of_ac06_10.jl:89-111 AC9: F* - ln sqrt(4) <= OFLI_B <= F* (within 1e-9) on regular and chaotic orbits at
    E = 1/8 and 1/6, 400 steps, tau = 10. Negative control: without the projection the upper bound fails."""
    starts = np.concatenate([hh_section_starts(3, 2, 1 / 8), hh_section_starts(5, 2, 1 / 6)])
    B = 1e-9
    frac = []
    for z0 in starts:
        G, F = _sandwich(z0, 400, 10)
        assert F - math.log(2.0) - B <= G <= F + B
        frac.append((F - G) / math.log(2.0))
    assert all(-1e-9 <= x <= 1 + 1e-9 for x in frac)
    with O.planted(O.planted.OFLI_NOPROJ):
        bad = [_sandwich(z0, 400, 10) for z0 in starts]
    assert any(G > F + 1e-3 for G, F in bad)


# ------------------------------------------------------------------------------------------------------ MEGNO
def _m5(sh, tau, q2w=3):
    """mg_M5 (test/megno/mg_helpers.jl:117-138) on the exact handoff sums sh[0..K-1] = s(k tau): Y, Ybar, Ybar20 at every
    handoff. q2w is the Q2 weight (3 in the definition; other values make the test-side defective double)."""
    Y, Yb, Y20 = [], [], []
    R1 = R2 = Q = Q2 = prev = mp.mpf(0)
    for kk, s in enumerate(sh, 1):
        t = mp.mpf(kk * tau)
        a = s - prev
        prev = s
        R1 += t * a
        R2 += t * t * a
        y = 2 * R1 / t
        Q += y
        Q2 += q2w * R2
        Y.append(y)
        Yb.append(Q / kk)
        Y20.append(tau * Q2 / t ** 3)
    return Y, Yb, Y20


def _exact_s(Apow, w, steps):
    """s(n) = (1/2) ln(||A^n w||^2 / ||w||^2) at the steps, A^n from the exact integer powers `Apow(n)`."""
    a, b = mp.mpf(w[0]), mp.mpf(w[1])
    q0 = a * a + b * b
    out = []
    for n in steps:
        M = Apow(n)
        y0 = M[0][0] * a + M[0][1] * b
        y1 = M[1][0] * a + M[1][1] * b
        out.append(mp.log((y0 * y0 + y1 * y1) / q0) / 2)
    return out


def _megno_worst(sys_runner, Apow, Gfirst, W0s, x0, N, tau, ns, q2w=3):
    """Largest ratio |series - M5| / 6 B_S(n), B_S = 8 eps (n + G_first), over Y, Ybar, Ybar20, both columns, the
    frames W0s and the end steps ns (mg_ac01_05.jl:141-206)."""
    worst = 0.0
    red = (MEGNO(),)
    ch, _ = make_chain(red)
    for n in ns:
        out = sys_runner(ch, np.repeat(np.asarray(x0)[None], len(W0s), axis=0), np.array(W0s), n, tau)
        f = _finals(out, red, 2, 2)[0]
        B = 6 * 8 * EPS * (n + Gfirst)
        for j, W0 in enumerate(W0s):
            for i in range(2):
                sh = _exact_s(Apow, W0[:, i], range(tau, n + 1, tau))
                Y, Yb, Y20 = _m5(sh, tau, q2w)
                for key, ref in (("Y", Y[-1]), ("Ybar", Yb[-1]), ("Ybar20", Y20[-1])):
                    worst = max(worst, float(abs(mp.mpf(f[key][j, i]) - ref)) / B)
    return worst


def _shear_runner(ch, x0, W0, n, tau):
    s = L.cl_system()
    s.kind = L.CL_SYS_STANDARD_MAP
    s.K = 0.0
    return O.run(s, ch, x0, W0, n, tau, 2, nthreads=1)


def _y5_runner(ch, x0, W0, n, tau):
    return O.run_linear_d(np.array([[2.0, 1.0], [1.0, 1.0]]), ch, x0, W0, n, tau, 2)


def _shear_pow(n):
    return [[1, n], [0, 1]]


def _y5_pow(n):
    M = [[1, 0], [0, 1]]
    for _ in range(n):
        M = [[2 * M[0][0] + M[1][0], 2 * M[0][1] + M[1][1]], [M[0][0] + M[1][0], M[0][1] + M[1][1]]]
    return M


@pytest.mark.parametrize("tau", [1, 10])
def test_megno_shear_all_series(tau):
    """mg_ac01_05.jl:141-206 AC2: the K = 0 shear, W0 = I and seeded frames, n in {10, 100, 1e4}: Y, Ybar, Ybar20 of both
    columns within 6 B_S(n) of M5 on the exact norms (G_first = ||M^tau||_F^2 = 2 + tau^2); column 1 of W0 = I exactly 0."""
    W0s = [np.eye(2)] + list(O.frames(20261012, 1, 4, 2, 2))
    ns = [n for n in (10, 100, 10000) if n % tau == 0]
    assert _megno_worst(_shear_runner, _shear_pow, 2 + tau * tau, W0s, [0.3, 0.7], 10000, tau, ns) <= 1
    red = (MEGNO(),)
    ch, _ = make_chain(red)
    f = _finals(_shear_runner(ch, np.array([[0.3, 0.7]]), np.eye(2)[None], 100, tau), red, 2, 2)[0]
    assert f["Y"][0, 0] == 0 and f["Ybar"][0, 0] == 0 and f["Ybar20"][0, 0] == 0


@pytest.mark.parametrize("tau", [1, 5])
def test_megno_y5_all_series(tau):
    """mg_ac01_05.jl:208-286 AC3(a): A = [[2, 1], [1, 1]] from the origin, seeded frames, n up to 200: the three series of
    both columns within 6 B_S(n), G_first = ||A^tau||_F^2."""
    W0s = list(O.frames(20261012 + 3, 1, 6, 2, 2))
    ns = [n for n in (1, 2, 4, 8, 20, 100, 200) if n % tau == 0]
    At = np.array(_y5_pow(tau), dtype=np.float64)
    assert _megno_worst(_y5_runner, _y5_pow, float(np.sum(At * At)), W0s, [0.0, 0.0], 200, tau, ns) <= 1


def test_megno_planted_q2_weight():
    """Negative control (mg_ac01_05.jl: the n^3 and weight doubles): with Q2 accumulating 2 R2 in place of 3 R2, Ybar20
    leaves the bound; the test-side M5 with weight 2 matches the defective oracle instead."""
    W0s = [np.eye(2)] + list(O.frames(20261012, 1, 2, 2, 2))
    with O.planted(O.planted.MEGNO_Q2):
        assert _megno_worst(_shear_runner, _shear_pow, 3, W0s, [0.3, 0.7], 1000, 1, [1000]) > 1e3
        assert _megno_worst(_shear_runner, _shear_pow, 3, W0s, [0.3, 0.7], 1000, 1, [1000], q2w=2) <= 1
    assert _megno_worst(_shear_runner, _shear_pow, 3, W0s, [0.3, 0.7], 1000, 1, [1000]) <= 1
