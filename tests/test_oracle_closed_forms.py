"""Pin the CPU oracle against the reference suites' closed forms, exact answers and invariants.

Each test names the reference test it restates. Values that only an import of the reference could
give (bitwise Julia sin/cos/log, StableRNG frames, Double64 arithmetic) are not claimed; see the
oracle's header. Runs on the CPU (no GPU needed)."""
import ctypes
import math

import mpmath as mp
import numpy as np
import pytest

from chaos_amd import _lib as L
from chaos_amd.reducers import (QRSpectrum, CGS, MGS, RGS, HouseholderQR, AugmentedQR, GivensQR, MEGNO, SALI, FLI,
                                GALI, Benettin, OFLI)
from chaos_amd.tangent import make_chain, decode_states, finals_of
from tests import oracle_lib as O

EPS = 2.0 ** -52
ALGS = [CGS(), MGS(), RGS(), HouseholderQR(), AugmentedQR(), GivensQR()]


def std(K):
    s = L.cl_system()
    s.kind = L.CL_SYS_STANDARD_MAP
    s.K = K
    return s


def finals(out, reducers, k, reducers2=()):
    _, allr = make_chain(reducers, reducers2)
    return finals_of(decode_states(out["states"], allr, 2, k), allr, 2, k)


def test_compensated_sum_ac9():
    """test/tangent-core/tc_ac09_11.jl:3-23: 1e8 fl(0.1) within eps; Neumaier [1, 1e100, 1, -1e100] = 2."""
    o = O.load()
    comp, plain = ctypes.c_double(), ctypes.c_double()
    n = 10 ** 8
    o.or_csum(n, 0.1, ctypes.byref(comp), ctypes.byref(plain))
    exact = mp.mpf(n) * mp.mpf(0.1)
    rel = lambda v: float(abs(mp.mpf(v) - exact) / exact)
    assert rel(comp.value) <= EPS
    assert rel(plain.value) > 100 * rel(comp.value) + 1e-10
    v = np.array([1.0, 1e100, 1.0, -1e100])
    out = ctypes.c_double()
    o.or_csum_list(4, v.ctypes.data, ctypes.byref(out))
    assert out.value == 2.0


def test_y5_fibonacci_bitwise_ac1():
    """tc_ac01_04.jl:4-17: A = [[2,1],[1,1]], W0 = I, W_n = A^n and x_n = A^n x0 bitwise, n = 1..35."""
    A = np.array([[2.0, 1.0], [1.0, 1.0]])
    ch, _ = make_chain(())
    for n in (1, 2, 7, 20, 35):
        out = O.run_linear(A, ch, [[3.0, 5.0]], [np.eye(2)], n, 1, 2)
        An = np.linalg.matrix_power(np.array([[2, 1], [1, 1]], dtype=object), n).astype(np.float64)
        assert np.array_equal(out["W"][0], An)
        assert np.array_equal(out["x"][0], An @ np.array([3.0, 5.0]))


@pytest.mark.parametrize("math_mode", [O.MATH_LIBM, O.MATH_CL])
def test_k0_shear_bitwise_ac3(math_mode):
    """tc_ac01_04.jl:52-64: K = 0 gives W_n = [[1,n],[0,1]] and x_n = x0 + n p0 bitwise (never wrapped)."""
    ch, _ = make_chain(())
    for n in (1, 10, 1000, 10 ** 6):
        out = O.run(std(0.0), ch, [[0.25, 0.5]], [np.eye(2)], n, 1, 2, math=math_mode, nthreads=1)
        assert np.array_equal(out["W"][0], np.array([[1.0, float(n)], [0.0, 1.0]]))
        assert np.array_equal(out["x"][0], np.array([0.25 + 0.5 * n, 0.5]))


def test_fixed_point_exact():
    """test/standard-map/sm_ac07_13.jl:136-148: the fixed point (0, 0) stays exact for 1e4 steps."""
    ch, _ = make_chain(())
    for K in (0.5, 2.0, 6.0):
        out = O.run(std(K), ch, [[0.0, 0.0]], [np.eye(2)], 10 ** 4, 1, 2, nthreads=1)
        assert np.array_equal(out["x"][0], [0.0, 0.0])


@pytest.mark.parametrize("alg", ALGS, ids=lambda a: a.name)
def test_qr_k0_shear_ac2(alg):
    """gs_ac01_05.jl:27-47: Q0 = (e2, e1), Y_n = [[1,n],[0,1]]: S1 = ln(n^2+1)/2 and S2 = -S1 to 8 n eps;
    lambda1(1e4) <= ln(1e4)/1e4 + 1e-12."""
    red = (QRSpectrum(alg),)
    ch, _ = make_chain(red)
    W0 = np.array([[0.0, 1.0], [1.0, 0.0]])
    for n in (1, 10, 1000, 10 ** 4):
        out = O.run(std(0.0), ch, [[0.3, 0.7]], [W0], n, 1, 2, nthreads=1)
        assert out["code"][0] == 0
        f = finals(out, red, 2)[0]
        S = f["sum"][0] + f["comp"][0]
        ref = mp.log(mp.mpf(n) ** 2 + 1) / 2
        assert abs(mp.mpf(S[0]) - ref) <= 8 * n * EPS
        assert abs(S[1] + S[0]) <= 8 * n * EPS
        if n == 10 ** 4:
            assert f["lambda"][0, 0] <= math.log(1e4) / 1e4 + 1e-12


@pytest.mark.parametrize("alg", ALGS, ids=lambda a: a.name)
def test_qr_cat_map_ac1(alg):
    """gs_ac01_05.jl:3-25: cat map [[2,1],[1,1]] from its fixed point: lambda1 = ln mu + ln(c+^2 + c-^2 mu^-4n)/(2n)
    within theta = 16 eps (G_max/tau + M_max), lambda1 + lambda2 within theta; 20 seeded frames."""
    A = np.array([[2.0, 1.0], [1.0, 1.0]])
    red = (QRSpectrum(alg),)
    ch, _ = make_chain(red)
    W0 = O.frames(20261012 + 1, 1, 20, 2, 2)
    s5 = mp.sqrt(5)
    mu = (3 + s5) / 2
    g = (s5 - 1) / 2
    nrm = mp.sqrt(1 + g * g)
    for tau in (1, 2, 4):
        for n in (40, 400):
            out = O.run_linear(A, ch, np.zeros((20, 2)), W0, n, tau, 2)
            f = finals(out, red, 2)[0]
            for j in range(20):
                q = W0[j][:, 0]
                cp = (mp.mpf(q[0]) + g * mp.mpf(q[1])) / nrm
                cm = (-g * mp.mpf(q[0]) + mp.mpf(q[1])) / nrm
                lam1 = mp.log(mu) + mp.log(cp ** 2 + cm ** 2 * mu ** (-4 * n)) / (2 * n)
                th = 16 * EPS * (f["G_max"][j] / tau + 2.0)
                assert abs(mp.mpf(f["lambda"][j, 0]) - lam1) <= th
                assert abs(f["lambda"][j, 0] + f["lambda"][j, 1]) <= th


def _shear_megno_closed(n):
    lp = mp.fsum(mp.log(mp.mpf(i) ** 2 + 1) for i in range(n))
    return mp.log(mp.mpf(n) ** 2 + 1) - lp / n


def test_megno_shear_ac2():
    """mg_ac01_05.jl:141-206: K = 0, W0 = I, tau = 1: column 1 exactly 0; column 2
    Y(n) = ln(n^2+1) - (1/n) sum_{i<n} ln(i^2+1) within 6*8 eps (n + G), and for n >= 100
    |Y - 2 - (ln n - ln 2 sinh pi)/n| <= 3/n^2 + bound."""
    red = (MEGNO(),)
    ch, _ = make_chain(red)
    mp.mp.prec = 200
    for n in (10, 100, 1000):
        out = O.run(std(0.0), ch, [[0.3, 0.7]], [np.eye(2)], n, 1, 2, nthreads=1)
        f = finals(out, red, 2)[0]
        assert f["Y"][0, 0] == 0 and f["Ybar"][0, 0] == 0 and f["Ybar20"][0, 0] == 0
        B = 6 * 8 * EPS * (n + 2.62)
        assert abs(mp.mpf(f["Y"][0, 1]) - _shear_megno_closed(n)) <= B
        if n >= 100:
            asym = 2 + (mp.log(n) - mp.log(2 * mp.sinh(mp.pi))) / n
            assert abs(mp.mpf(f["Y"][0, 1]) - asym) <= 3.0 / n ** 2 + B


def test_sali_fli_gali_shear():
    """test/sali/sl_ac01_03.jl (AC3), test/fli-ofli/fl_ac01_05.jl (AC2), test/gali AC2 on the K = 0 shear with W0 = I:
    u1 = e1, u2 = (n, 1)/sqrt(n^2+1): SALI = sqrt(2 - 2n/sqrt(n^2+1)), GALI_2 = 1/sqrt(n^2+1),
    FLI_B = S_2 = ln(n^2+1)/2, FLI_1 = 0, and the log sums of MEGNO, FLI, SALI bitwise equal (mg_ac06_08.jl:47)."""
    mp.mp.prec = 200
    chains = [(SALI(),), (FLI(),), (GALI(),), (MEGNO(replace=False), FLI(replace=False), SALI())]
    for n in (1, 10, 1000, 10 ** 5):
        res = {}
        for red in chains:
            ch, _ = make_chain(red)
            out = O.run(std(0.0), ch, [[0.3, 0.7]], [np.eye(2)], n, 1, 2, nthreads=1)
            res[red] = finals(out, red, 2)
        sali = res[chains[0]][0]["sali"][0]
        ref = mp.sqrt(2 - 2 * n / mp.sqrt(mp.mpf(n) ** 2 + 1))
        assert abs(mp.mpf(sali) - ref) <= 8 * EPS * max(1.0, n * 1e-0) * 4
        fl = res[chains[1]][0]
        assert fl["fli1"][0] == 0.0
        assert abs(mp.mpf(fl["fli"][0]) - mp.log(mp.mpf(n) ** 2 + 1) / 2) <= 8 * n * EPS
        ga = res[chains[2]][0]["gali"][0, 0]
        assert abs(mp.mpf(ga) - 1 / mp.sqrt(mp.mpf(n) ** 2 + 1)) <= 8 * EPS * 4 / math.sqrt(n * n + 1) + 4 * EPS
        mf = res[chains[3]]
        assert np.array_equal(mf[0]["sum"], mf[1]["sum"]) and np.array_equal(mf[1]["sum"], mf[2]["sum"])
        assert np.array_equal(mf[0]["comp"], mf[2]["comp"])


def test_benettin_equals_qr_k1():
    """A k = 1 QR spectrum and Benettin perform the same normalisation: equal sums on the same orbit."""
    x0 = np.array([[1.0, 2.0], [0.1, 0.2], [3.0, 0.5]])
    W0 = O.frames(5, 1, 3, 2, 1)
    rb, rq = (Benettin(),), (QRSpectrum(MGS()),)
    ob = O.run(std(2.0), make_chain(rb)[0], x0, W0, 2000, 1, 1, nthreads=1)
    oq = O.run(std(2.0), make_chain(rq)[0], x0, W0, 2000, 1, 1, nthreads=1)
    fb, fq = finals(ob, rb, 1)[0], finals(oq, rq, 1)[0]
    assert np.allclose(fb["lambda"][:, 0], fq["lambda"][:, 0], rtol=0, atol=1e-13)


def test_same_orbit_tangent_bound_ac4():
    """tc_ac01_04.jl:66-88 with tc_std_residual (tc_helpers.jl:151-166): along the realised orbit points the product
    differs from a 256-bit recomputation by at most 8 n eps prod ||M_i||_F ||W0||_F. Libm mode, so the orbit is
    re-traced here with the same glibc sin."""
    mp.mp.prec = 256
    rng = np.random.default_rng(20261012 + 4)
    ch, _ = make_chain(())
    for K, n in ((0.5, 100), (2 * math.pi, 20)):
        x0 = 2 * math.pi * rng.random((20, 2))
        W0 = O.frames(20261016, 1, 20, 2, 2)
        out = O.run(std(K), ch, x0, W0, n, 1, 2, math=O.MATH_LIBM, nthreads=1)
        for j in range(20):
            z = list(x0[j])
            Wref = mp.matrix(W0[j].tolist())
            prodM = mp.mpf(1)
            for _ in range(n):
                c = mp.mpf(K) * mp.cos(mp.mpf(z[0]))
                M = mp.matrix([[1 + c, 1], [c, 1]])
                prodM *= mp.sqrt(sum(M[i, jj] ** 2 for i in range(2) for jj in range(2)))
                Wref = M * Wref
                p = z[1] + K * math.sin(z[0])
                z = [z[0] + p, p]
            assert np.array_equal(out["x"][j], np.array(z))
            err = mp.sqrt(sum((mp.mpf(out["W"][j][a, b]) - Wref[a, b]) ** 2 for a in range(2) for b in range(2)))
            bound = 8 * n * EPS * prodM * mp.sqrt(sum(mp.mpf(v) ** 2 for v in W0[j].ravel()))
            assert err <= bound


def test_frames_properties():
    """tc_ac05_08.jl:37-79 properties of the seeded frames: orthonormal to 8 k eps, column j independent of k,
    distinct orbits give distinct frames."""
    for d, k in ((2, 1), (2, 2)):
        W = O.frames(1, 1, 500, d, k)
        for w in W:
            assert np.max(np.abs(w.T @ w - np.eye(k))) <= 8 * k * EPS
    W1, W2 = O.frames(7, 1, 200, 2, 1), O.frames(7, 1, 200, 2, 2)
    assert np.array_equal(W1[:, :, 0], W2[:, :, 0])
    assert len({tuple(w.ravel()) for w in W2}) == 200


def test_frames_haar_angle():
    """The first column's angle on the circle is uniform (KS against U(0, 2 pi)), the Haar property the reference
    tests with a KS statistic (tc_ac05_08.jl:37-79)."""
    from scipy import stats
    W = O.frames(11, 1, 20000, 2, 2)
    ang = np.mod(np.arctan2(W[:, 1, 0], W[:, 0, 0]), 2 * np.pi)
    assert stats.kstest(ang / (2 * np.pi), "uniform").pvalue > 1e-3
    # R diagonal positive: column 2 is column 1 rotated by +-90 degrees with either orientation
    det = W[:, 0, 0] * W[:, 1, 1] - W[:, 1, 0] * W[:, 0, 1]
    assert 0.45 < np.mean(det > 0) < 0.55


def test_quad_agrees_with_fp64_on_regular_orbits():
    """Regular orbits (K = 0.5, near the origin) agree between the FP64 oracle and the __float128 oracle to 1e-10 in
    lambda after 1e4 steps: the high-precision oracle is consistent with the FP64 restatement."""
    red = (QRSpectrum(MGS()),)
    ch, _ = make_chain(red)
    x0 = np.array([[1.0, 0.2], [0.5, -0.3], [2.5, 0.0]])  # inside the main island; (0.3, 0.1) lies in the chaotic layer
    W0 = O.frames(3, 1, 3, 2, 2)
    a = O.run(std(0.5), ch, x0, W0, 10000, 1, 2, nthreads=1)
    b = O.run(std(0.5), ch, x0, W0, 10000, 1, 2, nthreads=1, quad=True)
    fa, fb = finals(a, red, 2)[0], finals(b, red, 2)[0]
    assert np.max(np.abs(fa["lambda"] - fb["lambda"])) < 1e-10


def test_libm_and_portable_math_agree_on_regular_orbits():
    """The project's portable sin/cos/log and glibc give the same lambda on regular orbits to 1e-12."""
    red = (QRSpectrum(HouseholderQR()), )
    ch, _ = make_chain(red)
    x0 = np.array([[1.0, 0.2], [0.5, -0.3], [2.5, 0.0]])
    W0 = O.frames(3, 1, 3, 2, 2)
    a = O.run(std(0.5), ch, x0, W0, 10000, 1, 2, math=O.MATH_LIBM, nthreads=1)
    b = O.run(std(0.5), ch, x0, W0, 10000, 1, 2, math=O.MATH_CL, nthreads=1)
    fa, fb = finals(a, red, 2)[0], finals(b, red, 2)[0]
    assert np.max(np.abs(fa["lambda"] - fb["lambda"])) < 1e-12


def test_ofli_bounded_by_fli():
    """OFLI on a map (direction: the displacement F(x) - x): the orthogonal component is never longer than the
    column, so sin theta <= 1 and the OFLI basis maximum stays within the FLI log sums (test/ofli AC9-AC10 sandwich)."""
    red = (OFLI(replace=False), FLI())
    ch, _ = make_chain(red)
    x0 = np.array([[0.3, 0.1], [1.0, 2.2], [2.5, 4.0]])
    W0 = O.frames(3, 1, 3, 2, 2)
    out = O.run(std(1.2), ch, x0, W0, 3000, 1, 2, nthreads=1)
    fo, ff = finals(out, red, 2)
    assert np.all(fo["sinmin"] <= 1.0 + 4 * EPS)
    assert np.array_equal(fo["sum"], ff["sum"])
    assert np.all(fo["ofli"] <= ff["fli"] + 1e-12) or np.all(fo["ofli"] <= fo["fli"] + 1e-12)


def test_per_orbit_k_map_equals_fixed_k_runs():
    """The per-orbit-K standard map (state x, p, K) steps exactly as StandardMap(K) for each orbit's K."""
    from chaos_amd import StandardMap, StandardMapK
    from chaos_amd.inputs import torus_uniform
    from chaos_amd.tangent import make_chain
    from chaos_amd.reducers import MEGNO, QRSpectrum, MGS
    ch, _ = make_chain((MEGNO(replace=False), QRSpectrum(MGS())))
    Ks = np.array([0.5, 0.97, 2.0, 6.0])
    x2 = torus_uniform(5, 4)
    x3 = np.concatenate([x2, Ks[:, None]], axis=1)
    W0 = O.frames(5, 1, 4, 2, 2)
    a = O.run(StandardMapK().descriptor(), ch, x3, W0, 500, 1, 2)
    for i, K in enumerate(Ks):
        b = O.run(StandardMap(K).descriptor(), ch, x2[i:i + 1], W0[i:i + 1], 500, 1, 2)
        assert np.array_equal(a["states"][i], b["states"][0])
        assert np.array_equal(a["x"][i, :2], b["x"][0]) and a["x"][i, 2] == K


def test_readme_single_orbit_anchor():
    """README.md:19-21: K = 6, x0 = (0.1, 0.2), N = 1e5, k = 2, tau = 1, seed 1, MGS gives lambda = +-1.1348588476487835
    with Julia's libm and StableRNG frames. Bitwise agreement needs both (unpinnable here), so this is a statistical
    anchor: the finite-time exponent of a chaotic orbit at N = 1e5 scatters by about 3e-3 between orbit
    realisations; the pair sums to zero (symplectic pairing)."""
    ch, allr = make_chain((QRSpectrum(MGS()),))
    x0 = np.array([[0.1, 0.2]])
    W0 = O.frames(1, 1, 1, 2, 2)
    for math_mode in (O.MATH_CL, O.MATH_LIBM):
        r = O.run(L_std_map(6.0), ch, x0, W0, 100_000, 1, 2, math=math_mode)
        lam = finals_of(decode_states(r["states"], allr, 2, 2), allr, 2, 2)[0]["lambda"][0]
        assert abs(lam[0] - 1.1348588476487835) < 0.02
        assert abs(lam[0] + lam[1]) < 1e-12


def L_std_map(K):
    from chaos_amd import StandardMap
    return StandardMap(K).descriptor()


def test_same_orbit_quad_spectrum_matches_fp64():
    """Along one FP64 orbit, the QR spectrum with __float128 Jacobians, products and factorisation agrees with the
    FP64 spectrum to rounding (AC4, test/tangent-core/tc_ac01_04.jl:66-88): FP64 tangents lose nothing measurable
    over 5000 steps, for the chaotic K = 6 and the mixed K = 0.97."""
    from chaos_amd import StandardMap
    from chaos_amd.inputs import torus_uniform
    for alg, K in ((MGS(), 6.0), (HouseholderQR(), 0.97)):
        ch, allr = make_chain((QRSpectrum(alg),))
        x0 = torus_uniform(1, 16)
        W0 = O.frames(1, 1, 16, 2, 2)
        lq = O.same_orbit_quad(K, ch, x0, W0, 5000, 1, 2)
        r = O.run(StandardMap(K).descriptor(), ch, x0, W0, 5000, 1, 2, math=O.MATH_CL)
        l64 = finals_of(decode_states(r["states"], allr, 2, 2), allr, 2, 2)[0]["lambda"]
        assert np.max(np.abs(l64 - lq)) < 1e-12
        # and a wrong Jacobian (K perturbed by 1e-9) is visible
        lw = O.same_orbit_quad(K * (1 + 1e-9), ch, x0, W0, 5000, 1, 2)
        assert np.max(np.abs(l64 - lw)) > 1e-11
