"""ResidualMonitor in the oracle (src/Diagnostics/monitor.jl:44-54), pinned by what the mathematics fixes: the
one-step tangents of the standard map and of the Gauss method are symplectic with det 1, so both residuals stay at
rounding level; max|M| of the map is max(1 + K cos x, |K cos x|, 1) <= 1 + |K|; a tolerance of zero trips the
orthogonality check of a frame that is not exactly orthonormal. The bounds catch the plausible mistakes: with the
form transposed, M' Omega' M - Omega = -2 Omega for a symplectic M (residual 2), and a dropped or sign-flipped term
of the determinant moves it away from 1 by O(1)."""
import numpy as np

from chaos_amd import StandardMap, HenonHeilesGauss, QRSpectrum, MGS, GivensQR, ResidualMonitor
from chaos_amd.inputs import torus_uniform, hh_section_starts
from chaos_amd.tangent import make_chain, decode_states, finals_of
from tests import oracle_lib as O

EPS = 2.0 ** -52


def run_map(K, mon, N=400, k=2):
    ch, allr = make_chain((QRSpectrum(MGS()), mon) if k == 2 else (mon,))
    x0 = torus_uniform(4, 32)
    r = O.run(StandardMap(K).descriptor(), ch, x0, O.frames(4, 1, 32, 2, k), N, 1, k)
    return r, finals_of(decode_states(r["states"], allr, 2, k), allr, 2, k), allr


def test_map_residuals_are_rounding_level_and_M_max_bounded():
    K = 6.0
    r, fin, allr = run_map(K, ResidualMonitor(StandardMap(K), 64.0, 1e-12))
    f = fin[1]
    assert np.all(r["code"] == 0)
    assert np.all(f["det_max"] <= 4 * EPS * (1 + K)) and np.all(f["symp_max"] <= 4 * EPS * (1 + K) ** 2)
    # MGS loses orthogonality as eps * cond(W), and one step has cond(M) <= (1 + K)^2
    assert np.all(f["orth_max"] <= 4 * EPS * (1 + K) ** 2) and np.all(f["orth_max"] > 0)
    assert np.all(f["M_max"] <= 1 + K) and np.all(f["M_max"] >= 1)
    st = decode_states(r["states"], allr, 2, 2)[1]
    assert np.all(st["step"] == 400) and np.all(st["sum_max"] == 0) and np.all(st["G_max"] == 0)


def test_zero_orth_tolerance_trips_at_first_handoff():
    r, fin, allr = run_map(2.0, ResidualMonitor(StandardMap(2.0), 64.0, 0.0))
    st = decode_states(r["states"], allr, 2, 2)[1]
    tripped = st["code"] == 3
    assert tripped.sum() > 0
    assert np.all(r["code"][tripped] == 3) and np.all(r["hstep"][tripped] == 1)
    assert np.all(st["value"][tripped] > 0) and np.all(st["threshold"][tripped] == 0)


def test_k1_has_no_orthogonality_check():
    r, fin, allr = run_map(1.0, ResidualMonitor(StandardMap(1.0), 64.0, 0.0), k=1)
    assert np.all(r["code"] == 0) and np.all(fin[0]["orth_max"] == 0)


def test_gauss_step_jacobian_is_symplectic():
    hh = HenonHeilesGauss(h=0.01, tol=1e-15)
    ch, allr = make_chain((QRSpectrum(GivensQR()), ResidualMonitor(hh, 64.0, 1e-12)))
    x0 = hh_section_starts(2, 16, 1 / 8)
    r = O.run(hh.descriptor(), ch, x0, O.frames(2, 1, 16, 4, 4), 300, 10, 4)
    f = finals_of(decode_states(r["states"], allr, 4, 4), allr, 4, 4)[1]
    assert np.all(r["code"] == 0)
    assert np.all(f["det_max"] <= 64 * EPS * (1 + f["M_max"])) and np.all(f["det_max"] < 1e-13)
    assert np.all(f["symp_max"] < 1e-13) and np.all(f["M_max"] > 0.9)



# ------------------------------------------------ the reference's D34 checks (test/diagnostics/dg_d34.jl:16-103)
from chaos_amd import (Benettin, det_threshold, symp_threshold, sum_threshold, eps_sum, eps_pair, check_residual,
                       monitor_exponents, NumericalHealthError)
import math
import pytest

DET_TOL = 4.0  # the reference's det_residual_tol default (test/diagnostics/runtests.jl:6)


def test_thresholds_d34():
    """Powers of two make the thresholds exact (dg_d34.jl:16-37)."""
    M2 = np.array([[2.0, 1.0], [1.0, 1.0]])
    assert det_threshold(4.0, M2) == 4.0 * EPS * 3.0
    assert symp_threshold(4.0, M2) == 4.0 * EPS * 9.0
    assert symp_threshold(4.0, np.eye(4)) == 4.0 * EPS * 4 * 4
    assert sum_threshold(16.0, 5.0, 1, 2.0) == 16.0 * EPS * 7.0
    assert sum_threshold(16.0, 6.5, 2, 2.0) == 16.0 * EPS * (6.5 / 2 + 2)


def lin_run(delta, N):
    ch, allr = make_chain((Benettin(), ResidualMonitor(StandardMap(1.0), DET_TOL, math.inf)))
    A = np.array([[1.0 + delta, 0.0], [0.0, 1.0]])
    r = O.run_linear(A, ch, np.zeros((1, 2)), np.array([[[1.0], [0.0]]]), N, 1, 1)
    return r, decode_states(r["states"], allr, 2, 1)[1]


def test_det_raise_and_quiet_d34():
    """Tangent diag(1 + delta, 1) has det residual exactly delta: quiet for delta in (theta/4, theta/2] over 100
    handoffs, raised at the first handoff with value delta and threshold theta(delta) = 4 eps (2 + delta) for delta
    in [2 theta, 4 theta), theta = theta(8 eps) (dg_d34.jl:39-69)."""
    theta = lambda d: 4 * EPS * (2 + d)
    t0 = theta(8 * EPS)
    raise_ = [2.0 ** j for j in range(-130, 21) if 2 * t0 <= 2.0 ** j < 4 * t0]
    quiet = [2.0 ** j for j in range(-130, 21) if t0 / 4 < 2.0 ** j <= t0 / 2]
    assert raise_ and quiet
    for d in quiet:
        r, st = lin_run(d, 100)
        assert r["code"][0] == 0 and st["code"][0] == 0 and st["det_max"][0] == d
    for d in raise_:
        r, st = lin_run(d, 1)
        assert r["code"][0] == 3 and r["hstep"][0] == 1
        assert st["code"][0] == 1 and st["value"][0] == d and st["step"][0] == 1
        assert st["threshold"][0] == theta(d)  # exact: 4 eps is a power of two and 2 + delta is representable


def test_m_max_is_the_running_max_along_the_orbit_d34():
    """M_max equals max |DT| at the point after each window, recomputed by stepping the orbit (with the same glibc
    sin and cos as the oracle's libm mode); G_max is 0 in a plain run and k = 1 has no orthogonality
    (dg_d34.jl:71-92)."""
    K, N = 2.0, 100
    ch, allr = make_chain((Benettin(), ResidualMonitor(StandardMap(K), DET_TOL, math.inf)))
    r = O.run(StandardMap(K).descriptor(), ch, np.array([[1.0, 2.0]]), np.array([[[1.0], [0.0]]]), N, 1, 1,
              math=O.MATH_LIBM)
    fin = finals_of(decode_states(r["states"], allr, 2, 1), allr, 2, 1)[1]
    x, p, mref = 1.0, 2.0, 0.0
    for _ in range(N):
        p = p + K * math.sin(x)
        x = x + p
        c = K * math.cos(x)
        mref = max(mref, abs(1.0 + c), abs(c), 1.0)
    assert fin["M_max"][0] == mref and fin["G_max"][0] == 0 and fin["orth_max"][0] == 0 and fin["det_max"][0] >= 0


def test_monitor_exponents_running_maxima_d34():
    """monitor_exponents folds G and M_max in as running maxima; the end-of-run sum check passes
    (dg_d34.jl:93-102); a sum above its threshold and a recorded monitor failure both raise."""
    ch, allr = make_chain((Benettin(), ResidualMonitor(StandardMap(2.0), DET_TOL, math.inf)))
    r = O.run(StandardMap(2.0).descriptor(), ch, np.array([[1.0, 2.0]]), np.array([[[1.0], [0.0]]]), 100, 1, 1)
    st = decode_states(r["states"], allr, 2, 1)[1]
    lam = np.array([0.5, -0.5])
    st2 = monitor_exponents(st, lam, 6.8, 100, pairing_tol=16.0, tau=1, M_max=10.0, step=100)
    assert st2["G_max"][0] == 6.8 and st2["M_max"][0] == 10.0
    st3 = monitor_exponents(st2, lam, 1.0, 100, pairing_tol=16.0, tau=1, M_max=1.0, step=100)
    assert st3["G_max"][0] == 6.8 and st3["M_max"][0] == 10.0 and st3["sum_max"][0] == 0 and st3["pair_max"][0] == 0
    assert check_residual("sum", eps_sum(lam), sum_threshold(16.0, st3["G_max"][0], 1, st3["M_max"][0]), step=100,
                          orbit=1) is None
    with pytest.raises(NumericalHealthError) as e:
        monitor_exponents(st, np.array([0.5, -0.5 + 1e-12]), 1.0, 100, pairing_tol=16.0, tau=1, M_max=1.0, step=100,
                          orbit0=7)
    assert e.value.what == "sum" and e.value.orbit == 7 and e.value.threshold == 16.0 * EPS * (1.0 + st["M_max"][0])
    bad = st.copy()
    bad["code"], bad["value"], bad["threshold"], bad["step"] = 1, 3e-15, 1e-15, 40
    with pytest.raises(NumericalHealthError) as e:
        monitor_exponents(bad, lam, 1.0, 100, pairing_tol=16.0, tau=1, M_max=1.0, step=100)
    assert e.value.what == "det" and e.value.step == 40 and e.value.value == 3e-15


def test_eps_pair_and_sum_estimator_order():
    """Pairing in estimator order, no sorting (residuals.jl:16-18): lambda_i pairs with lambda_{M+1-i}, so a
    spectrum listed out of estimator order shows a residual though it sums to zero (dyadic values: exact)."""
    lam = np.array([[0.5, 0.25, -0.25, -0.5], [0.5, 0.25, 0.25, -0.5], [0.5, -0.5, 0.25, -0.25]])
    assert np.array_equal(eps_pair(lam), [0.0, 0.5, 0.25])
    assert np.array_equal(eps_sum(lam), [0.0, 0.5, 0.0])
