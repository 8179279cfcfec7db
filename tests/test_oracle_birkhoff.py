"""WeightedBirkhoff of the oracle (oracle/oracle.cpp, red_wbirk; src/Indicators/wbirkhoff_reducer.jl:4-116) and the
host functions (weighted_birkhoff, birkhoff_index), pinned by the reference's acceptance criteria
(test/weighted-birkhoff/wb_ac01_03.jl, wb_ac04_06.jl): constants and normalisation (AC1), the exact rotation (AC2),
the K = 0 standard map (AC3: trigonometric averages within 4 eps, the lifted momentum exactly), the reducer equal to
the pure function on the sampled series bit for bit (AC6), and the index's own rules; the portable exp against libm."""
import math

import mpmath as mp
import numpy as np
import pytest

from chaos_amd import (StandardMap, WrappedStandardMap, WeightedBirkhoff, CosObservable, SinObservable,
                       LiftedObservable, weighted_birkhoff, birkhoff_index, RunMonitor, torus_state)
from chaos_amd import _lib as L
from chaos_amd.birkhoff import weights as wb_weights
from chaos_amd.tangent import make_chain, decode_states, finals_of
from tests import oracle_lib as O

EPS = 2.0 ** -52
BITS = 256


def cp(p):
    return 4.0 ** p   # c_p of the reference suite (wb_helpers.jl:21)


def run_wb(red, sysm, x0, N, tau=1, extra_first=(), k=1, reset=False):
    chain, allr = make_chain(tuple(extra_first) + (red,), reset_frame=reset)
    W0 = np.tile(np.array([[1.0], [0.0]]) if k == 1 else np.eye(2), (len(x0), 1, 1))
    r = O.run(sysm.descriptor(), chain, x0, W0, N, tau, k)
    fin = finals_of(decode_states(r["states"], allr, 2, k), allr, 2, k)
    return r, fin[-1]


def test_portable_exp_within_one_ulp():
    """cl_exp (cl_math.h) against libm: at most 1 ulp on 3·10^5 arguments over the whole range, the subnormal
    results included; exp(0) = 1 exactly, overflow to inf, underflow to 0, NaN through."""
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(-745, 709.7, 200000), rng.uniform(-1, 1, 50000), rng.uniform(-745.1, -708, 50000),
                        -1.0 / np.linspace(1e-6, 0.25, 20000)])
    got, ref = O.exp(x), np.exp(x)
    assert np.max(np.abs(got.view(np.int64) - ref.view(np.int64))) <= 1
    sp = O.exp([0.0, 800.0, -800.0, np.nan, np.inf, -np.inf])
    assert sp[0] == 1.0 and sp[1] == np.inf and sp[2] == 0.0 and np.isnan(sp[3]) and sp[4] == np.inf and sp[5] == 0


def test_weights_library_equals_oracle_and_definition():
    """The library's host weights (the kernels' arithmetic) equal the oracle's bit for bit, w(0) = 0, the table is
    symmetric, and each weight agrees with exp(-1 / (t (1 - t))^p) at 256 bits (AC1d's definition) within the
    rounding of its argument carried through exp, (4 |x| + 4) eps relative for x = 1 / (t (1 - t))^p."""
    for N in (16, 100, 1000, 51200):
        for p in (1, 2, 3, 4):
            w = wb_weights(N, p)
            assert np.array_equal(w, O.wb_weights(N, p)) and w[0] == 0.0
            assert np.array_equal(w[1:], w[1:][::-1])
            if N <= 1000:
                with mp.workprec(BITS):
                    for jj in range(1, N, max(1, N // 97)):
                        t = mp.mpf(min(jj, N - jj)) / N
                        xa = 1 / (t * (1 - t)) ** p
                        ref = mp.e ** (-xa)
                        if ref > 2.0 ** -1000:
                            assert abs(mp.mpf(w[jj]) - ref) <= (4 * xa + 4) * EPS * ref, (N, p, jj)


@pytest.mark.parametrize("p", [1, 2])
def test_ac1d_normalisation(p):
    """AC1d: the weight total B of a horizon within c_p eps relative of the 256-bit sum of the definition; one B for
    both windows, so a constant series gives bitwise equal window averages."""
    for N in (16, 1000):
        ser = np.stack([np.full(2 * N, 0.7), np.sin(np.arange(1, 2 * N + 1, dtype=np.float64))], 1)
        r = weighted_birkhoff(ser, N, order=p)
        with mp.workprec(BITS):
            Bref = mp.fsum((lambda t: mp.e ** (-1 / (t * (1 - t)) ** p))(mp.mpf(min(jj, N - jj)) / N)
                           for jj in range(1, N))
        assert abs(mp.mpf(r["B"][0]) - Bref) <= cp(p) * EPS * Bref
        assert r["avg1"][0, 0] == r["avg2"][0, 0]


def test_ac1a_constants_pure_function():
    """AC1a (the pure function): a constant series c and 2c over windows N and 2N: both averages within 2 eps |c|,
    delta = 0 bitwise, below_floor set, zero clear, S = |c|, absdig = -log10(32 eps S) within 2 eps relative."""
    for c in (0.3, 1e3, 1e-3, math.pi):
        for N in (16, 1000):
            ser = np.stack([np.full(4 * N, c), np.full(4 * N, 2 * c)], 1)
            r = weighted_birkhoff(ser, (N, 2 * N))
            ix = birkhoff_index(r["avg1"], r["avg2"], r["S"])
            for k, cc in enumerate((c, 2 * c)):
                for h in range(2):
                    assert abs(r["avg1"][k, h] - cc) <= 2 * EPS * abs(cc) and abs(r["avg2"][k, h] - cc) <= 2 * EPS * abs(cc)
                    assert ix["delta"][k, h] == 0 and ix["below_floor"][k, h] and not ix["zero"][k, h]
                    assert r["S"][k, h] == abs(cc)
                    ref = -math.log10(32 * EPS * abs(cc))
                    assert abs(ix["absdig"][k, h] - ref) <= 2 * EPS * ref


def rotation(theta0, alpha, k, n):
    with mp.workprec(BITS):
        tp = 2 * mp.pi
        return (np.array([float(mp.cos(tp * k * (theta0 + j * alpha))) for j in range(n)]),
                np.array([float(mp.sin(tp * k * (theta0 + j * alpha))) for j in range(n)]))


@pytest.mark.parametrize("p", [1, 2])
def test_ac2_exact_rotation(p):
    """AC2: cos and sin of 2 pi theta_j, theta_j = theta_0 + j (sqrt 2 - 1), generated at 256 bits and rounded once
    (and cos 4 pi theta): both window averages within c_p eps (the limit is 0) and |delta| <= 2 c_p eps."""
    with mp.workprec(BITS):
        alpha = mp.sqrt(2) - 1
        th0s = (mp.mpf(1) / 7, mp.mpf("0.3"))
    for th0 in th0s:
        for k, Ns in ((1, (500, 1000, 4000)), (2, (1000, 4000))):
            for N in Ns:
                cs, sn = rotation(th0, alpha, k, 2 * N)
                for f in ((cs, sn) if k == 1 else (cs,)):
                    r = weighted_birkhoff(f, N, order=p)
                    ix = birkhoff_index(r["avg1"][0, 0], r["avg2"][0, 0], r["S"][0, 0])
                    assert abs(r["avg1"][0, 0]) <= cp(p) * EPS and abs(r["avg2"][0, 0]) <= cp(p) * EPS
                    assert abs(ix["delta"]) <= 2 * cp(p) * EPS


def test_ac3_k0_standard_map():
    """AC3: the wrapped map at K = 0 (q_j = q_0 + j p_0 mod 2 pi, p constant), p_0 in 1.0, 2.0, 3.3, 5.0, random
    q_0, N in 1000, 4000: the cos q and sin q averages within 4 eps and |delta| <= 8 eps; the lifted p average equals
    p_0 within 2 eps p_0; p_0 = pi: the average of cos 2q equals cos 2 q_0 within 16 eps."""
    sysm = WrappedStandardMap(0.0)
    rng = np.random.default_rng(25103)
    q0s = 2 * math.pi * rng.random(4)
    for N in (1000, 4000):
        red = WeightedBirkhoff(window=N, observables=(CosObservable(1), SinObservable(1), LiftedObservable(2, 4)))
        x0 = np.array([torus_state([q0, p0]) for p0 in (1.0, 2.0, 3.3, 5.0) for q0 in q0s])
        red.check_run(1, 2 * N, x0=x0[0])
        r, f = run_wb(red, sysm, x0, 2 * N)
        assert np.all(r["code"] == 0) and np.all(f["complete"])
        assert np.all(np.abs(f["avg1"][:, :2, 0]) <= 4 * EPS) and np.all(np.abs(f["avg2"][:, :2, 0]) <= 4 * EPS)
        assert np.all(np.abs(f["delta"][:, :2, 0]) <= 8 * EPS)
        pp = x0[:, 1]
        assert np.all(np.abs(f["avg1"][:, 2, 0] - pp) <= 2 * EPS * pp) and np.all(np.abs(f["avg2"][:, 2, 0] - pp) <= 2 * EPS * pp)
        red2 = WeightedBirkhoff(window=N, observables=(CosObservable(1, 2),))
        x1 = np.array([torus_state([q0, math.pi]) for q0 in q0s])
        r, f = run_wb(red2, sysm, x1, 2 * N)
        with mp.workprec(BITS):
            ref = np.array([float(mp.cos(2 * mp.mpf(q0))) for q0 in x1[:, 0]])
        assert np.all(np.abs(f["avg1"][:, 0, 0] - ref) <= 16 * EPS) and np.all(np.abs(f["delta"][:, 0, 0]) <= 8 * EPS)


def sampled_series(sysm, x0, nsamp, tau, obs):
    """The observables at the handoffs 1..nsamp of one orbit: its point after j tau steps from the oracle's runs,
    the values with the kernels' sin/cos (cl_math_host)."""
    lib = L.load()
    chain, _ = make_chain((), reset_frame=True)   # the deviation column restarts each handoff: no overflow stops the orbit
    pts = []
    for j in range(1, nsamp + 1):
        pts.append(O.run(sysm.descriptor(), chain, x0[None], np.array([[[1.0], [0.0]]]), j * tau, tau, 1)["x"][0])
    pts = np.array(pts)
    out = np.empty((nsamp, len(obs)))
    for k, o in enumerate(obs):
        kind, a, b = o.code()
        if kind == L.CL_OBS_LIFTED:
            out[:, k] = pts[:, a] + 2 * (3.141592653589793 + 1.2246467991473532e-16) * pts[:, b]
            continue
        arg = np.ascontiguousarray(float(b) * pts[:, a])
        sn, cs, lg = np.empty_like(arg), np.empty_like(arg), np.empty_like(arg)
        lib.cl_math_host(arg.size, arg.ctypes.data, sn.ctypes.data, cs.ctypes.data, lg.ctypes.data)
        out[:, k] = cs if kind == L.CL_OBS_COS else sn
    return out


@pytest.mark.parametrize("p,tau", [(1, 1), (2, 4)])
def test_ac6_reducer_equals_function(p, tau):
    """AC6: K = 4 observables, horizons 64, 256 and 400, K_std in 0.5, 2 pi, 10: the reducer's window averages, S and
    B are bit for bit those of the pure function on the series sampled at the handoffs; and the same for the single
    horizon with the weight vector given explicitly."""
    Ns = (64, 256, 400)
    obs = (CosObservable(1), SinObservable(1), CosObservable(2), SinObservable(2))
    for K in (0.5, 2 * math.pi, 10.0):
        sysm = WrappedStandardMap(K)
        x0 = np.array(torus_state([1.1, 2.3]))
        red = WeightedBirkhoff(window=Ns, observables=obs, birkhoff_order=p)
        red.check_run(tau, 800 * tau)
        r, f = run_wb(red, sysm, x0[None], 800 * tau, tau, extra_first=(RunMonitor(sysm),))
        assert r["code"][0] == 0 or K != 0.5
        if r["code"][0] != 0:   # the unnormalised deviation column overflowed on a chaotic orbit
            continue
        ser = sampled_series(sysm, x0, 800, tau, obs)
        g = weighted_birkhoff(ser, Ns, order=p)
        assert np.array_equal(f["avg1"][0], g["avg1"]) and np.array_equal(f["avg2"][0], g["avg2"])
        assert np.array_equal(f["S"][0], g["S"]) and np.array_equal(f["B"][0], g["B"])
        for h, N in enumerate(Ns):
            gw = weighted_birkhoff(ser[:, 0], N, order=p, weights_=wb_weights(N, p))
            assert gw["avg1"][0, 0] == g["avg1"][0, h] and gw["B"][0] == g["B"][h]


def test_frozen_horizons_and_failure():
    """Past 2 max(window) handoffs only the counters advance (nothing is evaluated: a non-finite point there does not
    fail the orbit); before it, a non-finite observable fails the orbit with its 1-based index and handoff step; a
    horizon not yet complete reads NaN (avg2 before 2N, avg1 before N)."""
    sysm = StandardMap(0.0)
    red = WeightedBirkhoff(window=(16, 32), observables=(LiftedObservable(1, 2), CosObservable(2)))
    # lifted x[1] + 2 pi x[2] on the lifted map is x + 2 pi p: finite; p = 1e308 overflows it at once
    r, f = run_wb(red, sysm, np.array([[0.5, 1e308]]), 10)
    chain, allr = make_chain((red,))
    st = decode_states(r["states"], allr, 2, 1)[0]
    assert r["code"][0] == 3 and st["code"][0] == 1 and st["n"][0] == 1 and st["handoffs"][0] == 0
    assert red.error(st[0])[0] == "observable" and red.error(st[0])[4] == 1
    r, f = run_wb(red, sysm, np.array([[0.5, 0.25]]), 70)
    r64, f64 = run_wb(red, sysm, np.array([[0.5, 0.25]]), 64)
    assert r["code"][0] == 0 and f["handoffs"][0] == 70 and f["complete"][0] and f["n"][0] == 70
    assert np.array_equal(f["avg1"], f64["avg1"]) and np.array_equal(f["avg2"], f64["avg2"])
    assert np.array_equal(f["S"], f64["S"]) and np.array_equal(f["B"], f64["B"])
    r2, f2 = run_wb(red, sysm, np.array([[0.5, 0.25]]), 20)
    assert np.isnan(f2["avg2"][0, 0, 0]) and not np.isnan(f2["avg1"][0, 0, 0]) and np.isnan(f2["avg1"][0, 0, 1])
    assert np.isnan(f2["delta"][0, 0, 1]) and not f2["below_floor"][0, 0, 1]


def test_index_rules():
    """birkhoff_index (wbirkhoff.jl:144-171): swapping the pair negates delta and keeps the digits; S = 0 gives NaN
    digits with zero and below_floor set; a denormal S floors phi at floatmin; an overflowing delta gives NaN digits
    and no flags; reldig is NaN where m <= phi and dig is then absdig."""
    a, b, S = 0.3, 0.3 + 1e-9, 1.0
    x, y = birkhoff_index(a, b, S), birkhoff_index(b, a, S)
    assert y["delta"] == -x["delta"] and x["absdig"] == y["absdig"] and x["reldig"] == y["reldig"] and x["dig"] == y["dig"]
    z = birkhoff_index(0.0, 0.0, 0.0)
    assert np.isnan(z["absdig"]) and z["zero"] and z["below_floor"]
    d = birkhoff_index(1e-310, 0.0, 1e-310)
    assert np.isfinite(d["absdig"]) and d["absdig"] == -math.log10(2.0 ** -1022)
    o = birkhoff_index(1.7e308, -1.7e308, 1.7e308)
    assert np.isnan(o["absdig"]) and not o["below_floor"] and not o["zero"]
    s = birkhoff_index(1e-20, -1e-20, 1.0)
    assert np.isnan(s["reldig"]) and s["dig"] == s["absdig"]
    with pytest.raises(ValueError):
        birkhoff_index(1.0, 1.0, 1.0, floor_multiplier=0)
    for bad in (dict(window=8, observables=(CosObservable(1),)), dict(window=16, observables=()),
                dict(window=16, observables=(CosObservable(1),), birkhoff_order=5),
                dict(window=16, observables=tuple(CosObservable(1) for _ in range(9)))):
        with pytest.raises(ValueError):
            WeightedBirkhoff(**bad)
    red = WeightedBirkhoff(window=(16, 64), observables=(CosObservable(1),))
    red.check_run(2, 256)
    with pytest.raises(ValueError):
        red.check_run(1, 100)
    with pytest.raises(ValueError):
        red.check_run(0, 128)


@pytest.mark.parametrize("K", [2 * math.pi, 10.0])
def test_window_mode_on_chaotic_orbits(K):
    """The window kinds' constant deviation column (reset after every handoff, run_indicator_device.jl:66-80): on
    chaotic orbits the column never overflows, the frame is e_1 after each handoff, RunMonitor's maxima are unchanged,
    and the reducer still equals the pure function on the sampled series bit for bit (AC6 on the chaotic classes)."""
    sysm = WrappedStandardMap(K)
    Ns = (250, 1000)
    obs = (CosObservable(1), SinObservable(1), CosObservable(2, 2), LiftedObservable(2, 4))
    red = WeightedBirkhoff(window=Ns, observables=obs, birkhoff_order=2)
    x0 = np.array(torus_state([0.7, 2.9]))
    r, f = run_wb(red, sysm, x0[None], 2000, 1, extra_first=(RunMonitor(sysm),), reset=True)
    assert r["code"][0] == 0 and np.array_equal(r["W"][0], [[1.0], [0.0]])
    r0, _ = run_wb(red, sysm, x0[None], 2000, 1, extra_first=(RunMonitor(sysm),))
    assert r0["code"][0] == 2   # without the reset the unnormalised column overflows
    assert np.array_equal(r["states"][0, :24], r0["states"][0, :24]) or r0["hstep"][0] < 2000
    ser = sampled_series(sysm, x0, 2000, 1, obs)
    g = weighted_birkhoff(ser, Ns, order=2)
    assert np.array_equal(f["avg1"][0], g["avg1"]) and np.array_equal(f["avg2"][0], g["avg2"])
    assert np.array_equal(f["S"][0], g["S"])
