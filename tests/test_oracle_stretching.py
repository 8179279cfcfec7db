"""StretchingNumbers of the oracle (oracle/oracle.cpp, red_stretch; src/Indicators/stretching.jl:34-147) pinned against
the reference's acceptance criteria (test/stretching/sn_ac01_03.jl): constant one-step matrices put every stretching
number of a column in one cell (AC2a); the cat map's counts equal those of its closed-form series cell by cell
(AC2c); the K = 0 shear's counts equal the closed form's and b = 0 gives exactly zero sums (AC3). The host helpers
(spectrum_cell, spectral_distance, check_run, check_frame, reducer_final) against their definitions, and the sums
against MEGNO's column log sums on the standard map, bit for bit."""
import math

import mpmath as mp
import numpy as np
import pytest

from chaos_amd import (StandardMap, WrappedStandardMap, MEGNO, StretchingNumbers, RunMonitor, stretching_range,
                       spectral_distance, spectrum_cell, torus_state, NumericalHealthError)
from chaos_amd.tangent import make_chain, decode_states, finals_of
from tests import oracle_lib as O

EPS = 2.0 ** -52
BITS = 256


def run(sr, A, x0, W0, N, tau, sysd=None, extra=()):
    chain, allr = make_chain((sr,) + tuple(extra))
    r = O.run_ws(sysd, chain, sr.workspace(2, 2), x0, W0, N, tau, 2, A=A)
    fin = finals_of(decode_states(r["states"], allr, 2, 2), allr, 2, 2, r["ws"])
    return r, fin


def cells_ref(zs, lo, hi, B):
    """Cells of mpmath values (sn_cell_ref, test/stretching/sn_helpers.jl:40-45)."""
    with mp.workprec(BITS):
        l, u = mp.mpf(lo), mp.mpf(hi)
        h = (u - l) / B
        out = []
        for z in zs:
            out.append(0 if z < l else (B + 1 if z > u else min(int(mp.floor((z - l) / h)) + 1, B)))
    return out


def hist_ref(zs, lo, hi, B):
    N = np.zeros(B + 2, dtype=np.int64)
    for c in cells_ref(zs, lo, hi, B):
        N[c] += 1
    return N


def edge_ok(zs, lo, hi, B, tol=1e-9, endslack=0.0):
    """The edge condition of the reference suite (sn_helpers.jl:100-122): every value at least tol from an interior
    cell edge (endslack from the two end edges)."""
    with mp.workprec(BITS):
        l, u = mp.mpf(lo), mp.mpf(hi)
        h = (u - l) / B
        for z in zs:
            if z < l or z > u:
                if abs(z - l) < tol or abs(z - u) < tol:
                    return False
                continue
            f = (z - l) / h
            c = int(mp.nint(f))
            dist = abs(f - c) * h
            if (c in (0, B) and dist < max(endslack, 0)) or (0 < c < B and dist < tol):
                return False
    return True


def pick_B(B0, lo, hi, *zl, endslack=0.0):
    for B in range(B0, B0 + 61):
        if all(edge_ok(z, lo, hi, B, endslack=endslack) for z in zl):
            return B
    raise AssertionError("no B satisfies the edge condition")


def D_ref(N1, N2, lo, hi, B):
    k = int(N1.sum())
    with mp.workprec(BITS):
        h = (mp.mpf(lo) - mp.mpf(hi)) / -B
        return mp.mpf(0) if k == 0 else mp.sqrt(mp.mpf(int(((N1 - N2) ** 2).sum()))) / (k * mp.sqrt(h))


def check_counts(fin, zs1, zs2, k, lo, hi, B):
    N1, N2 = hist_ref(zs1[:k], lo, hi, B), hist_ref(zs2[:k], lo, hi, B)
    c = fin["counts"][0]
    assert np.array_equal(c[:, 0], N1) and np.array_equal(c[:, 1], N2)
    Dr = D_ref(N1, N2, lo, hi, B)
    D = fin["D"][0]
    assert (D == 0) if Dr == 0 else abs(mp.mpf(D) - Dr) <= 4 * EPS * Dr


@pytest.mark.parametrize("tau", [1, 5])
def test_ac2a_constant_matrices(tau):
    """AC2a: diag(l1, l2), l in exp(1/2), exp(-1/2), exp(3/2) formed in float64, columns (1, 0) and (0, 1), range
    [-2, 2] with B = 1001: each count vector has the single entry k in the cell of ln|l|; D = sqrt(2 / h) when the
    cells differ and 0 when they coincide."""
    B, lo, hi = 1001, -2.0, 2.0
    ls = [math.exp(0.5), math.exp(-0.5), math.exp(1.5)]
    for l1 in ls:
        for l2 in ls:
            A = np.array([[l1, 0.0], [0.0, l2]])
            with mp.workprec(BITS):
                z1, z2 = mp.log(mp.mpf(l1)), mp.log(mp.mpf(l2))
            for k in (1, 10, 1000):
                sr = StretchingNumbers(range=(lo, hi), bins=B, interval=tau)
                r, fin = run(sr, A, np.zeros((1, 2)), np.eye(2)[None], k * tau, tau)
                assert r["code"][0] == 0
                check_counts(fin[0], [z1] * k, [z2] * k, k, lo, hi, B)
                c = fin[0]["counts"][0]
                assert np.count_nonzero(c[:, 0]) == 1 and c[:, 0].max() == k
                h = (hi - lo) / B
                if l1 == l2:
                    assert fin[0]["D"][0] == 0
                else:
                    assert abs(fin[0]["D"][0] - math.sqrt(2 / h)) <= 4 * EPS * math.sqrt(2 / h)


def y5_z(w, tau, K):
    """The cat map's stretching numbers of a column w (sn_y5_z, sn_helpers.jl:342-353): with w = c1 e1 + c2 e2 on the
    eigenvectors, z_k = ln mu + ln((1 + r mu^(-4 k tau)) / (1 + r mu^(-4 (k - 1) tau))) / (2 tau), r = (c2 / c1)^2."""
    with mp.workprec(BITS):
        mu = (3 + mp.sqrt(5)) / 2
        t = mu - 2
        nrm = mp.sqrt(1 + t * t)
        e1, e2 = (1 / nrm, t / nrm), (-t / nrm, 1 / nrm)
        wb = (mp.mpf(w[0]), mp.mpf(w[1]))
        c1, c2 = wb[0] * e1[0] + wb[1] * e1[1], wb[0] * e2[0] + wb[1] * e2[1]
        r = (c2 / c1) ** 2
        return [mp.log(mu) + mp.log((1 + r * mu ** (-4 * k * tau)) / (1 + r * mu ** (-4 * (k - 1) * tau))) / (2 * tau)
                for k in range(1, K + 1)]


def y5_range():
    """stretching_range of a 2 x 2 LinearSymplecticMap (stretching.jl:243-251): ln of the largest singular value,
    padded by 8 eps relative."""
    with mp.workprec(BITS):
        F, det = mp.mpf(4 + 1 + 1 + 1), mp.mpf(1)
        L = mp.log(mp.sqrt((F + mp.sqrt(F * F - 4 * det * det)) / 2))
        hi = float(L * (1 + 8 * mp.mpf(2) ** -52))
    return -hi, hi


@pytest.mark.parametrize("c1", [0.05, 0.6, 0.99])
@pytest.mark.parametrize("tau", [1, 2, 5])
def test_ac2c_cat_map_closed_form(c1, tau):
    """AC2c: the cat map [[2, 1], [1, 1]] with the column pair c1 e1 + c2 e2 and its quarter turn, the derived range,
    k up to 1000: the counts equal those of the closed-form series in every cell, no overflow counts, D within 4 eps
    relative of the exact formula."""
    with mp.workprec(BITS):
        mu = (3 + mp.sqrt(5)) / 2
        t = mu - 2
        nrm = mp.sqrt(1 + t * t)
        e1, e2 = (1 / nrm, t / nrm), (-t / nrm, 1 / nrm)
        c2 = mp.sqrt(1 - mp.mpf(c1) ** 2)
        w1 = (mp.mpf(c1) * e1[0] + c2 * e2[0], mp.mpf(c1) * e1[1] + c2 * e2[1])
        w2 = (-c2 * e1[0] + mp.mpf(c1) * e2[0], -c2 * e1[1] + mp.mpf(c1) * e2[1])
    W0 = np.array([[float(w1[0]), float(w2[0])], [float(w1[1]), float(w2[1])]])
    lo, hi = y5_range()
    zs1, zs2 = y5_z(W0[:, 0], tau, 1000), y5_z(W0[:, 1], tau, 1000)
    B = pick_B(101, lo, hi, zs1, zs2, endslack=2 * EPS * max(1.0, abs(hi)))
    for k in (1, 10, 100, 1000):
        sr = StretchingNumbers(range=(lo, hi), bins=B, interval=tau)
        r, fin = run(sr, np.array([[2.0, 1.0], [1.0, 1.0]]), np.zeros((1, 2)), W0[None], k * tau, tau)
        assert r["code"][0] == 0
        check_counts(fin[0], zs1, zs2, k, lo, hi, B)
        assert np.all(fin[0]["below"] == 0) and np.all(fin[0]["above"] == 0)


def shear_z(a, b, tau, K):
    """K = 0 shear [[1, 1], [0, 1]]: z_k = ln(|(a + k tau b, b)| / |(a + (k - 1) tau b, b)|) / tau."""
    with mp.workprec(BITS):
        a, b = mp.mpf(a), mp.mpf(b)
        return [(mp.log((a + k * tau * b) ** 2 + b * b) - mp.log((a + (k - 1) * tau * b) ** 2 + b * b)) / (2 * tau)
                for k in range(1, K + 1)]


@pytest.mark.parametrize("ab", [(0.6, 0.8), (-0.6, 0.8), (1.0, 0.0), (0.0, 1.0)])
@pytest.mark.parametrize("tau", [1, 3])
def test_ac3_shear(ab, tau):
    """AC3: the wrapped standard map at K = 0 (its one-step tangent is [[1, 1], [0, 1]] at every point), columns (a, b)
    and (-b, a), the derived range: counts equal the closed form's cell by cell at k = 10, 1000, 3000; b = 0 keeps
    column 1 at (1, 0), every z exactly 0, all counts in the central cell."""
    a, b = ab
    sysm = WrappedStandardMap(0.0)
    lo, hi = stretching_range(sysm)
    kmax = 3000
    zs1, zs2 = shear_z(a, b, tau, kmax), shear_z(-b, a, tau, kmax)
    B = pick_B(1001, lo, hi, zs1, zs2, endslack=2 * EPS)
    x0 = np.array([torus_state([0.3, 0.7])])
    W0 = np.array([[[a, -b], [b, a]]])
    for k in (10, 1000, kmax):
        sr = StretchingNumbers(range=(lo, hi), bins=B, interval=tau)
        r, fin = run(sr, None, x0, W0, k * tau, tau, sysd=sysm.descriptor())
        assert r["code"][0] == 0
        check_counts(fin[0], zs1, zs2, k, lo, hi, B)
        f = fin[0]
        if b == 0:
            assert f["sum"][0, 0] == 0 and f["comp"][0, 0] == 0 and f["sum2"][0, 0] == 0 and f["mean"][0, 0] == 0
            assert f["counts"][0][(B + 1) // 2, 0] == k
        if a == 0:
            assert f["sum"][0, 1] == 0 and f["comp"][0, 1] == 0 and f["counts"][0][(B + 1) // 2, 1] == k


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
def test_sums_are_megnos_and_counts_total(wrapped):
    """On the standard map at K = 6 the log sums of StretchingNumbers equal MEGNO's column sums bit for bit (the same
    column logs), every column's counts total the handoffs, 0 <= D <= sqrt(2 / h), mean = total(sum) / n, and with
    RunMonitor after it the chain is unchanged (the monitor only reads)."""
    sysm = (WrappedStandardMap if wrapped else StandardMap)(6.0)
    rng = np.random.default_rng(5)
    z = 2 * math.pi * rng.random((40, 2))
    x0 = np.array([torus_state(v) for v in z]) if wrapped else z
    W0 = O.frames(11, 1, 40, 2, 2)
    lo, hi = stretching_range(sysm)
    N, tau = 600, 3
    sr = StretchingNumbers(range=(lo, hi), bins=200, interval=3)
    r, fin = run(sr, None, x0, W0, N, tau, sysd=sysm.descriptor())
    chain, allr = make_chain((MEGNO(),))
    rm = O.run(sysm.descriptor(), chain, x0, W0, N, tau, 2)
    meg = finals_of(decode_states(rm["states"], allr, 2, 2), allr, 2, 2)[0]
    f = fin[0]
    assert np.all(r["code"] == 0)
    assert np.array_equal(f["sum"], meg["sum"]) and np.array_equal(f["comp"], meg["comp"])
    assert np.array_equal(r["W"], rm["W"]) and np.array_equal(r["x"], rm["x"])
    assert np.all(f["counts"].sum(axis=1) == N // tau)
    h = (hi - lo) / 200
    assert np.all((f["D"] >= 0) & (f["D"] <= math.sqrt(2 / h) * (1 + 4 * EPS)))
    assert np.array_equal(f["mean"], (f["sum"] + f["comp"]) / N)
    r2, fin2 = run(sr, None, x0, W0, N, tau, sysd=sysm.descriptor(), extra=(RunMonitor(sysm),))
    assert np.array_equal(r2["ws"], r["ws"]) and np.array_equal(fin2[0]["sum"], f["sum"])


def test_health_and_replace():
    """A zero column fails at the first handoff ("deviation matrix", the column, value 0) with the counts
    untouched. replace = false is the position before a later replacer: (StretchingNumbers(replace=false), MEGNO())
    counts and sums exactly what StretchingNumbers() alone does, and alone it leaves the frame unnormalised."""
    sysm = StandardMap(2.0)
    sr = StretchingNumbers(range=stretching_range(sysm), bins=50)
    W0 = np.array([[[1.0, 0.0], [0.0, 0.0]]])
    r, fin = run(sr, None, [[0.3, 0.4]], W0, 10, 1, sysd=sysm.descriptor())
    chain, allr = make_chain((sr,))
    st = decode_states(r["states"], allr, 2, 2)[0]
    assert r["code"][0] == 3 and r["hstep"][0] == 1 and st["code"][0] == 1 and st["column"][0] == 2
    assert sr.error(st[0])[:2] == ("deviation matrix", 0.0) and np.all(r["ws"] == 0)
    W1 = O.frames(2, 1, 1, 2, 2)
    a, fa = run(sr, None, [[0.3, 0.4]], W1, 10, 1, sysd=sysm.descriptor())
    keep = StretchingNumbers(range=stretching_range(sysm), bins=50, replace=False)
    b, fb = run(keep, None, [[0.3, 0.4]], W1, 10, 1, sysd=sysm.descriptor(), extra=(MEGNO(),))
    assert np.array_equal(a["ws"], b["ws"]) and np.array_equal(fa[0]["sum"], fb[0]["sum"])
    assert np.array_equal(a["W"], b["W"])
    c, fc = run(keep, None, [[0.3, 0.4]], W1, 10, 1, sysd=sysm.descriptor())
    assert np.all(np.abs(np.linalg.norm(a["W"][0], axis=0) - 1) < 1e-15)
    assert np.all(np.abs(np.linalg.norm(c["W"][0], axis=0) - 1) > 0.5)


def test_host_helpers():
    """spectrum_cell: z = lo is cell 1, z = hi is cell B (the last cell is closed), overflow cells 0 and B + 1;
    spectral_distance: symmetric, 0 on equal counts, sqrt(2 / h) for two single cells, errors on unequal totals;
    check_run and check_frame as the reference's."""
    assert spectrum_cell(-1.0, -1.0, 1.0, 10) == 1 and spectrum_cell(1.0, -1.0, 1.0, 10) == 10
    assert spectrum_cell(-1.5, -1.0, 1.0, 10) == 0 and spectrum_cell(1.5, -1.0, 1.0, 10) == 11
    assert spectrum_cell(0.0, -1.0, 1.0, 10) == 6
    a, b = np.array([0, 3, 0, 0]), np.array([0, 0, 3, 0])
    h = 0.5
    assert spectral_distance(a, b, h) == spectral_distance(b, a, h) == math.sqrt(18) / (3 * math.sqrt(h))
    assert spectral_distance(a, a, h) == 0 and spectral_distance(np.zeros(4, int), np.zeros(4, int), h) == 0
    with pytest.raises(ValueError):
        spectral_distance(a, np.array([0, 1, 0, 0]), h)
    sr = StretchingNumbers(range=(-1, 1), bins=10, interval=6)
    sr.check_run(3, 600)
    with pytest.raises(ValueError):
        sr.check_run(4, 600)
    with pytest.raises(ValueError):
        sr.check_run(1, 2 * 10 ** 9)
    assert sr.check_frame(np.eye(2)) == 0.0
    with pytest.raises(NumericalHealthError):
        sr.check_frame(np.array([[1.0, 1.0], [0.0, 0.0]]))
    with pytest.raises(NumericalHealthError):
        sr.check_frame(np.array([[2.0, 0.0], [0.0, 1.0]]))
    for bad in (dict(range=(1, 1)), dict(range=(0, np.inf)), dict(range=(-1, 1), bins=0),
                dict(range=(-1, 1), interval=0), dict(range=(-1, 1), min_per_cell=-1)):
        with pytest.raises(ValueError):
            StretchingNumbers(**bad)
    lo, hi = stretching_range(StandardMap(0.0))
    assert hi == -lo and abs(hi - math.log((1 + math.sqrt(5)) / 2)) <= 16 * EPS
