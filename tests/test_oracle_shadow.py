"""The two-trajectory estimator of the oracle (oracle/oracle.cpp, or_shadow; src/Lyapunov/two_trajectory.jl:45-82)
pinned against the reference's own acceptance criteria (test/lyap-two-trajectory/l2_analytic.jl, l2_runtime.jl):
closed forms on the linear cat map, the K = 0 shear and the hyperbolic fixed point at the origin, the minimal image
across the torus seam, the reference orbit untouched by the shadow, the start and health codes, segmented runs
equal to direct ones, and the wrapped and lifted modes against Benettin within the reference's path bound. The
host-side contract checks of `propagate_shadow` (argument errors before any device work) run here too."""
import math

import mpmath as mp
import numpy as np
import pytest

from chaos_amd import StandardMap, WrappedStandardMap, Benettin, torus_state, propagate_shadow
from chaos_amd.recorder import record_grid
from chaos_amd.tangent import make_chain, decode_states, finals_of
from chaos_amd.two_trajectory import unit_directions, check_d0
from tests import oracle_lib as O

EPS = 2.0 ** -52
TWO_PI = 2 * math.pi


def BS(n, Gfirst, K=0.0):
    """B_S(n) of the reference's suite (l2_helpers.jl:23): 8 eps (n + G_first) + 16 |K| n^2 eps."""
    return 8 * EPS * (n + Gfirst) + 16 * abs(K) * n * n * EPS


def sigma(K):
    S2 = 2 * K * K + 2 * abs(K) + 5
    return (math.sqrt(S2) + math.sqrt(S2 - 4)) / 2


def Enl(K, d0):
    return abs(K) * sigma(K) * d0 / math.sqrt(2)


def Ero(K, d0, M):
    return (abs(K) + 4) * M * EPS * sigma(K) / d0


def nX2(e, j=0):
    """n X2 as sum + comp of the end record (l2_nX2)."""
    return mp.mpf(e["sum"][j]) + mp.mpf(e["comp"][j])


def fib(n):
    a, b = 0, 1
    for _ in range(n):
        a, b = b, a + b
    return a


CAT = np.array([[2.0, 1.0], [1.0, 1.0]])


@pytest.mark.parametrize("tau", [1, 5, 10])
def test_ac1_linear_map_at_origin(tau):
    """AC1: cat map at the origin, u0 = (1, 0), d0 = 2^-20: |n X2 - ln ||A^n e1||| <= B_S(n) with
    A^n e1 = (F_{2n+1}, F_{2n}) and G_first = ||A^tau||_F^2."""
    for n in (20, 100, 1000):
        e = O.shadow(None, np.zeros((1, 2)), [[1.0, 0.0]], n, tau, 2.0 ** -20, A=CAT)
        assert e["code"][0] == 0 and e["step"][0] == n
        with mp.workprec(256):
            want = mp.log(mp.sqrt(mp.mpf(fib(2 * n + 1)) ** 2 + mp.mpf(fib(2 * n)) ** 2))
            err = float(abs(nX2(e) - want))
        G = float(fib(2 * tau + 1) ** 2 + 2 * fib(2 * tau) ** 2 + fib(2 * tau - 1) ** 2)
        assert err <= BS(n, G), (n, err)


@pytest.mark.parametrize("tau", [1, 10])
@pytest.mark.parametrize("d0", [2.0 ** -20, 1e-9])
def test_ac2_shear_at_origin(d0, tau):
    """AC2: StandardMap(0) at the origin (x' = x + p, p' = p): |n X2 - ln(||(a + n b, b)|| / ||(a, b)||)| <= B_S(n),
    G_first = 2 + tau^2, 20 directions."""
    rng = np.random.default_rng(25102)
    th = TWO_PI * rng.random(20)
    U0 = np.stack([np.cos(th), np.sin(th)], axis=1)
    u = unit_directions(U0, 20)
    for n in (10, 100, 10000):
        e = O.shadow(StandardMap(0.0).descriptor(), np.zeros((20, 2)), u, n, tau, d0)
        assert np.all(e["code"] == 0)
        with mp.workprec(256):
            for j in range(20):
                a, b = mp.mpf(U0[j, 0]), mp.mpf(U0[j, 1])
                want = mp.log(mp.sqrt((a + n * b) ** 2 + b ** 2) / mp.sqrt(a * a + b * b))
                assert float(abs(nX2(e, j) - want)) <= BS(n, 2 + tau * tau), (n, j)


@pytest.mark.parametrize("K", [1.0, 2.0, 6.0])
def test_ac3_hyperbolic_fixed_point(K):
    """AC3: StandardMap(K) at the fixed point (0, 0), 100 directions: |n X2 - (n ln mu + ln(|c1| ||v_u||))| <=
    B_S(n) + 2 n E_nl(d0) with mu = (2 + K + sqrt(K (K + 4))) / 2 and c1 the unstable component of u0."""
    rng = np.random.default_rng(25153)
    th = TWO_PI * rng.random(100)
    U0 = np.stack([np.cos(th), np.sin(th)], axis=1)
    u = unit_directions(U0, 100)
    G = (1 + K) ** 2 + 1 + K * K + 1
    for d0 in (1e-6, 1e-9):
        for n in (20, 100, 1000):
            e = O.shadow(StandardMap(K).descriptor(), np.zeros((100, 2)), u, n, 1, d0)
            assert np.all(e["code"] == 0)
            bound = BS(n, G) + 2 * n * Enl(K, d0)
            with mp.workprec(256):
                Kb = mp.mpf(K)
                s = mp.sqrt(Kb * (Kb + 4))
                mu = (2 + Kb + s) / 2
                vu = mp.sqrt(1 + ((s - Kb) / 2) ** 2)
                for j in range(100):
                    a, b = mp.mpf(U0[j, 0]), mp.mpf(U0[j, 1])
                    want = n * mp.log(mu) + mp.log(abs(a * (Kb + s) / 2 + b) / s * vu / mp.sqrt(a * a + b * b))
                    assert float(abs(nX2(e, j) - want)) <= bound, (d0, n, j)


def test_ac4_minimal_image():
    """AC4: wrapped K = 0, p = 0, d0 = 2^-20: the pair straddling x = 0 reads X2 = 0 like the pair at x = 1 (within
    2 E_ro); after one handoff the plain difference of the straddling pair is ~2 pi while the minimal image has norm
    d0; the mirror pair across p = 0 with u0 = (0, 1) follows ln sqrt(n^2 + 1)."""
    d0 = 2.0 ** -20
    sysd = WrappedStandardMap(0.0).descriptor()
    E = Ero(0.0, d0, TWO_PI)
    strad, plain = np.array(torus_state([-d0 / 2, 0.0])), np.array(torus_state([1.0, 0.0]))
    assert strad[0] > 6.28 and plain[0] == 1.0
    ux = np.array([[1.0, 0.0], [1.0, 0.0]])
    for n in (1, 10, 100):
        e = O.shadow(sysd, np.stack([strad, plain]), ux, n, 1, d0)
        X2 = (e["sum"] + e["comp"]) / n
        assert np.all(e["code"] == 0) and np.all(np.abs(X2) <= 2 * E) and abs(X2[0] - X2[1]) <= 2 * E
    e = O.shadow(sysd, strad[None], ux[:1], 1, 1, d0)
    xr, xs = e["x"][0], e["y"][0]
    pl = xs[:2] - xr[:2]
    assert math.log(math.hypot(*pl) / d0) >= 10
    tp = 2 * (3.141592653589793 + 1.2246467991473532e-16)
    sep = np.array([v - tp if v > tp / 2 else (v + tp if v <= -tp / 2 else v) for v in pl])
    assert np.all(np.abs(sep) <= math.pi) and abs(math.hypot(*sep) - d0) <= 4 * E * d0
    # the placement adds its reduction counts to the reference's: the lifted separation is the minimal image
    lifted = (xs[:2] + tp * xs[2:]) - (xr[:2] + tp * xr[2:])
    assert np.max(np.abs(lifted - sep)) <= 4 * EPS * 2 * tp
    mirror = np.array(torus_state([1.0, -d0 / 2]))
    for n in (1, 10, 100):
        e = O.shadow(sysd, mirror[None], [[0.0, 1.0]], n, 1, d0)
        with mp.workprec(256):
            err = float(abs(nX2(e) - mp.log(mp.sqrt(mp.mpf(n) ** 2 + 1))))
        assert err <= BS(n, 3) + 2 * n * E


def _snorm(v):
    m = max(abs(v[0]), abs(v[1]))
    return m * math.sqrt((v[0] / m) * (v[0] / m) + (v[1] / m) * (v[1] / m))


@pytest.mark.parametrize("tau", [1, 3])
def test_two_handoffs_by_hand(tau):
    """Two handoffs of the linear map A = [[2, 1], [1, 1]] away from the origin, worked in plain float64: the sum is
    log(d_1 / d_0+) + log(d_2 / d_1+) with d_n+ the realized separation after the placement x + delta d0 / d (not d0
    itself), dprev is d_2+ and the shadow is the placed point (1 ulp on each log: libm against numpy)."""
    x0, u0, d0 = np.array([0.7, -0.3]), np.array([0.6, 0.8]), 1e-7

    def fwd(z):
        return np.array([2.0 * z[0] + 1.0 * z[1], 1.0 * z[0] + 1.0 * z[1]])

    x, y = x0.copy(), x0 + d0 * u0
    dp = _snorm(y - x)
    logs = []
    for _ in range(2):
        for _ in range(tau):
            x, y = fwd(x), fwd(y)
        delta = y - x
        d = _snorm(delta)
        logs.append(math.log(d / dp))
        y = x + delta * (d0 / d)
        dp = _snorm(y - x)
    e = O.shadow(None, x0[None], u0[None], 2 * tau, tau, d0, A=CAT, math=O.MATH_LIBM)
    assert np.array_equal(e["x"][0], x) and np.array_equal(e["y"][0], y) and e["dprev"][0] == dp
    t = e["sum"][0] + e["comp"][0]
    assert abs(t - (logs[0] + logs[1])) <= 4 * EPS * (abs(logs[0]) + abs(logs[1]))


def starts(m, seed):
    rng = np.random.default_rng(seed)
    return TWO_PI * rng.random((m, 2))


@pytest.mark.parametrize("mode,K,tau", [("lifted", 0.5, 1), ("wrapped", TWO_PI, 4), ("lifted", TWO_PI, 4)])
def test_ac5_identical_reference_orbit(mode, K, tau):
    """AC5: the reference orbit of the two-trajectory run is the plain orbit, bit for bit the orbit of a Benettin run
    of the same start (N in 4, 8, 40, 400, 1000, 100 starts)."""
    wrapped = mode == "wrapped"
    sysm = (WrappedStandardMap if wrapped else StandardMap)(K)
    x0 = starts(100, 25250)
    if wrapped:
        x0 = np.array([torus_state(z) for z in x0])
    u = O.frames(25250, 1, 100, 2, 1)[:, :, 0]
    chain, _ = make_chain((Benettin(),))
    for N in (4, 8, 40, 400, 1000):
        e = O.shadow(sysm.descriptor(), x0, u, N, tau, 1e-8)
        l1 = O.run(sysm.descriptor(), chain, x0, u[:, :, None], N, tau, 1)
        assert np.array_equal(e["x"].view(np.uint64), l1["x"].view(np.uint64))


def test_ac10_start_and_health_codes():
    """AC10: d0 = 1e-20 at a start of magnitude ~1 is not resolved (code 2, step 0, value 0); a non-finite start
    reads a NaN separation (code 2); a lifted orbit at |p| ~ 1e308 loses the separation at its first handoff (code 2,
    value 0); a wrapped orbit whose winding count reaches 2^53 turns NaN (code 1, "state", with the first non-finite
    value); a wrapped d0 beyond pi realizes the shorter image (code 3, value |d0 - 2 pi|)."""
    sysd = StandardMap(2.0).descriptor()
    e = O.shadow(sysd, [[1.0, 2.0], [1.0, 2.0]], [[1.0, 0.0], [0.6, 0.8]], 4, 1, 1e-20)
    assert np.all(e["code"] == 2) and np.all(e["step"] == 0) and np.all(e["value"] == 0.0)
    e = O.shadow(sysd, [[0.3, np.nan], [np.nan, 0.3]], [[0.6, 0.8], [0.6, 0.8]], 4, 1, 1e-8)
    assert np.all(e["code"] == 2) and np.all(e["step"] == 0) and np.all(np.isnan(e["value"]))
    e = O.shadow(sysd, [[0.4, 1e308]], [[0.6, 0.8]], 10, 1, 1e-8)
    assert e["code"][0] == 2 and e["step"][0] == 1 and e["value"][0] == 0.0
    e = O.shadow(WrappedStandardMap(1.0).descriptor(), [[1.0, 1.0, 9007199254740990.0, 5.0]], [[0.6, 0.8]], 10, 1,
                 1e-8)
    assert e["code"][0] == 1 and e["step"][0] == 1 and np.isnan(e["value"][0]) and np.all(np.isnan(e["x"][0]))
    e = O.shadow(WrappedStandardMap(1.0).descriptor(), [np.array(torus_state([1.0, 1.0]))], [[1.0, 0.0]], 4, 1, 5.0)
    tp = 2 * (3.141592653589793 + 1.2246467991473532e-16)
    assert e["code"][0] == 3 and e["step"][0] == 0 and e["value"][0] == abs(5.0 - tp)


def test_ac10_lost_resolution_mid_run():
    """A lifted orbit with a large momentum outgrows the resolution of d0: the realized separation leaves
    [d0/2, 2 d0] (code 3) or collapses to zero (code 2) at a handoff after the start, and the record keeps the
    step, the offending separation and the last good sum."""
    sysd = StandardMap(1.0).descriptor()
    x0 = np.array([[0.5, p] for p in (3e7, 1e8, 3e8)])
    u = unit_directions([[0.6, 0.8]] * 3, 3)
    e = O.shadow(sysd, x0, u, 1000, 1, 1e-8)
    assert np.all(np.isin(e["code"], (2, 3))) and np.all(e["step"] > 0) and np.all(e["step"] < 1000)
    for j in range(3):
        v = e["value"][j]
        assert (v == 0.0) if e["code"][j] == 2 else (v > 0 and not (0.5e-8 <= v <= 2e-8))


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
def test_segmented_runs_equal_direct(wrapped):
    """Launches that carry the end record (D44) give the direct run's record bit for bit, for any split; an orbit
    that failed in an earlier launch keeps its record."""
    sysm = (WrappedStandardMap if wrapped else StandardMap)(3.0)
    x0 = starts(40, 7)
    x0[3] = [0.4, 1e308]
    if wrapped:
        x0 = np.array([torus_state(z) if np.all(np.isfinite(z)) else [z[0], z[1], 0, 0] for z in x0])
    u = O.frames(7, 1, 40, 2, 1)[:, :, 0]
    N, tau = 600, 3
    direct = O.shadow(sysm.descriptor(), x0, u, N, tau, 1e-7)
    for L in (3, 51, 300):
        e, n0 = None, 0
        while n0 < N:
            st = min(L, N - n0)
            e = O.shadow(sysm.descriptor(), x0, u, N, tau, 1e-7, n0=n0, steps=st,
                         ends=None if e is None else e["raw"])
            n0 += st
        assert np.array_equal(e["raw"], direct["raw"]), L
    assert direct["code"][3] != 0 and np.sum(direct["code"] != 0) == 1


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
def test_shadow_record_rows_equal_direct_runs(wrapped):
    """The recorder's L2 rows (_record_shadow!, run_record.jl:152-172): row g holds total(sum) / grid[g] and the
    point of the direct run to grid[g]; a start out of range fails at row 1, an overflow at the first row at or
    after its handoff."""
    sysm = (WrappedStandardMap if wrapped else StandardMap)(2.5)
    m, N, tau = 8, 240, 2
    x0 = starts(m, 11)
    x0[2] = [0.4, 1e308]
    if wrapped:
        x0 = np.array([torus_state(z) if np.all(np.isfinite(z)) else [z[0], z[1], 0, 0] for z in x0])
    x0[5, 1] = np.nan
    grid = record_grid(N, tau, points_per_decade=5, first_steps=6, extra_fractions=[0.1, 0.5])
    u = np.broadcast_to([math.cos(math.pi / 4), math.sin(math.pi / 4)], (m, 2))
    rec = O.shadow_record(sysm.descriptor(), x0, u, N, tau, 1e-8, grid)
    s = x0.shape[1]
    assert rec["fail"][5] == 1 and np.all(rec["R"][5] == 0)
    for g, n in enumerate(grid):
        d = O.shadow(sysm.descriptor(), x0, u, n, tau, 1e-8)
        for i in range(m):
            if rec["fail"][i] and g >= rec["fail"][i] - 1:
                assert np.all(rec["R"][i, g] == 0)
                continue
            assert d["code"][i] == 0
            assert rec["R"][i, g, 0] == (d["sum"][i] + d["comp"][i]) / n
            assert np.array_equal(rec["R"][i, g, 8:8 + s], d["x"][i]) and np.all(rec["R"][i, g, 1:8] == 0)
    f2 = int(rec["fail"][2])
    hit = int(O.shadow(sysm.descriptor(), x0[2:3], u[2:3], N, tau, 1e-8)["step"][0])
    assert f2 == next(g for g, n in enumerate(grid) if n >= hit) + 1
    assert np.sum(rec["fail"] > 0) == 2


def psi(sysm, x0, w0, n):
    """Psi_n of the reference's suite (test/lyap-benettin/lb_psi.jl): sum over j of ||M_n ... M_j|| ||w_{j-1}|| /
    ||w_n||, tangents at the realized points (a tolerance scale, plain float64)."""
    K = sysm.K
    z = np.array(x0, dtype=np.float64)
    Ms = []
    for _ in range(n):
        c = K * math.cos(z[0])
        Ms.append(np.array([[1 + c, 1.0], [c, 1.0]]))
        p = z[1] + K * math.sin(z[0])
        z = np.array([z[0] + p, p] + list(z[2:]))
    lw = np.zeros(n + 1)
    w = np.array(w0, dtype=np.float64)
    a = np.linalg.norm(w)
    w, acc = w / a, math.log(a)
    lw[0] = acc
    for k in range(n):
        w = Ms[k] @ w
        a = np.linalg.norm(w)
        acc += math.log(a)
        w = w / a
        lw[k + 1] = acc
    P, lP, out = np.eye(2), 0.0, 0.0
    for j in range(n - 1, -1, -1):
        Q = P @ Ms[j]
        f = np.linalg.norm(Q)
        P, lP = Q / f, lP + math.log(f)
        out += math.exp(lP + math.log(np.linalg.norm(P, 2)) + lw[j] - lw[n])
    return out


@pytest.mark.parametrize("mode", ["wrapped", "lifted"])
def test_ac12_two_trajectory_against_benettin(mode):
    """AC12: K = 5, n = 1000, d0 = 1e-8, 100 starts: |X2 - X_L1| <= Delta_2T Psi_n / n + B_S(n) / n, Delta_2T =
    2 (E_nl + E_ro) with M_state 2 pi (wrapped) or the largest lifted coordinate of the orbit; every lifted start is
    valid at n = 1000."""
    K, d0, n, m = 5.0, 1e-8, 1000, 100
    wrapped = mode == "wrapped"
    sysm = (WrappedStandardMap if wrapped else StandardMap)(K)
    z0 = starts(m, 25600)
    x0 = np.array([torus_state(z) for z in z0]) if wrapped else z0
    u = O.frames(25600, 1, m, 2, 1)[:, :, 0]
    chain, allr = make_chain((Benettin(),))
    l1 = O.run(sysm.descriptor(), chain, x0, u[:, :, None], n, 1, 1)
    X = finals_of(decode_states(l1["states"], allr, 2, 1), allr, 2, 1)[0]["lambda"][:, 0]
    e = O.shadow(sysm.descriptor(), x0, u, n, 1, d0)
    assert np.all(e["code"] == 0)
    X2 = (e["sum"] + e["comp"]) / n
    lifted = StandardMap(K)
    for j in range(m):
        if wrapped:
            M = TWO_PI
        else:
            z, M = z0[j].copy(), max(1.0, K)
            for _ in range(n):
                M = max(M, abs(z[0]), abs(z[1]))
                p = z[1] + K * math.sin(z[0])
                z = np.array([z[0] + p, p])
            M = max(M, abs(z[0]), abs(z[1]))
        assert Ero(K, d0, M) <= 0.5
        c = K * math.cos(z0[j][0])
        G = (1 + c) ** 2 + 1 + c * c + 1
        bound = 2 * (Enl(K, d0) + Ero(K, d0, M)) * psi(lifted, z0[j], u[j], n) / n + BS(n, G, K) / n
        assert abs(X2[j] - X[j]) <= bound, (j, X2[j], X[j], bound)


def test_contract_checks_before_device_work():
    """AC10 contract checks: N not a positive multiple of tau, a d0 that is not finite and positive or would be
    narrowed, a wrong state length or U0 count raise before any device work (these run without a GPU)."""
    sysm = StandardMap(2.0)
    x0 = [[0.3, 0.4]]
    with pytest.raises(ValueError):
        propagate_shadow(sysm, x0, 7, d0=1e-8, tau=5, seed=1)
    with pytest.raises(ValueError):
        propagate_shadow(sysm, x0, 0, d0=1e-8, tau=1, seed=1)
    for bad in (0.0, -1e-6, float("nan"), float("inf")):
        with pytest.raises(ValueError):
            propagate_shadow(sysm, x0, 4, d0=bad, tau=1, seed=1)
    with pytest.raises(ValueError):
        check_d0(np.longdouble("1e-8"))
    assert check_d0(np.float32(0.5)) == 0.5 and check_d0(3) == 3.0
    with pytest.raises(ValueError):
        propagate_shadow(WrappedStandardMap(1.0), x0, 4, d0=1e-8, seed=1)
    with pytest.raises(ValueError):
        propagate_shadow(sysm, x0, 4, d0=1e-8, seed=1, U0=[[1.0, 0.0], [0.0, 1.0]])


def test_u0_normalised():
    """U0 is normalised by its safe norm: (3, 4) and (0.6, 0.8) give the same estimate within 1e-9 (AC10)."""
    u = unit_directions([[3.0, 4.0], [0.6, 0.8]], 2)
    assert abs(u[0, 0] - 0.6) <= EPS and abs(u[0, 1] - 0.8) <= EPS
    e = O.shadow(StandardMap(2.0).descriptor(), [[0.3, 0.4], [0.3, 0.4]], u, 100, 1, 1e-6)
    X2 = (e["sum"] + e["comp"]) / 100
    assert abs(X2[0] - X2[1]) <= 1e-9
    n = unit_directions([[np.nan, 1.0], [0.0, 0.0]], 2)
    assert np.all(np.isnan(n))
