"""The recorder (src/Experiments/run_record.jl, run_device.jl, run_grid.jl) on the CPU: the grid and the start
frame against the reference's own checks (test/experiments/ex_plan.jl:212-230, ex_chunks.jl:37-52), and the oracle's
segmented recording against direct runs to each grid step (the recorder's defining property: the value stored at a
grid step is the final value of the same estimator run directly to that step, run_record.jl:1-3)."""
import math

import mpmath as mp
import numpy as np
import pytest

from chaos_amd import StandardMap, WrappedStandardMap, RunMonitor, torus_state
from chaos_amd.recorder import record_grid, deviation_frame, estimator, vectors, Records, METHODS, VARIATIONAL
from chaos_amd.tangent import make_chain, decode_states, finals_of
from tests import oracle_lib as O

KW = dict(points_per_decade=20, first_steps=100, extra_fractions=[0.1, 0.5])


def test_record_grid_d53d():
    """ex_plan.jl:212-230: strictly increasing multiples of tau, every early step, N, N/10 and N/2 on the grid; 201
    points at N = 10^7 and 221 at 10^8 (tau = 1)."""
    for N in (10 ** 3, 10 ** 4, 10 ** 5, 10 ** 6, 10 ** 7):
        for tau in (1, 2, 4, 8, 16):
            if N % tau:
                continue
            g = record_grid(N, tau, **KW)
            assert all(b > a for a, b in zip(g, g[1:])) and all(n % tau == 0 for n in g)
            gs = set(g)
            assert all(tau * -(-s // tau) in gs for s in range(1, 101))
            assert N in gs and tau * -(-(N // 10) // tau) in gs and tau * -(-(N // 2) // tau) in gs
    assert len(record_grid(10 ** 7, 1, **KW)) == 201
    assert len(record_grid(10 ** 8, 1, **KW)) == 221


def test_record_grid_small_by_hand():
    """N = 64, 5 points per decade, 10 first steps, fractions 0.1 and 0.5 (the grid of hot_loop_bytes,
    run_record.jl:308): 1..10, round(10^(j/5)) = 16, 25, 40, 63, round(6.4) = 6, round(32) = 32, and 64."""
    assert record_grid(64, 1, points_per_decade=5, first_steps=10, extra_fractions=[0.1, 0.5]) == \
        list(range(1, 11)) + [16, 25, 32, 40, 63, 64]
    assert record_grid(64, 4, points_per_decade=5, first_steps=10, extra_fractions=[0.1, 0.5]) == \
        [4, 8, 12, 16, 28, 32, 40, 64]


def test_deviation_frame_ex_chunks():
    """ex_chunks.jl:37-52: first column within 4 eps of (cos a, sin a) at 1024 bits, second column the exact
    counterclockwise quarter turn, k = 1 the first column."""
    for text in ("pi/4", "0.3 + pi/5", "2pi/3", "-1.25", "(pi - 0.5)/7"):
        W = deviation_frame(text, 2)
        with mp.workprec(1024):
            a = eval(text.replace("2pi", "2*pi").replace("pi", "mp.pi"), {"mp": mp})
            c, s = mp.cos(a), mp.sin(a)
        tol = 4 * 2.0 ** -52
        assert abs(mp.mpf(W[0, 0]) - c) <= tol and abs(mp.mpf(W[1, 0]) - s) <= tol
        assert W[0, 1] == -W[1, 0] and W[1, 1] == W[0, 0]
        assert np.array_equal(deviation_frame(text, 1)[:, 0], W[:, 0])


def test_estimators_and_vectors():
    assert set(METHODS) == {"L1", "L2", "L3", "L4", "L4b", "L5", "L5b"} and set(VARIATIONAL) == set(METHODS) - {"L2"}
    assert vectors("L1") == 1 and vectors("L2") == 1 and vectors("L4") == 2
    for bad in ("L9", "L2"):
        with pytest.raises(ValueError):
            estimator(bad)


def starts(m, wrapped, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 2 * np.pi, (m, 2))
    x[: m // 3] *= 20.0
    return np.array([torus_state(z) for z in x]) if wrapped else x


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
@pytest.mark.parametrize("method", ["L1", "L4", "L5", "L3"])
def test_oracle_record_equals_direct_runs(method, wrapped):
    """Each row of the segmented recording equals the finals of a direct run to that grid step, bit for bit: the
    estimator's exponents, G_max and orth_max, the monitor's maxima, the point; the sum and pairing columns are the
    running maxima of |lambda_1 + lambda_2| (k = 2)."""
    k = vectors(method)
    sys = (WrappedStandardMap if wrapped else StandardMap)(2.5)
    m, N, tau = 6, 240, 2
    grid = record_grid(N, tau, points_per_decade=5, first_steps=6, extra_fractions=[0.1, 0.5])
    x0 = starts(m, wrapped)
    W0 = np.broadcast_to(deviation_frame("pi/4", k), (m, 2, k))
    chain, allr = make_chain((estimator(method), RunMonitor(sys)))
    rec = O.record(sys.descriptor(), chain, x0, W0, N, tau, k, grid)
    assert np.all(rec["fail"] == 0)
    s = x0.shape[1]
    run_sum = np.zeros(m)
    for g, n in enumerate(grid):
        d = O.run(sys.descriptor(), chain, x0, W0, n, tau, k)
        fin = finals_of(decode_states(d["states"], allr, 2, k), allr, 2, k)
        row = rec["R"][:, g]
        lam = fin[0]["lambda"]
        assert np.array_equal(row[:, :k], lam)
        assert np.array_equal(row[:, k], fin[1]["det_max"]) and np.array_equal(row[:, k + 1], fin[1]["symp_max"])
        if method != "L1":
            assert np.array_equal(row[:, k + 2], fin[0]["orth_max"]) and np.array_equal(row[:, k + 5], fin[0]["G_max"])
            e = np.abs(lam[:, 0] + lam[:, 1])
            run_sum = e if g == 0 else np.maximum(run_sum, e)
            assert np.array_equal(row[:, k + 3], run_sum) and np.array_equal(row[:, k + 4], run_sum)
        else:
            assert np.all(row[:, k + 2] == 0) and np.all(row[:, k + 5] == 1.0) and np.all(row[:, k + 3:k + 5] == 0)
        assert np.array_equal(row[:, k + 6], fin[1]["M_max"])
        assert np.array_equal(row[:, k + 7:k + 7 + s], d["x"])
    end = O.run(sys.descriptor(), chain, x0, W0, N, tau, k)
    assert np.array_equal(rec["states"], end["states"]) and np.array_equal(rec["x"], end["x"])


def test_oracle_record_failures():
    """An orbit whose frame has a zero column fails at the first handoff (row 1, no rows); under Benettin, an orbit
    whose state overflows fails at the first grid row at or after the overflowing handoff; the others record every
    row."""
    sys = StandardMap(6.0)
    m, N, tau = 5, 400, 1
    grid = record_grid(N, tau, points_per_decade=5, first_steps=4, extra_fractions=[0.5])
    x0 = starts(m, False, seed=8)
    W0 = np.array(np.broadcast_to(deviation_frame("pi/4", 2), (m, 2, 2)))
    W0[1][:, 1] = 0.0
    chain, allr = make_chain((estimator("L4"), RunMonitor(sys)))
    rec = O.record(sys.descriptor(), chain, x0, W0, N, tau, 2, grid)
    assert rec["fail"][1] == 1 and np.all(rec["R"][1] == 0)
    ok = [i for i in range(m) if i != 1]
    assert np.all(rec["fail"][ok] == 0)
    # Benettin with a huge vector: the renormalisation keeps it finite, so plant the overflow in the state
    x1 = x0.copy()
    x1[2] = [0.5, 1e308]
    chain1, _ = make_chain((estimator("L1"), RunMonitor(sys)))
    W1 = np.broadcast_to(deviation_frame("pi/4", 1), (m, 2, 1))
    r1 = O.record(sys.descriptor(), chain1, x1, W1, N, tau, 1, grid)
    direct = O.run(sys.descriptor(), chain1, x1[2:3], W1[2:3], N, tau, 1)
    fail_step = int(direct["hstep"][0])
    want = next(g for g, n in enumerate(grid) if n >= fail_step) + 1
    assert r1["fail"][2] == want and np.all(r1["R"][2, want - 1:] == 0)


def test_unpack_blanks_from_the_failing_row():
    """_unpack! / _blank_failed!: rows before the failing row, NaN from it on, G_max and M_max of the last good
    row; winding counts from the wrapped state."""
    G, k, s = 4, 2, 4
    R = np.arange(2 * G * (k + 7 + s), dtype=np.float64).reshape(2, G, k + 7 + s)
    rec = Records(method="L4", grid=np.array([1, 2, 3, 4]), R=R, fail=np.array([0, 3]), code=np.zeros(2, np.int8),
                  hstep=np.zeros(2, np.int64), x=np.zeros((2, s)), W=np.zeros((2, 2, k)), states=None, k=k, s=s)
    a, b = rec.unpack(0), rec.unpack(1)
    assert np.array_equal(a["lyap"], R[0, :, :2]) and a["gmax"] == R[0, 3, k + 5] and a["failrow"] == 0
    assert np.all(np.isnan(b["lyap"][2:])) and np.array_equal(b["lyap"][:2], R[1, :2, :2])
    assert b["gmax"] == R[1, 1, k + 5] and b["mmax"] == R[1, 1, k + 6] and np.array_equal(b["t"], [1, 2, 0, 0])
    assert np.array_equal(b["winding"][:2], R[1, :2, k + 9:k + 11].astype(np.int64))


def test_run_monitor_equals_residual_monitor_maxima():
    """RunMonitor (run_record.jl:37-40) keeps the same running maxima of the det and symplecticity residuals and
    of max |M| as ResidualMonitor (monitor.jl:44-54, pinned by the reference's D34 checks) when the latter's
    tolerances never trip; it never fails."""
    from chaos_amd import ResidualMonitor, QRSpectrum, MGS
    sys = StandardMap(3.5)
    m, N = 16, 500
    x0 = starts(m, False, seed=21)
    W0 = np.broadcast_to(deviation_frame("pi/3", 2), (m, 2, 2))
    ch, allr = make_chain((QRSpectrum(MGS()), RunMonitor(sys), ResidualMonitor(sys, 1e30, math.inf)))
    r = O.run(sys.descriptor(), ch, x0, W0, N, 1, 2)
    assert np.all(r["code"] == 0)
    fin = finals_of(decode_states(r["states"], allr, 2, 2), allr, 2, 2)
    for key in ("det_max", "symp_max", "M_max"):
        assert np.array_equal(fin[1][key], fin[2][key]), key
    assert np.all(fin[1]["M_max"] <= 1 + 3.5) and np.all(fin[1]["det_max"] < 1e-14)
