"""The two-trajectory estimator on the device (`cl_shadow_device`, `cl_propagate_shadow_chunk`,
`cl_shadow_record_rows`) against the oracle (or_shadow, or_shadow_record) bit for bit: end records (reference and
shadow states, compensated sum, previous separation, code, step, value), errors of `propagate_shadow`, launch
splits, and the recorder's L2 rows (src/Lyapunov/two_trajectory.jl:45-143, src/Experiments/run_device.jl:74-98).
Every test prints its case before touching the GPU."""
import ctypes
import math
import sys

import numpy as np
import pytest

from chaos_amd import StandardMap, WrappedStandardMap, torus_state, propagate_shadow, NumericalHealthError
from chaos_amd import _lib as L
from chaos_amd.recorder import record, record_grid, deviation_frame
from chaos_amd.two_trajectory import seeded_directions, unit_directions
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def ensemble(m, wrapped, seed):
    """Random starts with planted cases: a NaN start (code 2 at step 0), p ~ 1e308 (separation lost at step 1), a
    large momentum (resolution lost mid-run), a winding count near 2^53 (wrapped: NaN state)."""
    rng = np.random.default_rng(seed)
    z = 2 * math.pi * rng.random((m, 2))
    z[: m // 5] *= 40.0
    z[7] = [0.4, 1e308]
    z[8] = [0.5, 3e7]
    if wrapped:
        x = np.array([torus_state(v) for v in z])
        x[7] = [0.4, 1.0, 0.0, 0.0]
        x[8] = [1.0, 1.0, 9007199254740990.0, 5.0]
        x[3, 0] = np.nan
        return x
    z[3, 1] = np.nan
    return z


def same_ends(dev, orc, s):
    raw_d = dev.states if hasattr(dev, "states") else dev
    a, b = raw_d.view(np.uint64), orc["raw"].view(np.uint64)
    # NaN values compare as NaN (payloads are not specified), everything else bit for bit
    f_a, f_b = a.view(np.float64), b.view(np.float64)
    both_nan = np.isnan(f_a) & np.isnan(f_b)
    code_col = 2 * s + 3
    both_nan[:, code_col] = False
    both_nan[:, code_col + 1] = False
    diff = (a != b) & ~both_nan
    assert not diff.any(), f"end records differ at {np.argwhere(diff)[:8].tolist()}"


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
@pytest.mark.parametrize("tau", [1, 4])
def test_device_shadow_equals_oracle(wrapped, tau):
    m, N, d0 = 3000, 2000, 1e-8
    sysm = (WrappedStandardMap if wrapped else StandardMap)(3.5)
    x0 = ensemble(m, wrapped, 31 + tau)
    s = x0.shape[1]
    u = seeded_directions(77, m)
    orc = O.shadow(sysm.descriptor(), x0, u, N, tau, d0)
    assert np.sum(orc["code"] != 0) >= 2
    for L_ in (None, 4 * 37):
        say("shadow", "wrapped" if wrapped else "lifted", "tau", tau, "m", m, "N", N, "launch", L_)
        res = propagate_shadow(sysm, x0, N, d0=d0, tau=tau, seed=77, launch_steps=L_, check=False)
        same_ends(res, orc, s)
        ok = orc["code"] == 0
        assert np.array_equal(res.finals[0]["X2"][ok], ((orc["sum"] + orc["comp"]) / N)[ok])


def test_host_chunk_equals_device():
    """cl_propagate_shadow_chunk (host AoS buffers, the _shadow_chunk! launch) gives the device path's records."""
    from chaos_amd.tangent import _context
    m, N, tau, d0 = 777, 900, 3, 1e-7
    sysm = StandardMap(1.7)
    x0 = ensemble(m, False, 5)
    u = seeded_directions(9, m)
    say("shadow host chunk", "m", m, "N", N)
    lib, ctx = _context()
    sysd = sysm.descriptor()
    ends = np.zeros((m, 10), dtype=np.uint64)
    xa, ua = np.ascontiguousarray(x0), np.ascontiguousarray(u)
    for n0, st in ((0, 300), (300, 600)):
        rc = lib.cl_propagate_shadow_chunk(ctx, ctypes.byref(sysd), m, 2, n0, st, N, tau, d0, xa.ctypes.data,
                                           ua.ctypes.data, ends.ctypes.data)
        assert rc == 0, lib.cl_last_error(ctx).decode()
    res = propagate_shadow(sysm, x0, N, d0=d0, tau=tau, seed=9, check=False)
    assert np.array_equal(np.isnan(ends.view(np.float64)), np.isnan(res.states.view(np.float64)))
    fin = ~np.isnan(ends.view(np.float64))
    assert np.array_equal(ends[fin], res.states[fin])
    same_ends(ends, O.shadow(sysd, x0, u, N, tau, d0), 2)


def test_errors_follow_the_reference():
    """_check_starts before any step (lowest orbit whose start is out of range, step 0), then the lowest failing
    orbit of the run ("state" / "separation" with its handoff and value); argument errors before any launch."""
    sysm = StandardMap(2.0)
    x0 = np.array([[0.3, 0.4], [0.5, 3e7], [1.0, 2.0], [0.7, 0.1]])
    say("shadow errors: start check")
    with pytest.raises(NumericalHealthError) as ei:
        propagate_shadow(sysm, np.array([[0.3, 0.4], [1.0, 2.0], [1.0, 2.0]]), 10, d0=1e-20, seed=3)
    e = ei.value
    assert e.what == "separation" and e.step == 0 and e.orbit == 1 and e.value == 0.0 and e.threshold == 0.0
    say("shadow errors: run")
    orc = O.shadow(sysm.descriptor(), x0, seeded_directions(3, 4), 1000, 1, 1e-8)
    j = int(np.nonzero(orc["code"])[0][0])
    with pytest.raises(NumericalHealthError) as ei:
        propagate_shadow(sysm, x0, 1000, d0=1e-8, seed=3)
    e = ei.value
    assert e.orbit == j + 1 and e.step == orc["step"][j] and e.what == "separation"
    assert e.value == orc["value"][j] and e.threshold == (1e-8 if orc["code"][j] == 3 else 0.0)
    say("shadow errors: wrapped count overflow")
    w = WrappedStandardMap(1.0)
    xw = np.array([torus_state([0.2, 0.3]), [1.0, 1.0, 9007199254740990.0, 5.0]])
    with pytest.raises(NumericalHealthError) as ei:
        propagate_shadow(w, xw, 10, d0=1e-8, seed=3)
    e = ei.value
    assert e.what == "state" and e.orbit == 2 and e.step == 1 and math.isnan(e.value) and e.threshold == math.inf


def test_closed_forms_on_device():
    """The hyperbolic fixed point (AC3) and the shear (AC2) on the device equal the oracle's records, which the CPU
    suite pins against the closed forms."""
    rng = np.random.default_rng(3)
    th = 2 * math.pi * rng.random(100)
    U0 = np.stack([np.cos(th), np.sin(th)], axis=1)
    for K, d0, n in ((6.0, 1e-6, 1000), (0.0, 2.0 ** -20, 10000)):
        say("shadow closed form", "K", K, "d0", d0, "n", n)
        res = propagate_shadow(StandardMap(K), np.zeros((100, 2)), n, d0=d0, seed=1, U0=U0)
        orc = O.shadow(StandardMap(K).descriptor(), np.zeros((100, 2)), unit_directions(U0, 100), n, 1, d0)
        same_ends(res, orc, 2)


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
def test_recorder_l2_equals_oracle(wrapped):
    m, N, tau = 400, 1500, 3
    grid = record_grid(N, tau, points_per_decade=10, first_steps=20, extra_fractions=[0.1, 0.5])
    sysm = (WrappedStandardMap if wrapped else StandardMap)(2.0)
    x0 = ensemble(m, wrapped, 12)
    s = x0.shape[1]
    u = np.broadcast_to(deviation_frame("pi/4", 1)[:, 0], (m, 2))
    orc = O.shadow_record(sysm.descriptor(), x0, u, N, tau, 1e-8, grid)
    assert np.sum(orc["fail"] > 0) >= 2
    for L_ in (None, 21):
        say("recorder L2", "wrapped" if wrapped else "lifted", "m", m, "N", N, "launch", L_)
        rec = record(sysm, x0, N, method="L2", tau=tau, grid=grid, eps="1e-8", launch_steps=L_)
        assert np.array_equal(rec.fail, orc["fail"])
        assert np.array_equal(rec.R.view(np.uint64), orc["R"].view(np.uint64))
        same_ends(rec.states, orc["ends"], s)
        assert rec.R.shape == (m, len(grid), 1 + 7 + s)


def test_full_size_sample():
    """The study size of the two-trajectory bench (2^20 orbits of the lifted map at K = 6, 1000 steps, d0 = 1e-8):
    a sample of 4096 orbits equals the oracle bit for bit, and the median X2 is near ln(K / 2)."""
    m, N = 1 << 20, 1000
    rng = np.random.default_rng(2024)
    x0 = 2 * math.pi * rng.random((m, 2))
    say("shadow full size", "m", m, "N", N)
    res = propagate_shadow(StandardMap(6.0), x0, N, d0=1e-8, seed=5, check=False)
    idx = np.sort(rng.choice(m, 4096, replace=False))
    u = seeded_directions(5, m)[idx]
    orc = O.shadow(StandardMap(6.0).descriptor(), x0[idx], u, N, 1, 1e-8)
    same_ends(res.states[idx], orc, 2)
    ok = res.code == 0
    assert ok.mean() > 0.99 and abs(np.median(res.finals[0]["X2"][ok]) - math.log(3.0)) < 0.1
