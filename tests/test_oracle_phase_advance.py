"""WeightedPhaseAdvance (src/Indicators/naff_reducer.jl:142-254) and the portable atan2 its planar increment uses, on
the CPU: atan2 against libm; the oracle's restatement pinned by closed forms (a rotation map's planar increments are
its angle, a free rotation's lifted increments its momentum, so nu = angle / 2 pi and the windows agree below the
floor), the branch-cut count, the failure of a non-finite sample, the run-length rule and the final's NaN rules."""
import math

import numpy as np
import pytest

from chaos_amd import (StandardMap, WrappedStandardMap, WeightedPhaseAdvance, PlanarObservable, LiftedAngleObservable,
                       RunMonitor, torus_state)
from chaos_amd.tangent import make_chain, decode_states
from tests import oracle_lib as O


def ulps(a, b):
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(b), 1e-300))


def test_atan2_against_libm():
    rng = np.random.default_rng(1)
    n = 400_000
    mag = lambda: 10.0 ** rng.uniform(-300, 300, n) * rng.choice([-1, 1], n)
    for y, x in ((mag(), mag()), (rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)),
                 (rng.uniform(-1, 1, n) * 1e-8, rng.choice([-1.0, 1.0], n))):
        a, b = O.atan2(y, x), np.arctan2(y, x)
        assert np.max(ulps(a, b)) <= 2.0
        assert np.mean(a == b) > 0.75
    sp = np.array([[0.0, 1.0], [0.0, -1.0], [-0.0, 1.0], [-0.0, -1.0], [0.0, 0.0], [0.0, -0.0], [-0.0, -0.0],
                   [1.0, 0.0], [-1.0, 0.0], [np.inf, 1.0], [1.0, np.inf], [1.0, -np.inf], [np.inf, np.inf],
                   [np.inf, -np.inf], [-np.inf, -np.inf]])
    assert np.array_equal(O.atan2(sp[:, 0], sp[:, 1]).view(np.uint64), np.arctan2(sp[:, 0], sp[:, 1]).view(np.uint64))
    assert np.isnan(O.atan2(np.array([np.nan]), np.array([1.0]))[0])


def pa_run_linear(theta, m, windows, order=2):
    r = WeightedPhaseAdvance(window=windows, observable=PlanarObservable(1, 2), birkhoff_order=order)
    N = 2 * max(windows) + 1
    chain, allr = make_chain((r,))
    A = np.array([[math.cos(theta), -math.sin(theta)], [math.sin(theta), math.cos(theta)]])
    rng = np.random.default_rng(3)
    x0 = rng.uniform(-1, 1, (m, 2))
    W1 = np.broadcast_to(np.array([[1.0], [0.0]]), (m, 2, 1))
    out = O.run_linear(A, chain, x0, W1, N, 1, 1)
    return r, r.final(decode_states(out["states"], allr, 2, 1)[0]), out


@pytest.mark.parametrize("theta", [0.3, -1.1, 2.9])
def test_rotation_map_frequency_is_its_angle(theta):
    r, f, out = pa_run_linear(theta, 32, (40, 100))
    assert np.all(out["code"] == 0) and np.all(f["complete"])
    nu = theta / (2 * math.pi)
    assert np.max(np.abs(f["nu1"] - nu)) < 1e-13 and np.max(np.abs(f["nu2"] - nu)) < 1e-13
    assert np.all(f["below_floor"]) and np.all(f["D"] == math.log10(32 * 2.0 ** -52))
    # |theta| > 0.9 pi: every increment of the windows is counted as near the branch cut
    assert np.all(f["near_pi"] == (200 if abs(theta) > 0.9 * math.pi else 0))


def test_free_rotation_lifted_angle():
    """K = 0: p stays, x' = x + p, so the lifted increments are p and nu = p / 2 pi."""
    m, W = 16, (32,)
    r = WeightedPhaseAdvance(window=W, observable=LiftedAngleObservable(1), birkhoff_order=1)
    chain, allr = make_chain((r,), reset_frame=True)
    p = np.linspace(0.1, 2.0, m)
    x0 = np.column_stack([np.zeros(m), p])
    W1 = np.broadcast_to(np.array([[1.0], [0.0]]), (m, 2, 1))
    out = O.run(StandardMap(0.0).descriptor(), chain, x0, W1, 65, 1, 1)
    f = r.final(decode_states(out["states"], allr, 2, 1)[0])
    assert np.max(np.abs(f["nu1"][:, 0] - p / (2 * math.pi))) < 1e-13 and np.all(f["near_pi"] == 0)


def test_partial_run_nan_rules_and_failure():
    r, f, _ = pa_run_linear(0.5, 4, (40,))
    r2 = WeightedPhaseAdvance(window=(40,), observable=PlanarObservable(1, 2), birkhoff_order=2)
    chain, allr = make_chain((r2,))
    A = np.array([[math.cos(0.5), -math.sin(0.5)], [math.sin(0.5), math.cos(0.5)]])
    x0 = np.array([[0.5, 0.1], [np.nan, 0.2]])
    W1 = np.broadcast_to(np.array([[1.0], [0.0]]), (2, 2, 1))
    out = O.run_linear(A, chain, x0, W1, 60, 1, 1)   # 60 handoffs: window 1 complete, window 2 not
    st = decode_states(out["states"], allr, 2, 1)[0]
    f = r2.final(st)
    assert np.isfinite(f["nu1"][0, 0]) and np.isnan(f["nu2"][0, 0]) and np.isnan(f["D"][0, 0])
    assert not f["below_floor"][0, 0] and not f["complete"][0]
    assert out["code"][1] == 1 and out["hstep"][1] == 1   # a non-finite point fails the body's check first
    with pytest.raises(ValueError):
        r2.check_run(1, 80)
    r2.check_run(1, 81)
    r2.check_run(3, 243)


def test_wrapped_chain_with_monitor_runs():
    """The study's pa spec (windows 250, 1000, order 2, planar (q, p)) on the wrapped map at K = 0.5 after RunMonitor:
    complete runs, finite frequencies in [-1/2, 1/2] and distances."""
    m = 24
    sysm = WrappedStandardMap(0.5)
    r = WeightedPhaseAdvance(window=(250, 1000), observable=PlanarObservable(1, 2), birkhoff_order=2)
    chain, allr = make_chain((RunMonitor(sysm), r), reset_frame=True)
    rng = np.random.default_rng(2)
    x0 = np.array([torus_state(v) for v in rng.uniform(0.5, 1.5, (m, 2))])
    W1 = np.broadcast_to(np.array([[1.0], [0.0]]), (m, 2, 1))
    out = O.run(sysm.descriptor(), chain, x0, W1, 2001, 1, 1)
    f = r.final(decode_states(out["states"], allr, 2, 1)[1])
    assert np.all(out["code"] == 0) and np.all(f["complete"]) and np.all(np.isfinite(f["delta"]))
    assert np.all(np.abs(f["nu1"]) <= 0.5) and np.all(np.abs(f["nu2"]) <= 0.5)
