"""The recorder's device path against the oracle's recording (oracle/oracle.cpp, or_record): record rows, failing
rows, final points, frames, chain states, status codes and failing steps bit for bit
(src/Experiments/run_device.jl:9-72, run_record.jl:160-172). Every test prints its case before touching the GPU."""
import sys

import numpy as np
import pytest

from chaos_amd import StandardMap, WrappedStandardMap, RunMonitor, torus_state
from chaos_amd.recorder import record, record_grid, deviation_frame, estimator, vectors, VARIATIONAL
from chaos_amd.tangent import make_chain
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def planted(m, k, wrapped, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 2 * np.pi, (m, 2))
    x[: m // 4] *= 30.0
    x[5] = [np.nan, 0.3]       # the state: fails at the first handoff
    x[9] = [0.4, 1e308]        # overflows within the first steps
    W = np.array(np.broadcast_to(deviation_frame("pi/4", k), (m, 2, k)))
    if k == 2:
        W[11][:, 1] = 0.0      # the factorisation: a zero column
    if wrapped:
        x = np.array([torus_state(z) if np.all(np.isfinite(z)) else np.array([z[0], z[1], 0.0, 0.0]) for z in x])
    return x, W


def same(rec, orc):
    assert np.array_equal(rec.fail, orc["fail"]), "failing rows differ"
    assert np.array_equal(rec.R.view(np.uint64), orc["R"].view(np.uint64)), "record rows differ"
    assert np.array_equal(rec.code, orc["code"]) and np.array_equal(rec.hstep, orc["hstep"])
    assert np.array_equal(rec.x.view(np.uint64), orc["x"].view(np.uint64))
    assert np.array_equal(rec.W.view(np.uint64), orc["W"].view(np.uint64))
    assert np.array_equal(rec.states, orc["states"])


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
@pytest.mark.parametrize("method", VARIATIONAL)
def test_device_record_equals_oracle(method, wrapped):
    m, N, tau = 300, 1500, 3
    k = vectors(method)
    grid = record_grid(N, tau, points_per_decade=10, first_steps=20, extra_fractions=[0.1, 0.5])
    sysm = (WrappedStandardMap if wrapped else StandardMap)(2.0)
    x0, W0 = planted(m, k, wrapped, 4)
    # the default orthogonality tolerance, and a tight one that stops orbits at varied rows
    for tol in ((8.0, 0.6) if k == 2 else (8.0,)):
        chain, _ = make_chain((estimator(method, tol), RunMonitor(sysm)))
        orc = O.record(sysm.descriptor(), chain, x0, W0, N, tau, k, grid)
        assert np.sum(orc["fail"] > 0) >= 2
        for L_ in (None, 21):
            say("recorder", method, "wrapped" if wrapped else "lifted", "tol", tol, "m", m, "N", N, "launch", L_)
            rec = record(sysm, x0, N, method=method, tau=tau, grid=grid, W0=W0, orthogonality_tol=tol,
                         launch_steps=L_)
            same(rec, orc)


def test_device_record_study_grid():
    """The study grid (20 points per decade, every step to 100, fractions 0.1 and 0.5) on a chaotic ensemble."""
    m, N = 512, 20000
    say("recorder study grid", "m", m, "N", N)
    sysm = StandardMap(6.0)
    rng = np.random.default_rng(12)
    x0 = rng.uniform(0, 2 * np.pi, (m, 2))
    rec = record(sysm, x0, N, method="L4", tau=1)
    grid = record_grid(N, 1, points_per_decade=20, first_steps=100, extra_fractions=[0.1, 0.5])
    assert np.array_equal(rec.grid, grid)
    chain, _ = make_chain((estimator("L4"), RunMonitor(sysm)))
    orc = O.record(sysm.descriptor(), chain, x0, np.broadcast_to(deviation_frame("pi/4", 2), (m, 2, 2)), N, 1, 2,
                   grid)
    same(rec, orc)
    # chaotic sea at K = 6: lambda_1 near ln(K/2) at the last row, the pair summing to rounding level
    lam = rec.R[:, -1, :2]
    assert abs(np.median(lam[:, 0]) - np.log(3.0)) < 0.1 and np.max(np.abs(lam.sum(axis=1))) < 1e-9
