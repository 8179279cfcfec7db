"""The indicator recorder's device path (record_indicators: the propagate launches ending at the grid steps and
cl_ind_record_rows) against the oracle's recorder (or_ind_record) bit for bit: rows, failing rows, points, frames,
states, codes (src/Experiments/run_indicator_device.jl:31-91). Every test prints its case before touching the GPU."""
import math
import sys

import numpy as np
import pytest

from chaos_amd import (StandardMap, WrappedStandardMap, RunMonitor, WeightedBirkhoff, CosObservable, SinObservable,
                       LiftedObservable, torus_state)
from chaos_amd.recorder import IND_MEMBER_COLUMNS, ind_members, record_grid, record_indicators
from chaos_amd.tangent import make_chain
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def starts(m, wrapped, seed):
    rng = np.random.default_rng(seed)
    z = 2 * math.pi * rng.random((m, 2))
    z[1] = [np.nan, 0.5]               # fails in the first segment
    if wrapped:
        return np.array([torus_state(v) if np.all(np.isfinite(v)) else [v[0], v[1], 0.0, 0.0] for v in z])
    z[2] = [0.5, 1e306]                # overflows after some 180 steps: a failing row inside the grid
    return z


@pytest.mark.parametrize("members", [("FLI", "MEGNO", "SALI"), ("MEGNO", "FLI", "SALI"), ("MEGNO",), ("FLI",),
                                     ("SALI",)])
@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
@pytest.mark.parametrize("tau,launch", [(1, None), (2, 38)])
def test_tangent_kind_equals_oracle(members, wrapped, tau, launch):
    m, N = 1000, 400
    sysm = (WrappedStandardMap if wrapped else StandardMap)(6.0)
    x0 = starts(m, wrapped, 11 + tau)
    s = x0.shape[1]
    grid = record_grid(N, tau, points_per_decade=8, first_steps=20, extra_fractions=(0.1, 0.5))
    W0 = O.frames(6, 0, m, 2, 2)
    chain, _ = make_chain(ind_members(members) + (RunMonitor(sysm),))
    ncol = sum(len(IND_MEMBER_COLUMNS[n]) for n in members) + 3 + s
    orc = O.ind_record(sysm.descriptor(), chain, x0, W0, N, tau, 2, grid, ncol)
    say("ind tangent", members, "wrapped" if wrapped else "lifted", "tau", tau, "launch", launch, "m", m, "N", N,
        "rows", len(grid))
    rec = record_indicators(sysm, x0, N, members=members, tau=tau, grid=grid, W0=W0, launch_steps=launch)
    assert rec.R.shape == orc["R"].shape
    assert np.array_equal(rec.fail, orc["fail"]) and np.array_equal(rec.code, orc["code"])
    assert np.array_equal(rec.hstep, orc["hstep"])
    assert np.array_equal(rec.R.view(np.uint64), orc["R"].view(np.uint64))
    assert np.array_equal(rec.states, orc["states"])
    ok = rec.code == 0
    assert np.array_equal(rec.x[ok].view(np.uint64), orc["x"][ok].view(np.uint64))
    assert np.array_equal(rec.W[ok].view(np.uint64), orc["W"][ok].view(np.uint64))
    assert rec.fail[1] == 1 and (wrapped or 1 < rec.fail[2] < len(grid))
    u = rec.unpack(2)
    assert np.all(np.isnan(u["series"][max(u["failrow"] - 1, 0):])) if u["failrow"] else True


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
def test_birkhoff_window_kind_equals_oracle(wrapped):
    m, N, tau = 1500, 2000, 1
    sysm = (WrappedStandardMap if wrapped else StandardMap)(2 * math.pi)
    rng = np.random.default_rng(21)
    z = 2 * math.pi * rng.random((m, 2))
    x0 = np.array([torus_state(v) for v in z]) if wrapped else z
    x0[4] = [np.nan, 0.1, 0.0, 0.0] if wrapped else [np.nan, 0.1]
    lift = LiftedObservable(2, 4) if wrapped else LiftedObservable(1, 2)
    wb = WeightedBirkhoff(window=(250, 1000), observables=(CosObservable(1), SinObservable(2), lift),
                          birkhoff_order=2)
    grid = record_grid(N, tau, points_per_decade=10, first_steps=10, extra_fractions=(0.5,))
    chain, _ = make_chain((RunMonitor(sysm), wb), reset_frame=True)
    W1 = np.broadcast_to(np.array([[1.0], [0.0]]), (m, 2, 1))
    orc = O.ind_record(sysm.descriptor(), chain, x0, W1, N, tau, 1, grid, 3 + x0.shape[1])
    say("ind birkhoff window", "wrapped" if wrapped else "lifted", "m", m, "N", N, "rows", len(grid))
    rec = record_indicators(sysm, x0, N, reader=wb, tau=tau, grid=grid, launch_steps=300)
    assert np.array_equal(rec.fail, orc["fail"]) and np.array_equal(rec.code, orc["code"])
    assert np.array_equal(rec.R.view(np.uint64), orc["R"].view(np.uint64))
    assert np.array_equal(rec.states, orc["states"])
    assert rec.fail[4] == 1 and np.mean(rec.fail == 0) > 0.99


@pytest.mark.parametrize("members", [("RLI", "MEGNO"), ("RLI",)])
@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
@pytest.mark.parametrize("tau,launch", [(1, None), (2, 50)])
def test_pair_kind_equals_oracle(members, wrapped, tau, launch):
    """The pair kind (PairMap + RLI, the study's spec RLI, MEGNO): the block frame stepped block by block, RLI's sums,
    the monitor on orbit A, rows of orbit A's point; planted failures (a non-finite half, a broken block frame)."""
    from chaos_amd import PairMap, pair_start, pair_frame
    m, N = 800, 300
    base = (WrappedStandardMap if wrapped else StandardMap)(6.0)
    rng = np.random.default_rng(31 + tau)
    u = np.array([math.cos(0.3), math.sin(0.3)])
    z = np.array([pair_start(base, 2 * math.pi * rng.random(2), u, 1e-10)[0] for _ in range(m)])
    s = base.state_length
    z[3, s] = np.nan                              # orbit B not finite: fails at the first handoff
    W = np.broadcast_to(pair_frame(u), (m, 4, 2)).copy()
    W[5, 2, 0] = 1e-200                           # off-block entry: "pairing" at the first handoff
    psys = PairMap(base)
    grid = record_grid(N, tau, points_per_decade=8, first_steps=20, extra_fractions=(0.5,))
    chain, _ = make_chain(ind_members(members, pair=True) + (RunMonitor(psys),))
    ncol = sum(len(IND_MEMBER_COLUMNS[n]) for n in members) + 3 + s
    orc = O.ind_record(psys.descriptor(), chain, z, W, N, tau, 2, grid, ncol)
    say("ind pair", members, "wrapped" if wrapped else "lifted", "tau", tau, "launch", launch, "m", m, "N", N)
    rec = record_indicators(psys, z, N, members=members, tau=tau, grid=grid, W0=W, launch_steps=launch)
    assert rec.R.shape == orc["R"].shape
    assert np.array_equal(rec.fail, orc["fail"]) and np.array_equal(rec.code, orc["code"])
    assert np.array_equal(rec.hstep, orc["hstep"])
    assert np.array_equal(rec.R.view(np.uint64), orc["R"].view(np.uint64))
    assert np.array_equal(rec.states, orc["states"])
    ok = rec.code == 0
    assert np.array_equal(rec.x[ok].view(np.uint64), orc["x"][ok].view(np.uint64))
    assert np.array_equal(rec.W[ok].view(np.uint64), orc["W"][ok].view(np.uint64))
    assert rec.fail[3] == 1 and rec.fail[5] == 1 and np.mean(ok) > 0.99


@pytest.mark.parametrize("obs", ["planar", "lifted"])
@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
def test_phase_advance_window_kind_equals_oracle(obs, wrapped):
    """The phase-advance window kind (RunMonitor, WeightedPhaseAdvance; the study's pa spec: windows 250 and 1000,
    order 2, planar (q, p)) and the lifted-angle observable, against the oracle bit for bit."""
    from chaos_amd import WeightedPhaseAdvance, PlanarObservable, LiftedAngleObservable
    m, tau = 1200, 1
    windows = (250, 1000) if obs == "planar" else (64,)
    N = 2 * max(windows) + 1
    sysm = (WrappedStandardMap if wrapped else StandardMap)(0.9)
    rng = np.random.default_rng(41)
    z = 2 * math.pi * rng.random((m, 2))
    x0 = np.array([torus_state(v) for v in z]) if wrapped else z
    x0[6] = [np.nan, 0.1, 0.0, 0.0] if wrapped else [np.nan, 0.1]
    o = PlanarObservable(1, 2) if obs == "planar" else (LiftedAngleObservable(2, 4) if wrapped else LiftedAngleObservable(2))
    pa = WeightedPhaseAdvance(window=windows, observable=o, birkhoff_order=2)
    grid = record_grid(N, tau, points_per_decade=10, first_steps=10, extra_fractions=(0.5,))
    chain, _ = make_chain((RunMonitor(sysm), pa), reset_frame=True)
    W1 = np.broadcast_to(np.array([[1.0], [0.0]]), (m, 2, 1))
    orc = O.ind_record(sysm.descriptor(), chain, x0, W1, N, tau, 1, grid, 3 + x0.shape[1])
    say("ind phase advance", obs, "wrapped" if wrapped else "lifted", "m", m, "N", N, "rows", len(grid))
    rec = record_indicators(sysm, x0, N, reader=pa, tau=tau, grid=grid, launch_steps=400)
    assert np.array_equal(rec.fail, orc["fail"]) and np.array_equal(rec.code, orc["code"])
    assert np.array_equal(rec.R.view(np.uint64), orc["R"].view(np.uint64))
    assert np.array_equal(rec.states, orc["states"])
    assert rec.fail[6] == 1 and np.mean(rec.fail == 0) > 0.99


def test_phase_advance_planted_branch_cases():
    """The planar increment's device-only branches against the oracle bit for bit: a start at the origin (a fixed
    point of the standard map: every sample is 0, the increment NaN), a start about 1e-150 from it (the product of the
    samples falls below 1e-280 and both are rescaled by powers of two first), and starts near the origin at K = -5,
    where the tangent at the origin has two negative eigenvalues, so consecutive samples point nearly opposite and the
    increments fall near the branch cut (near_pi counts above 0)."""
    from chaos_amd import WeightedPhaseAdvance, PlanarObservable
    from chaos_amd.tangent import decode_states, finals_of
    m, tau, windows = 256, 1, (64,)
    N = 2 * max(windows) + 1
    sysm = StandardMap(-5.0)
    rng = np.random.default_rng(43)
    x0 = 1e-3 * rng.standard_normal((m, 2))
    x0[0] = [0.0, 0.0]
    x0[1] = [1e-150, -2e-150]
    x0[2] = [3e-151, 1e-151]
    pa = WeightedPhaseAdvance(window=windows, observable=PlanarObservable(1, 2), birkhoff_order=2)
    grid = record_grid(N, tau, points_per_decade=10, first_steps=10, extra_fractions=(0.5,))
    chain, _ = make_chain((RunMonitor(sysm), pa), reset_frame=True)
    W1 = np.broadcast_to(np.array([[1.0], [0.0]]), (m, 2, 1))
    orc = O.ind_record(sysm.descriptor(), chain, x0, W1, N, tau, 1, grid, 3 + 2)
    say("ind phase advance planted", "m", m, "N", N)
    rec = record_indicators(sysm, x0, N, reader=pa, tau=tau, grid=grid)
    assert np.array_equal(rec.fail, orc["fail"]) and np.array_equal(rec.code, orc["code"])
    assert np.array_equal(rec.R.view(np.uint64), orc["R"].view(np.uint64))
    assert np.array_equal(rec.states, orc["states"])
    fin = finals_of(decode_states(rec.states, rec.reducers, 2, 1), rec.reducers, 2, 1)[1]
    assert np.any(np.asarray(fin["near_pi"]) > 0)
    assert np.all(np.isnan(np.asarray(fin["nu1"])[0]))  # the origin: every increment NaN
