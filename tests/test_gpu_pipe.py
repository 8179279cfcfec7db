"""The wave-specialised pipeline for small single-QR map ensembles (chaos_amd/csrc/pipe_kernel.hpp): the four
stage waves, their LDS hand-off and the fused rerun of failed orbits reproduce the fused kernel and the oracle bit
for bit (src/Tangent/orbit.jl:48-58, src/Lyapunov/qr_spectrum.jl:54-96). CL_PIPE=1 forces the pipeline for every
eligible launch, CL_PIPE=0 forbids it. Every test prints its case before touching the GPU."""
import ctypes
import os
import sys

import numpy as np
import pytest

from chaos_amd import (StandardMap, StandardMapK, WrappedStandardMap, QRSpectrum, CGS, MGS, RGS, HouseholderQR,
                       AugmentedQR, GivensQR, propagate, torus_state)
from chaos_amd.inputs import torus_uniform
from chaos_amd.tangent import make_chain
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

ALGS = [CGS, MGS, RGS, HouseholderQR, AugmentedQR, GivensQR]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def with_pipe(flag, fn):
    os.environ["CL_PIPE"] = flag
    try:
        return fn()
    finally:
        os.environ.pop("CL_PIPE", None)


def same(a, b):
    for key in ("x", "W", "states", "code", "hstep"):
        u, v = np.ascontiguousarray(getattr(a, key)), np.ascontiguousarray(getattr(b, key))
        assert u.shape == v.shape and np.array_equal(u.view(np.uint8), v.view(np.uint8)), key


def same_oracle(res, orc):
    assert np.array_equal(res.code, orc["code"]) and np.array_equal(res.hstep, orc["hstep"])
    assert np.array_equal(res.x.view(np.uint64), orc["x"].view(np.uint64))
    assert np.array_equal(res.W.view(np.uint64), orc["W"].view(np.uint64))
    assert np.array_equal(res.states, orc["states"])


@pytest.mark.parametrize("k", [2, 1])
@pytest.mark.parametrize("alg", ALGS, ids=lambda a: a.__name__)
@pytest.mark.parametrize("system", [StandardMap, WrappedStandardMap], ids=["std", "wrapped"])
def test_pipeline_equals_fused_and_oracle(system, alg, k):
    """A ragged group count (777 orbits: 12 full groups and one of 9), chaotic and regular K, tau with one and
    several handoffs per chunk, and the largest tau the rings hold."""
    m, N = 777, 960
    for K, tau in ((6.0, 1), (0.97, 3), (2.0, 16)):
        say("pipeline", system.__name__, alg.__name__, "k", k, "K", K, "tau", tau, "m", m, "N", N)
        x0 = torus_uniform(31, m)
        x0[: m // 5] *= 30.0
        if system is WrappedStandardMap:
            x0 = np.array([torus_state(z) for z in x0])
        chain = (QRSpectrum(alg()),)
        run = lambda: propagate(system(K), x0, N, k=k, tau=tau, seed=5, reducers=chain, check=False,
                                keep_states=True)
        p = with_pipe("1", run)
        f = with_pipe("0", run)
        same(p, f)
        ch, allr = make_chain(chain)
        orc = O.run(system(K).descriptor(), ch, x0, O.frames(5, 1, m, 2, k), N, tau, k, math=O.MATH_CL)
        same_oracle(p, orc)


def test_pipeline_per_orbit_k():
    m, N = 500, 800
    say("pipeline per-orbit K", "m", m, "N", N)
    rng = np.random.default_rng(4)
    x0 = np.concatenate([torus_uniform(8, m), rng.uniform(0.5, 6.0, (m, 1))], axis=1)
    for alg in (MGS, HouseholderQR):
        run = lambda: propagate(StandardMapK(), x0, N, k=2, seed=9, reducers=(QRSpectrum(alg()),), check=False,
                                keep_states=True)
        same(with_pipe("1", run), with_pipe("0", run))


def test_pipeline_failures_are_rerun_exactly():
    """Failures planted in every stage: a non-finite start (the orbit stage), a zero frame column (the
    factorisation), a frame that overflows in the first step (the frame status), a tight orthogonality tolerance
    (0.5) that most orbits exceed at varied steps (the checks stage). Failed orbits are rerun by the fused kernel:
    codes, failing steps, states and frames equal the fused kernel's and the oracle's, in one launch or many."""
    m, N, tau = 200, 600, 2
    say("pipeline failures", "m", m, "N", N)
    x0 = torus_uniform(12, m)
    x0[3] = [np.nan, 0.5]
    x0[70] = [0.2, np.inf]
    W0 = O.frames(12, 1, m, 2, 2).copy()
    W0[5][:, 1] = 0.0
    W0[130][:, 0] = 0.0
    W0[9] *= 1.5e308
    for tol in (8.0, 0.5):
        chain = (QRSpectrum(MGS(), orthogonality_tol=tol),)
        ch, allr = make_chain(chain)
        orc = O.run(StandardMap(6.0).descriptor(), ch, x0, W0, N, tau, 2, math=O.MATH_CL)
        codes = set(orc["code"].tolist())
        assert {0, 1, 2, 3} <= codes, codes
        if tol < 1:  # the checks stage trips at about a hundred different steps
            assert len(set(orc["hstep"][orc["code"] == 3].tolist())) > 50
        for L_ in (None, 14):
            say("pipeline failures", "tol", tol, "launch", L_)
            run = lambda: propagate(StandardMap(6.0), x0, N, k=2, tau=tau, seed=12, reducers=chain, W0=W0,
                                    check=False, keep_states=True, launch_steps=L_)
            p = with_pipe("1", run)
            same(p, with_pipe("0", run))
            same_oracle(p, orc)


def test_pipeline_device_entry_continues_failed_orbits():
    """Launch after launch through the device entry point: orbits that failed in an earlier launch stay as they
    were (the pipeline leaves them to the fused kernel, which skips them)."""
    m, N = 130, 400
    say("pipeline continuation", "m", m, "N", N)
    x0 = torus_uniform(2, m)
    x0[7] = [np.inf, 0.0]
    chain = (QRSpectrum(HouseholderQR()),)
    run = lambda: propagate(StandardMap(3.0), x0, N, k=2, tau=4, seed=2, reducers=chain, check=False,
                            keep_states=True, launch_steps=40)
    same(with_pipe("1", run), with_pipe("0", run))


def test_pipeline_tau_above_the_ring_falls_back():
    m, N, tau = 100, 340, 17
    say("tau above ring", tau)
    x0 = torus_uniform(3, m)
    run = lambda: propagate(StandardMap(2.0), x0, N, k=2, tau=tau, seed=3, reducers=(QRSpectrum(MGS()),),
                            check=False, keep_states=True)
    same(with_pipe("1", run), with_pipe("0", run))


@pytest.mark.parametrize("alg", [CGS, MGS, RGS, HouseholderQR, AugmentedQR], ids=lambda a: a.__name__)
def test_pipeline_monitored_chain(alg):
    """The recorder's chains (a QR spectrum, then RunMonitor) through the pipeline's seventh wave: equal to the
    fused kernel and the oracle, lifted and wrapped, with failures and a launch split."""
    from chaos_amd import RunMonitor
    m, N, tau = 400, 600, 2
    for system in (StandardMap, WrappedStandardMap):
        x0 = torus_uniform(14, m)
        x0[: m // 5] *= 25.0
        if system is WrappedStandardMap:
            x0 = np.array([torus_state(z) for z in x0])
        x0[3, :2] = [np.inf, 0.1]
        s = system(3.0)
        chain = (QRSpectrum(alg(), orthogonality_tol=0.6), RunMonitor(s))
        for L_ in (None, 10):
            say("pipeline monitored", system.__name__, alg.__name__, "launch", L_)
            run = lambda: propagate(s, x0, N, k=2, tau=tau, seed=6, reducers=chain, check=False, keep_states=True,
                                    launch_steps=L_)
            p = with_pipe("1", run)
            same(p, with_pipe("0", run))
            ch, allr = make_chain(chain)
            same_oracle(p, O.run(s.descriptor(), ch, x0, O.frames(6, 1, m, 2, 2), N, tau, 2, math=O.MATH_CL))
