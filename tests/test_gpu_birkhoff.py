"""WeightedBirkhoff on the device against the oracle (or_propagate with red_wbirk) bit for bit: the horizon sums held
in the state array, the register tail (n, handoffs, code), points, frames, codes; launch splits; the host chunk's
Julia-layout states (src/Indicators/wbirkhoff_reducer.jl:4-116). Every test prints its case before touching the
GPU."""
import ctypes
import math
import sys

import numpy as np
import pytest

from chaos_amd import (StandardMap, WrappedStandardMap, WeightedBirkhoff, CosObservable, SinObservable,
                       LiftedObservable, RunMonitor, propagate, torus_state)
from chaos_amd import _lib as L
from chaos_amd.tangent import make_chain, decode_states, finals_of
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def starts(m, wrapped, seed):
    rng = np.random.default_rng(seed)
    z = 2 * math.pi * rng.random((m, 2))
    z[3] = [0.5, 1e307]          # the lifted observable overflows (x + 2 pi p): an "observable" failure
    z[4] = [np.nan, 0.1]         # a non-finite state
    if wrapped:
        x = np.array([torus_state(v) if np.all(np.isfinite(v)) else [v[0], v[1], 0.0, 0.0] for v in z])
        x[3] = [0.5, 1.0, 0.0, 1e306]   # the winding count makes the lifted p overflow
        return x
    return z


def observables(wrapped):
    lift = LiftedObservable(2, 4) if wrapped else LiftedObservable(1, 2)
    return (CosObservable(1), SinObservable(1), CosObservable(2, 3), SinObservable(2), lift)


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
@pytest.mark.parametrize("mon", [False, True], ids=["alone", "monitor"])
@pytest.mark.parametrize("p,tau", [(1, 1), (2, 3)])
def test_device_birkhoff_equals_oracle(wrapped, mon, p, tau):
    m = 2000
    Ns = (16, 64, 100)
    N = 2 * max(Ns) * tau
    sysm = (WrappedStandardMap if wrapped else StandardMap)(1.5)
    x0 = starts(m, wrapped, 7 + p + tau)
    red = WeightedBirkhoff(window=Ns, observables=observables(wrapped), birkhoff_order=p)
    red.check_run(tau, N)
    reds = (RunMonitor(sysm), red) if mon else (red,)
    chain, allr = make_chain(reds)
    W0 = O.frames(3, 1, m, 2, 1)
    orc = O.run(sysm.descriptor(), chain, x0, W0, N, tau, 1)
    assert np.sum(orc["code"] != 0) >= 2
    for L_ in (None, 37 * tau):
        say("birkhoff", "wrapped" if wrapped else "lifted", "monitor" if mon else "alone", "p", p, "tau", tau,
            "m", m, "N", N, "launch", L_)
        res = propagate(sysm, x0, N, k=1, tau=tau, seed=3, reducers=reds, W0=W0, launch_steps=L_, check=False,
                        keep_states=True)
        assert np.array_equal(res.code, orc["code"]) and np.array_equal(res.hstep, orc["hstep"])
        assert np.array_equal(res.states, orc["states"])
        ok = orc["code"] == 0
        assert np.array_equal(res.x[ok].view(np.uint64), orc["x"][ok].view(np.uint64))
        assert np.array_equal(res.W[ok].view(np.uint64), orc["W"][ok].view(np.uint64))
    fin = res.finals[-1]
    assert np.all(fin["complete"][ok]) and np.all(np.isfinite(fin["avg2"][ok]))


def test_host_chunk_states():
    """cl_propagate_chunk with the Julia-layout states of (RunMonitor, WeightedBirkhoff) over two launches."""
    from chaos_amd.tangent import _context
    m, tau = 500, 2
    Ns = (16, 50)
    N = 2 * max(Ns) * tau
    sysm = WrappedStandardMap(0.8)
    x0 = starts(m, True, 2)
    red = WeightedBirkhoff(window=Ns, observables=observables(True))
    chain, allr = make_chain((RunMonitor(sysm), red))
    W0 = O.frames(5, 1, m, 2, 1)
    say("birkhoff host chunk", "m", m, "N", N)
    lib, ctx = _context()
    sysd = sysm.descriptor()
    nb = lib.cl_state_bytes(ctypes.byref(sysd), ctypes.byref(chain), L.CL_ELT_F64, 2, 1)
    assert nb == 24 + 8 * 2 * (2 + 5 * 5) + 16
    x = np.ascontiguousarray(x0.copy())
    W = np.ascontiguousarray(W0.transpose(0, 2, 1).copy())
    st = np.zeros((m, nb), dtype=np.uint8)
    code = np.zeros(m, dtype=np.int8)
    h = np.zeros(m, dtype=np.int64)
    for n0, steps in ((0, 60), (60, N - 60)):
        rc = lib.cl_propagate_chunk(ctx, ctypes.byref(sysd), ctypes.byref(chain), L.CL_ELT_F64, m, 4, 2, 1, n0, steps,
                                    N, tau, x.ctypes.data, W.ctypes.data, None, st.ctypes.data, code.ctypes.data,
                                    h.ctypes.data, None)
        assert rc == 0, lib.cl_last_error(ctx).decode()
    orc = O.run(sysd, chain, x0, W0, N, tau, 1)
    assert np.array_equal(st, orc["states"]) and np.array_equal(code, orc["code"]) and np.array_equal(h, orc["hstep"])


def test_full_size_sample():
    """2^20 orbits of the wrapped map at K = 0.5 (mostly regular), horizons 100..1600 (3200 handoffs), cos and sin of
    both coordinates: a sample of 1024 orbits equals the oracle bit for bit; regular orbits converge (many horizons
    below the floor or with more than 8 digits)."""
    m = 1 << 20
    Ns = (100, 200, 400, 800, 1600)
    N = 2 * max(Ns)
    sysm = WrappedStandardMap(0.5)
    rng = np.random.default_rng(17)
    x0 = np.array([torus_state(v) for v in 2 * math.pi * rng.random((m, 2))])
    obs = (CosObservable(1), SinObservable(1), CosObservable(2), SinObservable(2))
    red = WeightedBirkhoff(window=Ns, observables=obs)
    say("birkhoff full size", "m", m, "N", N)
    res = propagate(sysm, x0, N, k=1, tau=1, seed=9, reducers=(red,), check=False, keep_states=True)
    assert np.mean(res.code == 0) > 0.99
    idx = np.sort(rng.choice(m, 1024, replace=False))
    chain, _ = make_chain((red,))
    orc = O.run(sysm.descriptor(), chain, x0[idx], O.frames(9, 1, m, 2, 1)[idx], N, 1, 1)
    assert np.array_equal(res.states[idx], orc["states"]) and np.array_equal(res.code[idx], orc["code"])
    dig = res.finals[0]["dig"][res.code == 0][:, :, -1]
    assert np.median(dig) > 8


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
def test_window_mode_equals_oracle(wrapped):
    """The window kinds' constant deviation column (CL_CHAIN_RESET_FRAME): (RunMonitor, WeightedBirkhoff) on chaotic
    orbits at K = 2 pi over windows 250 and 1000 (order 2, the reference's bk spec) equals the oracle bit for bit."""
    m = 3000
    sysm = (WrappedStandardMap if wrapped else StandardMap)(2 * math.pi)
    x0 = starts(m, wrapped, 31)
    obs = (CosObservable(1), SinObservable(1), CosObservable(2, 2),
           LiftedObservable(2, 4) if wrapped else LiftedObservable(1, 2))
    red = WeightedBirkhoff(window=(250, 1000), observables=obs, birkhoff_order=2)
    chain, _ = make_chain((RunMonitor(sysm), red), reset_frame=True)
    W0 = np.broadcast_to(np.array([[1.0], [0.0]]), (m, 2, 1))
    orc = O.run(sysm.descriptor(), chain, x0, W0, 2000, 1, 1)
    say("birkhoff window mode", "wrapped" if wrapped else "lifted", "m", m)
    res = propagate(sysm, x0, 2000, k=1, tau=1, seed=1, reducers=(RunMonitor(sysm), red), check=False,
                    keep_states=True, reset_frame=True, launch_steps=700)
    assert np.array_equal(res.code, orc["code"]) and np.array_equal(res.states, orc["states"])
    assert np.mean(orc["code"] == 0) > 0.99
