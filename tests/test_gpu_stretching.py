"""StretchingNumbers on the device (`cl_propagate_device_ws`, `cl_propagate_chunk` with a workspace) against the
oracle (or_propagate_ws) bit for bit: the UInt32 counts of every cell, the chain states (log sums, sums of squares,
handoffs, codes), points, frames, status codes (src/Indicators/stretching.jl:34-147; the workspace protocol,
src/Tangent/propagate.jl:29-50). Every test prints its case before touching the GPU."""
import ctypes
import math
import sys

import numpy as np
import pytest

from chaos_amd import (StandardMap, WrappedStandardMap, StretchingNumbers, RunMonitor, propagate, stretching_range,
                       torus_state)
from chaos_amd import _lib as L
from chaos_amd.tangent import make_chain, decode_states, finals_of
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def ensemble(m, wrapped, seed):
    rng = np.random.default_rng(seed)
    z = 2 * math.pi * rng.random((m, 2))
    z[: m // 4] *= 30.0
    z[5] = [np.nan, 0.2]      # the state: fails at the first handoff
    z[9] = [0.4, 1e308]       # overflows
    if wrapped:
        x = np.array([torus_state(v) if np.all(np.isfinite(v)) else [v[0], v[1], 0.0, 0.0] for v in z])
        return x
    return z


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
@pytest.mark.parametrize("mon", [False, True], ids=["alone", "monitor"])
@pytest.mark.parametrize("replace", [True, False])
def test_device_stretching_equals_oracle(wrapped, mon, replace):
    m, N, tau = 1500, 1200, 2
    sysm = (WrappedStandardMap if wrapped else StandardMap)(3.0)
    x0 = ensemble(m, wrapped, 3 + 2 * wrapped + mon)
    W0 = O.frames(21, 1, m, 2, 2)
    W0[13][:, 1] = 0.0        # a zero column: "deviation matrix" at the first handoff
    sr = StretchingNumbers(range=stretching_range(sysm), bins=257, interval=tau, replace=replace)
    reds = (sr, RunMonitor(sysm)) if mon else (sr,)
    chain, allr = make_chain(reds)
    orc = O.run_ws(sysm.descriptor(), chain, sr.workspace(2, 2), x0, W0, N, tau, 2)
    assert np.sum(orc["code"] != 0) >= 3
    for L_ in (None, 2 * 97):
        say("stretching", "wrapped" if wrapped else "lifted", "monitor" if mon else "alone", "replace", replace,
            "m", m, "N", N, "launch", L_)
        res = propagate(sysm, x0, N, k=2, tau=tau, seed=21, reducers=reds, W0=W0, launch_steps=L_, check=False,
                        keep_states=True)
        assert np.array_equal(res.code, orc["code"]) and np.array_equal(res.hstep, orc["hstep"])
        assert np.array_equal(res.states, orc["states"])
        assert np.array_equal(res.workspace, orc["ws"])
        ok = orc["code"] == 0
        assert np.array_equal(res.x[ok].view(np.uint64), orc["x"][ok].view(np.uint64))
        assert np.array_equal(res.W[ok].view(np.uint64), orc["W"][ok].view(np.uint64))
        fin = finals_of(decode_states(orc["states"], allr, 2, 2), allr, 2, 2, orc["ws"])[0]
        assert np.array_equal(res.finals[0]["D"], fin["D"])


def test_host_chunk_with_workspace():
    """cl_propagate_chunk with host buffers in the reference layouts (ws[i][entries]), two launches carrying the
    workspace, equals the oracle's run."""
    from chaos_amd.tangent import _context
    m, N, tau = 600, 900, 3
    sysm = StandardMap(1.3)
    x0 = ensemble(m, False, 8)
    W0 = O.frames(4, 1, m, 2, 2)
    sr = StretchingNumbers(range=stretching_range(sysm), bins=64, interval=3)
    chain, allr = make_chain((sr,))
    nws = sr.workspace(2, 2)
    say("stretching host chunk", "m", m, "N", N)
    lib, ctx = _context()
    assert lib.cl_workspace_entries(ctypes.byref(chain), 2, 2) == nws
    sysd = sysm.descriptor()
    x = np.ascontiguousarray(x0.copy())
    W = np.ascontiguousarray(W0.transpose(0, 2, 1).copy())
    nb = lib.cl_state_bytes(ctypes.byref(sysd), ctypes.byref(chain), L.CL_ELT_F64, 2, 2)
    st = np.zeros((m, nb), dtype=np.uint8)
    code = np.zeros(m, dtype=np.int8)
    h = np.zeros(m, dtype=np.int64)
    ws = np.full((m, nws), 7, dtype=np.uint32)   # zeroed by the first launch
    for n0, steps in ((0, 300), (300, 600)):
        rc = lib.cl_propagate_chunk(ctx, ctypes.byref(sysd), ctypes.byref(chain), L.CL_ELT_F64, m, 2, 2, 2, n0, steps,
                                    N, tau, x.ctypes.data, W.ctypes.data, None, st.ctypes.data, code.ctypes.data,
                                    h.ctypes.data, ws.ctypes.data)
        assert rc == 0, lib.cl_last_error(ctx).decode()
    orc = O.run_ws(sysd, chain, nws, x0, W0, N, tau, 2)
    assert np.array_equal(ws, orc["ws"]) and np.array_equal(st, orc["states"]) and np.array_equal(code, orc["code"])


def test_refuses_workspace_chain_without_workspace():
    """cl_propagate_device refuses a chain with a workspace reducer; cl_propagate_device_ws refuses one without."""
    from chaos_amd.tangent import _context
    import torch
    lib, ctx = _context()
    sysm = StandardMap(1.0)
    sysd = sysm.descriptor()
    sr = StretchingNumbers(range=(-1, 1), bins=8)
    chain, _ = make_chain((sr,))
    say("stretching refusals")
    dx = torch.zeros((2, 4), dtype=torch.float64, device="cuda")
    dW = torch.zeros((4, 4), dtype=torch.float64, device="cuda")
    dst = torch.zeros((13, 4), dtype=torch.int64, device="cuda")
    dc = torch.zeros(4, dtype=torch.int8, device="cuda")
    dh = torch.zeros(4, dtype=torch.int64, device="cuda")
    rc = lib.cl_propagate_device(ctx, ctypes.byref(sysd), ctypes.byref(chain), 1, 4, 2, 2, 2, 0, 1, 1, 1,
                                 dx.data_ptr(), dW.data_ptr(), None, dst.data_ptr(), dc.data_ptr(), dh.data_ptr(), None)
    assert rc == L.CL_ERR_ARGUMENT
    chain2, _ = make_chain(())
    rc = lib.cl_propagate_device_ws(ctx, ctypes.byref(sysd), ctypes.byref(chain2), 1, 4, 2, 2, 2, 0, 1, 1, 1,
                                    dx.data_ptr(), dW.data_ptr(), None, dst.data_ptr(), dc.data_ptr(), dh.data_ptr(),
                                    dst.data_ptr(), None)
    assert rc == L.CL_ERR_ARGUMENT
    torch.cuda.synchronize()


@pytest.mark.parametrize("block,bins", [(128, 100), (256, 200), (64, 500)])
def test_staged_and_direct_counts(block, bins):
    """The counts are staged in 8-bit LDS counters (flushed every 255 handoffs and at the launch's end) when
    2 (bins + 2) bytes per lane fit in 64 KiB per group, and go to the workspace directly otherwise: 128 lanes x 100
    cells and 64 x 500 stage, 256 x 200 does not. Both equal the oracle bit for bit over 600 handoffs, a launch
    split included."""
    from chaos_amd.autotune import set_tuning
    m, N = 4096, 600
    sysm = StandardMap(2.5)
    x0 = ensemble(m, False, 17)
    W0 = O.frames(5, 1, m, 2, 2)
    sr = StretchingNumbers(range=stretching_range(sysm), bins=bins, interval=1)
    chain, _ = make_chain((sr,))
    orc = O.run_ws(sysm.descriptor(), chain, sr.workspace(2, 2), x0, W0, N, 1, 2)
    set_tuning(block=block)
    try:
        for L_ in (None, 260):
            say("stretching staging", "block", block, "bins", bins, "launch", L_)
            res = propagate(sysm, x0, N, k=2, tau=1, seed=5, reducers=(sr,), W0=W0, launch_steps=L_, check=False,
                            keep_states=True)
            assert np.array_equal(res.workspace, orc["ws"]) and np.array_equal(res.states, orc["states"])
            assert np.array_equal(res.code, orc["code"]) and np.array_equal(res.hstep, orc["hstep"])
    finally:
        set_tuning()


def test_full_size_sample():
    """2^20 orbits of the wrapped map at K = 6 (the study's wrapped mode), 1000 steps, tau = 1, 100 cells: a sample
    of 2048 orbits equals the oracle bit for bit; every column's counts total the handoffs; the spectrum's mean is
    the finite-time exponent near ln(K / 2)."""
    m, N = 1 << 20, 1000
    sysm = WrappedStandardMap(6.0)
    rng = np.random.default_rng(99)
    x0 = np.array([torus_state(v) for v in 2 * math.pi * rng.random((m, 2))])
    sr = StretchingNumbers(range=stretching_range(sysm), bins=100, interval=1)
    say("stretching full size", "m", m, "N", N)
    res = propagate(sysm, x0, N, k=2, tau=1, seed=7, reducers=(sr,), check=False, keep_states=True)
    assert np.all(res.code == 0)
    assert np.all(res.workspace.reshape(m, 2, 102).sum(axis=2) == N)
    idx = np.sort(rng.choice(m, 2048, replace=False))
    chain, _ = make_chain((sr,))
    orc = O.run_ws(sysm.descriptor(), chain, sr.workspace(2, 2), x0[idx], O.frames(7, 1, m, 2, 2)[idx], N, 1, 2)
    # frames are keyed by orbit id: the sample's ids are idx + 1
    assert np.array_equal(res.workspace[idx], orc["ws"]) and np.array_equal(res.states[idx], orc["states"])
    assert abs(np.median(res.finals[0]["mean"][:, 0]) - math.log(3.0)) < 0.1
