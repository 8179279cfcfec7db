"""The FP32-tangent element type (CL_ELT_F64_F32TAN) on the GPU against the oracle's FP32-tangent policy, bit for bit:
the standard map, the wrapped map and the per-orbit-K map with the QR spectra, the indicator chains and the fused
two-frame pass (split passes and the fused kernel), launch splits, and the host chunk entry point."""
import ctypes
import os
import sys

import numpy as np
import pytest

from chaos_amd import StandardMap, WrappedStandardMap, StandardMapK, propagate, _lib as L
from chaos_amd.reducers import QRSpectrum, MGS, HouseholderQR, MEGNO, FLI, SALI, OFLI, Benettin, GALI
from chaos_amd.tangent import make_chain, _context
from chaos_amd.inputs import grid_centres
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu
F32 = L.CL_ELT_F64_F32TAN


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def starts(sys_, m, seed=3):
    x = grid_centres(int(np.ceil(np.sqrt(m))))[:m]
    if sys_.state_length == 4:
        x = np.concatenate([x, np.zeros_like(x)], axis=1)
    if sys_.state_length == 3:
        x = np.concatenate([x, np.random.default_rng(seed).uniform(0.5, 6.0, (m, 1))], axis=1)
    return x


def same(g, o, two=False):
    assert np.array_equal(g.x, o["x"], equal_nan=True)
    assert np.array_equal(g.W, o["W"], equal_nan=True)
    if two:
        assert np.array_equal(g.W2, o["W2"], equal_nan=True)
    assert np.array_equal(g.states, o["states"])
    assert np.array_equal(g.code, o["code"]) and np.array_equal(g.hstep, o["hstep"])


CHAINS = [
    ((QRSpectrum(MGS()),), (), 2),
    ((QRSpectrum(HouseholderQR()),), (), 2),
    ((MEGNO(),), (), 2),
    ((SALI(),), (), 2),
    ((GALI(),), (), 2),
    ((MEGNO(replace=False), FLI()), (), 2),
    ((MEGNO(replace=False), FLI(replace=False), OFLI(replace=False), SALI()), (), 2),
    ((MEGNO(replace=False), FLI(replace=False), OFLI(replace=False), SALI()), (QRSpectrum(HouseholderQR()),), 2),
    ((MEGNO(replace=False), FLI(replace=False)), (QRSpectrum(MGS()),), 2),
    ((QRSpectrum(MGS()),), (), 1),
    ((Benettin(),), (), 1),
]


@pytest.mark.parametrize("ci", range(len(CHAINS)))
@pytest.mark.parametrize("sysname", ["std", "wrapped"])
def test_f32tan_bitwise(ci, sysname):
    red, red2, k = CHAINS[ci]
    sys_ = StandardMap(2.0) if sysname == "std" else WrappedStandardMap(2.0)
    say("f32tan", sysname, ci)
    m = 300
    x0 = starts(sys_, m)
    W0 = O.frames(5, 1, m, 2, k)
    g = propagate(sys_, x0, 700, k=k, tau=1, seed=5, reducers=red, reducers2=red2, W0=W0, keep_states=True,
                  check=False, elt="f32tan")
    ch, _ = make_chain(red, red2)
    o = O.run(sys_.descriptor(), ch, x0, W0, 700, 1, k, W2=W0 if red2 else None, elt=F32)
    same(g, o, bool(red2))
    assert g.provenance["precision"] == "Float64 orbit, Float32 tangents"


def test_f32tan_fused_pass_and_splits():
    """The two-frame pass through the split passes (default) and the fused kernel (CL_SPLIT=0) give the same bits; a
    run split into launches of 70 steps too (the Float32 frames travel through HBM as doubles, exactly)."""
    say("f32tan fused vs split")
    sys_ = StandardMap(6.0)
    red = (MEGNO(replace=False), FLI(replace=False), OFLI(replace=False), SALI())
    red2 = (QRSpectrum(HouseholderQR()),)
    x0 = starts(sys_, 500)
    a = propagate(sys_, x0, 1400, k=2, tau=1, seed=2, reducers=red, reducers2=red2, keep_states=True, elt="f32tan")
    os.environ["CL_SPLIT"] = "0"
    try:
        b = propagate(sys_, x0, 1400, k=2, tau=1, seed=2, reducers=red, reducers2=red2, keep_states=True, elt="f32tan")
    finally:
        del os.environ["CL_SPLIT"]
    c = propagate(sys_, x0, 1400, k=2, tau=1, seed=2, reducers=red, reducers2=red2, keep_states=True, elt="f32tan",
                  launch_steps=70)
    for y in (b, c):
        assert np.array_equal(a.states, y.states) and np.array_equal(a.W, y.W) and np.array_equal(a.W2, y.W2)


def test_f32tan_per_orbit_k_and_tau():
    say("f32tan per-orbit K, tau 3")
    sys_ = StandardMapK()
    x0 = starts(sys_, 257)
    red = (QRSpectrum(MGS()),)
    W0 = O.frames(9, 1, 257, 2, 2)
    g = propagate(sys_, x0, 900, k=2, tau=3, seed=9, reducers=red, W0=W0, keep_states=True, elt="f32tan")
    o = O.run(sys_.descriptor(), make_chain(red)[0], x0, W0, 900, 3, 2, elt=F32)
    same(g, o)


def test_f32tan_host_chunk_entry_point():
    """cl_propagate_chunk with the FP32-tangent type on host AoS buffers equals the oracle."""
    say("f32tan chunk")
    lib, ctx = _context()
    sys_ = StandardMap(1.2)
    red = (MEGNO(replace=False), FLI(replace=False), SALI())
    ch, _ = make_chain(red)
    m = 100
    x0 = starts(sys_, m)
    W0 = O.frames(4, 1, m, 2, 2)
    x = x0.copy()
    W = np.ascontiguousarray(W0.transpose(0, 2, 1))  # [m][k][d]
    nb = lib.cl_state_bytes(ctypes.byref(sys_.descriptor()), ctypes.byref(ch), F32, 2, 2)
    assert nb > 0
    st = np.zeros((m, nb), dtype=np.uint8)
    code = np.zeros(m, dtype=np.int8)
    hs = np.zeros(m, dtype=np.int64)
    rc = lib.cl_propagate_chunk(ctx, ctypes.byref(sys_.descriptor()), ctypes.byref(ch), F32, m, 2, 2, 2, 0, 600, 600, 1,
                                x.ctypes.data, W.ctypes.data, None, st.ctypes.data, code.ctypes.data, hs.ctypes.data,
                                None)
    assert rc == 0, lib.cl_last_error(ctx)
    o = O.run(sys_.descriptor(), ch, x0, W0, 600, 1, 2, elt=F32)
    assert np.array_equal(x, o["x"]) and np.array_equal(W.transpose(0, 2, 1), o["W"])
    assert np.array_equal(st, o["states"]) and np.array_equal(code, o["code"]) and np.array_equal(hs, o["hstep"])
