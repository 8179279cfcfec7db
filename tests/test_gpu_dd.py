"""The double-double (compensated) Henon-Heiles orbit, CL_ELT_DD, on the GPU against the oracle's compensated policy,
bit for bit: the C4 chain (MEGNO and GALI on one frame, QR(Givens) on the other), the QR spectrum alone with and
without the monitor, OFLI, launch splits and the host chunk entry point."""
import ctypes
import sys

import numpy as np
import pytest

from chaos_amd import HenonHeilesGauss, propagate, _lib as L
from chaos_amd.reducers import QRSpectrum, GivensQR, MEGNO, GALI, OFLI, ResidualMonitor
from chaos_amd.tangent import make_chain, _context
from chaos_amd.inputs import hh_section_starts
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu
DD = L.CL_ELT_DD
SYS = HenonHeilesGauss(h=0.01, tol=1e-15)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


CHAINS = [
    ((MEGNO(replace=False), GALI()), (QRSpectrum(GivensQR()),)),
    ((QRSpectrum(GivensQR()),), ()),
    ((MEGNO(replace=False), GALI()), ()),
    ((OFLI(SYS),), ()),
    ((), ()),
]


@pytest.mark.parametrize("ci", range(len(CHAINS)))
def test_dd_bitwise(ci):
    red, red2 = CHAINS[ci]
    say("dd", ci)
    m = 130
    z = np.concatenate([hh_section_starts(2, m // 2, 1 / 8), hh_section_starts(4, m - m // 2, 1 / 6)])
    W0 = O.frames(2, 1, m, 4, 4)
    g = propagate(SYS, z, 400, k=4, tau=10, seed=2, reducers=red, reducers2=red2, W0=W0, keep_states=True,
                  check=False, elt="dd")
    assert g.x.shape == (m, 8)
    ch, _ = make_chain(red, red2)
    zz = np.concatenate([z, np.zeros_like(z)], axis=1)
    o = O.run(SYS.descriptor(), ch, zz, W0, 400, 10, 4, W2=W0 if red2 else None, elt=DD)
    assert np.array_equal(g.x, o["x"], equal_nan=True) and np.array_equal(g.W, o["W"], equal_nan=True)
    if red2:
        assert np.array_equal(g.W2, o["W2"], equal_nan=True)
    assert np.array_equal(g.states, o["states"])
    assert np.array_equal(g.code, o["code"]) and np.array_equal(g.hstep, o["hstep"])
    assert g.provenance["precision"].startswith("compensated")


def test_dd_monitor_and_launch_split():
    say("dd monitor, split")
    m = 64
    z = hh_section_starts(8, m, 1 / 12)
    red = (QRSpectrum(GivensQR()), ResidualMonitor(SYS, 1e3, 1e3))
    a = propagate(SYS, z, 600, k=4, tau=10, seed=8, reducers=red, keep_states=True, elt="dd")
    b = propagate(SYS, z, 600, k=4, tau=10, seed=8, reducers=red, keep_states=True, elt="dd", launch_steps=200)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.W, b.W) and np.array_equal(a.states, b.states)
    ch, _ = make_chain(red)
    W0 = O.frames(8, 1, m, 4, 4)
    o = O.run(SYS.descriptor(), ch, np.concatenate([z, np.zeros_like(z)], axis=1), W0, 600, 10, 4, elt=DD)
    assert np.array_equal(a.states, o["states"]) and np.array_equal(a.x, o["x"])


def test_dd_host_chunk_entry_point():
    """cl_propagate_chunk with the DD type: host AoS state [m][8] (value, correction)."""
    say("dd chunk")
    lib, ctx = _context()
    red = (QRSpectrum(GivensQR()),)
    ch, _ = make_chain(red)
    m = 40
    z = hh_section_starts(9, m, 1 / 8)
    x = np.ascontiguousarray(np.concatenate([z, np.zeros_like(z)], axis=1))
    x0 = x.copy()
    W0 = O.frames(9, 1, m, 4, 4)
    W = np.ascontiguousarray(W0.transpose(0, 2, 1))
    nb = lib.cl_state_bytes(ctypes.byref(SYS.descriptor()), ctypes.byref(ch), DD, 4, 4)
    st = np.zeros((m, nb), dtype=np.uint8)
    code = np.zeros(m, dtype=np.int8)
    hs = np.zeros(m, dtype=np.int64)
    rc = lib.cl_propagate_chunk(ctx, ctypes.byref(SYS.descriptor()), ctypes.byref(ch), DD, m, 8, 4, 4, 0, 300, 300, 10,
                                x.ctypes.data, W.ctypes.data, None, st.ctypes.data, code.ctypes.data, hs.ctypes.data,
                                None)
    assert rc == 0, lib.cl_last_error(ctx)
    o = O.run(SYS.descriptor(), ch, x0, W0, 300, 10, 4, elt=DD)
    assert np.array_equal(x, o["x"]) and np.array_equal(st, o["states"])
