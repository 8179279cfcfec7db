"""cl_propagate_mixed_noalloc, the entry point that allocates no device memory (its staging lives in the library's
scratch array), against cl_propagate_mixed on the same lattices, bit for bit: two calls in a row, the second smaller
(the scratch keeps the first call's contents, so anything stale would show), and the size limit."""
import ctypes
import sys

import numpy as np
import pytest

from chaos_amd import propagate_mixed, StandardMapChain, QRSpectrum, HouseholderQR, _lib as L
from chaos_amd.tangent import _context, make_chain

pytestmark = pytest.mark.gpu
K = 16


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _bind():
    return L.load()


def lattices(seed, sizes, lo, hi, tau):
    rng = np.random.default_rng(seed)
    return [(N, rng.uniform(0, 2 * np.pi, (m, 2 * N)), tau * rng.integers(lo // tau, hi // tau + 1, m))
            for N, m in sizes]


def same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x.x, y.x) and np.array_equal(x.W, y.W) and np.array_equal(x.states, y.states)
        assert np.array_equal(x.code, y.code) and np.array_equal(x.hstep, y.hstep)


def test_noalloc_equals_allocating_entry_point_twice():
    _bind()
    say("noalloc call 1 (5 lattices)")
    big = lattices(21, [(8, 40), (16, 30), (32, 20), (64, 10), (128, 4)], 50, 300, 10)
    ref = propagate_mixed(big, K=1.5, eps=0.1, k=K, tau=10, seed=2, chunk=100, keep_states=True)
    got = propagate_mixed(big, K=1.5, eps=0.1, k=K, tau=10, seed=2, chunk=100, keep_states=True, _noalloc=True)
    same(ref, got)
    say("noalloc call 2 (smaller)")
    small = lattices(22, [(16, 5), (128, 1)], 20, 120, 10)
    ref = propagate_mixed(small, K=1.5, eps=0.1, k=K, tau=10, seed=3, chunk=0, keep_states=True)
    got = propagate_mixed(small, K=1.5, eps=0.1, k=K, tau=10, seed=3, chunk=0, keep_states=True, _noalloc=True)
    same(ref, got)


def test_noalloc_rejects_oversize_work_list():
    """About 2e5 orbits fit the 4 MiB scratch; 3e5 orbits of N = 8 are refused before any device work."""
    lib = _bind()
    _, ctx = _context()
    say("noalloc oversize")
    m = 300000
    Ns = np.full(m, 10, dtype=np.int64)
    lat = (L.cl_lattice * 1)()
    lat[0].N, lat[0].m, lat[0].Ns = 8, m, Ns.ctypes.data
    lat[0].dx = lat[0].dW = lat[0].dstates = lat[0].dcode = lat[0].dhstep = 8  # never dereferenced: refused first
    sysd = StandardMapChain(8, 1.5, 0.1, "ring").descriptor()
    ch, _ = make_chain((QRSpectrum(HouseholderQR()),))
    rc = lib.cl_propagate_mixed_noalloc(ctx, ctypes.byref(sysd), ctypes.byref(ch), L.CL_ELT_F64, 1, lat, K, 10, 0, None)
    assert rc == L.CL_ERR_ARGUMENT
    assert b"scratch" in lib.cl_last_error(ctx)
