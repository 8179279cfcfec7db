"""The persistent mixed-lattice launch (cl_propagate_mixed) on the GPU: every lattice's result equals the per-lattice
ragged kernel and the oracle bit for bit, for one chunk and for graph-captured chunks (and the same chunks launched
directly), with retired orbits (reached counts, a non-finite start) skipped by the later chunks; the symplectic
pairing residual at N = 8, k = 2N (src/Diagnostics/residuals.jl:16-24, src/Diagnostics/monitor.jl:42-52)."""
import os
import sys

import numpy as np
import pytest

from chaos_amd import StandardMapChain, QRSpectrum, HouseholderQR, propagate_ragged, propagate_mixed, autotune
from chaos_amd.tangent import make_chain
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu
K = 16
R = (QRSpectrum(HouseholderQR()),)


@pytest.fixture(params=[(1, O.RED_BLOCKED), (2, O.RED_SITES)], ids=["blocked", "sites"])
def order(request):
    """The chain's QR order on the context (cl_tuning.chain_qr) and the oracle mode that restates it."""
    qr, mode = request.param
    autotune.set_tuning(chain_qr=qr)
    yield mode
    autotune.set_tuning()


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def lattices(seed, sizes, lo, hi, tau):
    rng = np.random.default_rng(seed)
    out = []
    for N, m in sizes:
        x0 = rng.uniform(0, 2 * np.pi, (m, 2 * N))
        Ns = tau * rng.integers(lo // tau, hi // tau + 1, m)
        out.append((N, x0, Ns))
    return out


def check_equal(a, b):
    assert np.array_equal(a.x, b.x, equal_nan=True) and np.array_equal(a.W, b.W, equal_nan=True)
    assert np.array_equal(a.states, b.states)
    assert np.array_equal(a.code, b.code) and np.array_equal(a.hstep, b.hstep)


@pytest.mark.parametrize("boundary", ["ring", "open"])
@pytest.mark.parametrize("chunk", [0, 60, 130])
def test_mixed_equals_ragged_and_oracle(boundary, chunk, order):
    say("mixed", boundary, chunk, order)
    lat = lattices(7, [(8, 11), (16, 9), (32, 7), (64, 5), (128, 3)], 30, 400, 10)
    got = propagate_mixed(lat, K=1.5, eps=0.1, boundary=boundary, k=K, tau=10, seed=3, chunk=chunk, keep_states=True)
    for (N, x0, Ns), g in zip(lat, got):
        sysm = StandardMapChain(N, 1.5, 0.1, boundary)
        r = propagate_ragged(sysm, x0, Ns, k=K, tau=10, seed=3, reducers=R, keep_states=True)
        check_equal(g, r)
        assert np.array_equal(g.hstep, Ns)
        if N <= 32:
            W0 = O.frames(3, 1, len(x0), 2 * N, K)
            o = O.run_chain(sysm.descriptor(), make_chain(R)[0], x0, W0, 0, 10, K, Ns=Ns, mode=order)
            assert np.array_equal(g.x, o["x"]) and np.array_equal(g.W, o["W"]) and np.array_equal(g.states, o["states"])


def test_mixed_graph_replays_equal_direct_launches():
    """Chunks captured in a graph and the same chunks launched directly give the same bits; a repeated call too."""
    say("mixed graph vs direct")
    lat = lattices(9, [(8, 6), (64, 3), (128, 2)], 20, 300, 10)
    a = propagate_mixed(lat, K=1.5, eps=0.1, k=K, tau=10, seed=4, chunk=50, keep_states=True)
    os.environ["CL_GRAPH"] = "0"
    try:
        b = propagate_mixed(lat, K=1.5, eps=0.1, k=K, tau=10, seed=4, chunk=50, keep_states=True)
    finally:
        del os.environ["CL_GRAPH"]
    c = propagate_mixed(lat, K=1.5, eps=0.1, k=K, tau=10, seed=4, chunk=50, keep_states=True)
    for x, y, z in zip(a, b, c):
        check_equal(x, y)
        check_equal(x, z)


def test_mixed_retires_failed_orbit(order):
    """A non-finite start fails at its first handoff (code 1, the reference's state status) and is skipped after;
    its neighbours in the same work item run on unchanged."""
    say("mixed failure")
    lat = lattices(11, [(16, 8), (128, 2)], 40, 200, 10)
    N, x0, Ns = lat[0]
    x0 = x0.copy()
    x0[3, 5] = np.nan
    lat[0] = (N, x0, Ns)
    got = propagate_mixed(lat, K=1.5, eps=0.1, k=K, tau=10, seed=6, chunk=40, keep_states=True, check=False)
    r = propagate_ragged(StandardMapChain(N, 1.5, 0.1, "ring"), x0, Ns, k=K, tau=10, seed=6, reducers=R,
                         keep_states=True, check=False)
    check_equal(got[0], r)
    assert got[0].code[3] == 1 and got[0].hstep[3] == 10
    assert np.all(np.delete(got[0].code, 3) == 0)


def test_mixed_pairing_residual_n8():
    """N = 8, k = 16 = 2N: the full spectrum; its sum vanishes to rounding (det M = 1) and the pairing residual
    max_i |lambda_i + lambda_{2N+1-i}| falls as 1/n: below 16/n on every orbit and 4/n on average (the oracle gives
    5.9/n at most on 16 orbits at n = 2000 and at n = 20000, so the constant is the orbit's, not a drift)."""
    say("mixed pairing")
    lat = lattices(13, [(8, 64)], 2000, 2000, 10)
    g = propagate_mixed(lat, K=1.5, eps=0.1, k=K, tau=10, seed=1, chunk=500)[0]
    lam = g.finals[0]["lambda"]
    assert np.max(np.abs(lam.sum(axis=1))) < 1e-9
    pair = np.max(np.abs(lam + lam[:, ::-1]), axis=1)
    assert np.max(pair) < 16.0 / 2000 and np.mean(pair) < 4.0 / 2000


def test_mixed_f32tan_equals_ragged(order):
    """Float32 tangents in the persistent launch: every lattice equals its ragged kernel bit for bit."""
    say("mixed f32tan", order)
    lat = lattices(17, [(8, 9), (16, 7), (32, 5), (64, 3), (128, 2)], 30, 300, 10)
    got = propagate_mixed(lat, K=1.5, eps=0.1, k=K, tau=10, seed=5, chunk=100, keep_states=True, elt="f32tan")
    for (N, x0, Ns), g in zip(lat, got):
        r = propagate_ragged(StandardMapChain(N, 1.5, 0.1, "ring"), x0, Ns, k=K, tau=10, seed=5, reducers=R,
                             keep_states=True, elt="f32tan")
        check_equal(g, r)
