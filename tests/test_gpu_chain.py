"""The coupled standard-map chain kernels on the GPU against the oracle, bit for bit, for every built N, both
boundaries, launch splits, per-orbit step counts and the host chunk entry point, under both orders of the
HouseholderQR: the blocked factorisation on the matrix cores (the oracle's RED_BLOCKED) and the lane-butterfly sweep
(RED_SITES)."""
import ctypes
import sys

import numpy as np
import pytest

from chaos_amd import StandardMapChain, QRSpectrum, HouseholderQR, propagate, propagate_ragged, autotune, _lib as L
from chaos_amd.tangent import make_chain, _context
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu
K = 16


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


def starts(seed, m, N):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 2 * np.pi, (m, 2 * N))


@pytest.mark.parametrize("ring", [True, False], ids=["ring", "open"])
@pytest.mark.parametrize("N", [8, 16, 32, 64, 128])
def test_chain_qr_bitwise(N, ring, order):
    m = 13 if N <= 64 else 3
    say("chain qr", N, ring, m, order)
    sysm = StandardMapChain(N, 1.5, 0.1, "ring" if ring else "open")
    x0 = starts(N + ring, m, N)
    W0 = O.frames(5, 1, m, 2 * N, K)
    chain, _ = make_chain((QRSpectrum(HouseholderQR()),))
    g = propagate(sysm, x0, 300, k=K, tau=10, seed=5, reducers=(QRSpectrum(HouseholderQR()),), W0=W0,
                  keep_states=True)
    o = O.run_chain(sysm.descriptor(), chain, x0, W0, 300, 10, K, mode=order)
    assert np.array_equal(g.x, o["x"])
    assert np.array_equal(g.W, o["W"])
    assert np.array_equal(g.states, o["states"])
    assert np.array_equal(g.code, o["code"]) and np.array_equal(g.hstep, o["hstep"])
    lam = g.finals[0]["lambda"]
    assert np.all(lam[:, 0] > 0.05)


@pytest.mark.parametrize("N", [8, 128])
def test_chain_no_reducer_and_launch_split(N, order):
    say("chain split", N, order)
    sysm = StandardMapChain(N, 1.5, 0.1, "ring")
    x0 = starts(3, 5, N)
    W0 = O.frames(3, 1, 5, 2 * N, K)
    g = propagate(sysm, x0, 200, k=K, tau=10, seed=3, W0=W0)
    o = O.run_chain(sysm.descriptor(), make_chain(())[0], x0, W0, 200, 10, K)
    assert np.array_equal(g.x, o["x"]) and np.array_equal(g.W, o["W"])
    r = (QRSpectrum(HouseholderQR()),)
    a = propagate(sysm, x0, 200, k=K, tau=10, seed=3, reducers=r, W0=W0, keep_states=True)
    b = propagate(sysm, x0, 200, k=K, tau=10, seed=3, reducers=r, W0=W0, keep_states=True, launch_steps=70)
    assert np.array_equal(a.states, b.states) and np.array_equal(a.W, b.W) and np.array_equal(a.x, b.x)


@pytest.mark.parametrize("N", [16, 128])
def test_chain_ragged_bitwise(N, order):
    say("chain ragged", N, order)
    m = 9
    sysm = StandardMapChain(N, 1.5, 0.1, "ring")
    x0 = starts(11, m, N)
    W0 = O.frames(11, 1, m, 2 * N, K)
    Ns = 10 * np.random.default_rng(1).integers(3, 40, m)
    r = (QRSpectrum(HouseholderQR()),)
    g = propagate_ragged(sysm, x0, Ns, k=K, tau=10, seed=11, reducers=r, W0=W0, keep_states=True)
    o = O.run_chain(sysm.descriptor(), make_chain(r)[0], x0, W0, 0, 10, K, Ns=Ns, mode=order)
    assert np.array_equal(g.x, o["x"]) and np.array_equal(g.W, o["W"])
    assert np.array_equal(g.states, o["states"]) and np.array_equal(g.hstep, Ns)


def test_chain_origin_spectrum(order):
    """At the origin the chain's spectrum is +-ln mu_j of its modes (see tests/test_oracle_chain.py)."""
    from tests.test_oracle_chain import origin_exponents
    N = 32
    say("chain origin", N, order)
    sysm = StandardMapChain(N, 1.5, 0.1, "ring")
    g = propagate(sysm, np.zeros((2, 2 * N)), 4000, k=K, tau=10, seed=2, reducers=(QRSpectrum(HouseholderQR()),))
    assert np.all(g.x == 0)
    want = origin_exponents(N, 1.5, 0.1, True)[:K]
    assert np.max(np.abs(g.finals[0]["lambda"] - want)) < 10.0 / 4000


@pytest.mark.parametrize("d,k", [(6, 3), (16, 16), (48, 5), (128, 16), (256, 16)])
def test_device_frames_of_chain_sizes_equal_host(d, k):
    """The frames of d > 4 (one 64-lane group per orbit, the frame in LDS) equal the host rule bit for bit."""
    import torch
    m = 37
    say("chain frames", d, k)
    lib, ctx = _context()
    dW = torch.empty((m, k, d), dtype=torch.float64, device="cuda")
    assert lib.cl_initial_deviations_device(ctx, 11, 5, m, d, k, dW.data_ptr(), None) == 0
    torch.cuda.synchronize()
    hW = np.empty((m, k, d))
    assert lib.cl_initial_deviations(11, 5, m, d, k, hW.ctypes.data) == 0
    assert np.array_equal(dW.cpu().numpy(), hW)


def test_chain_host_chunk_and_frames(order):
    import torch
    N, m = 64, 4
    say("chain chunk", N, order)
    lib, ctx = _context()
    sysm = StandardMapChain(N, 1.5, 0.1, "open")
    sd = sysm.descriptor()
    chain, _ = make_chain((QRSpectrum(HouseholderQR()),))
    d = 2 * N
    # device frames equal the host rule
    dW = torch.empty((m, K, d), dtype=torch.float64, device="cuda")
    assert lib.cl_initial_deviations_device(ctx, 7, 1, m, d, K, dW.data_ptr(), None) == 0
    torch.cuda.synchronize()
    hW = np.empty((m, K, d))
    assert lib.cl_initial_deviations(7, 1, m, d, K, hW.ctypes.data) == 0
    assert np.array_equal(dW.cpu().numpy(), hW)
    x0 = starts(4, m, N)
    x = x0.copy()
    W = hW.copy()
    words = 2 * K + 9
    st = np.zeros((m, words), dtype=np.uint64)
    code = np.zeros(m, dtype=np.int8)
    h = np.zeros(m, dtype=np.int64)
    rc = lib.cl_propagate_chunk(ctx, ctypes.byref(sd), ctypes.byref(chain), 1, m, d, d, K, 0, 100, 100, 10,
                                x.ctypes.data, W.ctypes.data, None, st.ctypes.data, code.ctypes.data, h.ctypes.data,
                                None)
    assert rc == 0
    o = O.run_chain(sd, chain, x0, hW.transpose(0, 2, 1), 100, 10, K, mode=order)
    assert np.array_equal(x, o["x"]) and np.array_equal(W.transpose(0, 2, 1), o["W"])
    assert np.array_equal(st.view(np.uint8).reshape(m, -1), o["states"])


@pytest.mark.parametrize("ring", [True, False], ids=["ring", "open"])
@pytest.mark.parametrize("N", [8, 16, 32, 64, 128])
def test_chain_f32tan_bitwise(N, ring, order):
    """Float32 tangents (elt f32tan): the Float64 orbit, the frame in Float32 between handoffs, each factorisation in
    FP64 on the widened frame; bit for bit against the oracle's policy (elt 2), launch splits included."""
    m = 7 if N <= 64 else 3
    say("chain f32tan", N, ring, m, order)
    sysm = StandardMapChain(N, 1.5, 0.1, "ring" if ring else "open")
    x0 = starts(50 + N, m, N)
    W0 = O.frames(9, 1, m, 2 * N, K)
    r = (QRSpectrum(HouseholderQR()),)
    g = propagate(sysm, x0, 300, k=K, tau=10, seed=9, reducers=r, W0=W0, keep_states=True, elt="f32tan")
    o = O.run_chain(sysm.descriptor(), make_chain(r)[0], x0, W0, 300, 10, K, mode=order, elt=2)
    assert np.array_equal(g.x, o["x"]) and np.array_equal(g.W, o["W"])
    assert np.array_equal(g.states, o["states"])
    assert np.array_equal(g.code, o["code"]) and np.array_equal(g.hstep, o["hstep"])
    assert g.provenance["precision"] == "Float64 orbit, Float32 tangents"
    if N in (8, 128):
        s = propagate(sysm, x0, 300, k=K, tau=10, seed=9, reducers=r, W0=W0, keep_states=True, elt="f32tan",
                      launch_steps=70)
        assert np.array_equal(s.states, g.states) and np.array_equal(s.W, g.W)


@pytest.mark.parametrize("N", [16, 128])
def test_chain_f32tan_ragged(N, order):
    say("chain f32tan ragged", N, order)
    m = 9
    sysm = StandardMapChain(N, 1.5, 0.1, "ring")
    x0 = starts(12, m, N)
    W0 = O.frames(12, 1, m, 2 * N, K)
    Ns = 10 * np.random.default_rng(2).integers(3, 40, m)
    r = (QRSpectrum(HouseholderQR()),)
    g = propagate_ragged(sysm, x0, Ns, k=K, tau=10, seed=12, reducers=r, W0=W0, keep_states=True, elt="f32tan")
    o = O.run_chain(sysm.descriptor(), make_chain(r)[0], x0, W0, 0, 10, K, Ns=Ns, mode=order, elt=2)
    assert np.array_equal(g.x, o["x"]) and np.array_equal(g.W, o["W"])
    assert np.array_equal(g.states, o["states"]) and np.array_equal(g.hstep, Ns)
