"""GPU parity: the HIP kernels against the CPU oracle on the same seeded inputs.

With the kernels' sin/cos/log evaluated bit-identically on host and device, the oracle (MATH_CL
mode) and the GPU must agree bit for bit on final points, frames, every reducer-state byte, status
codes and failing steps, for chaotic orbits too. Every test prints its case before touching the GPU."""
import ctypes
import math
import sys

import numpy as np
import pytest

from chaos_amd import (StandardMap, WrappedStandardMap, QRSpectrum, CGS, MGS, RGS, HouseholderQR, AugmentedQR,
                       GivensQR, Benettin, MEGNO, SALI, FLI, OFLI, GALI, ResidualMonitor, propagate, NumericalHealthError,
                       torus_state)
from chaos_amd import _lib as L
from chaos_amd.inputs import torus_uniform
from chaos_amd.tangent import make_chain, decode_states, finals_of, _context
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def sysd(kind, K):
    s = L.cl_system()
    s.kind = kind
    s.K = K
    return s


CHAINS_K2 = [
    (), (QRSpectrum(CGS()),), (QRSpectrum(MGS()),), (QRSpectrum(RGS()),), (QRSpectrum(HouseholderQR()),),
    (QRSpectrum(AugmentedQR()),), (QRSpectrum(GivensQR()),), (MEGNO(),), (MEGNO(replace=False),), (SALI(),), (FLI(),),
    (OFLI(),), (GALI(),), (MEGNO(replace=False), SALI()), (FLI(replace=False), SALI()),
    (MEGNO(replace=False), FLI(replace=False), SALI()),
    (MEGNO(replace=False), FLI(replace=False), OFLI(replace=False), SALI()),
]
CHAINS_K1 = [(), (Benettin(),), (QRSpectrum(CGS()),), (QRSpectrum(MGS()),), (QRSpectrum(RGS()),),
             (QRSpectrum(HouseholderQR()),), (QRSpectrum(AugmentedQR()),), (QRSpectrum(GivensQR()),), (OFLI(),)]


def cname(c):
    return "+".join(type(r).__name__ + (f"({r.alg.name})" if hasattr(r, "alg") else "") for r in c) or "none"


def starts(system, m, seed):
    x = torus_uniform(seed, m)
    x[: m // 4] *= 40.0  # lifted points far from the torus for a quarter of the orbits
    if system is WrappedStandardMap:
        x = np.array([torus_state(z) for z in x])
    return x


def compare(res, orc, allr, d, k):
    assert np.array_equal(res.code, orc["code"]), "status codes differ"
    assert np.array_equal(res.hstep, orc["hstep"]), "failing steps differ"
    assert np.array_equal(res.x.view(np.uint64), orc["x"].view(np.uint64)), "final points differ"
    assert np.array_equal(res.W.view(np.uint64), orc["W"].view(np.uint64)), "frames differ"
    if "W2" in orc:
        assert np.array_equal(res.W2.view(np.uint64), orc["W2"].view(np.uint64)), "second frames differ"
    assert np.array_equal(res.states, orc["states"]), "reducer states differ"


@pytest.mark.parametrize("system", [StandardMap, WrappedStandardMap], ids=["std", "wrapped"])
@pytest.mark.parametrize("chain", CHAINS_K2, ids=cname)
def test_chain_bitwise_k2(system, chain):
    m, N = 384, 1500
    for K, tau in ((0.97, 1), (6.0, 3)):
        say("chain", cname(chain), system.__name__, "K", K, "tau", tau, "m", m, "N", N)
        x0 = starts(system, m, 17)
        res = propagate(system(K), x0, N, k=2, tau=tau, seed=99, reducers=chain, check=False, keep_states=True)
        ch, allr = make_chain(chain)
        W0 = O.frames(99, 1, m, 2, 2)
        orc = O.run(system(K).descriptor(), ch, x0, W0, N, tau, 2, math=O.MATH_CL)
        compare(res, orc, allr, 2, 2)


@pytest.mark.parametrize("chain", CHAINS_K1, ids=cname)
def test_chain_bitwise_k1(chain):
    m, N = 256, 2000
    say("chain k=1", cname(chain), "m", m, "N", N)
    x0 = starts(StandardMap, m, 5)
    res = propagate(StandardMap(2.5), x0, N, k=1, tau=2, seed=3, reducers=chain, check=False, keep_states=True)
    ch, allr = make_chain(chain)
    orc = O.run(StandardMap(2.5).descriptor(), ch, x0, O.frames(3, 1, m, 2, 1), N, 2, 1, math=O.MATH_CL)
    compare(res, orc, allr, 2, 1)


def test_two_frames_bitwise():
    """Indicators on frame 1 and a QR spectrum on frame 2 in one pass equal the oracle's two-frame run and, chain by
    chain, the separate single-frame GPU runs."""
    m, N = 512, 3000
    ind = (MEGNO(replace=False), FLI(replace=False), OFLI(replace=False), SALI())
    for qr in (QRSpectrum(HouseholderQR()), QRSpectrum(MGS())):
        say("two frames", qr.alg.name, "m", m, "N", N)
        x0 = starts(StandardMap, m, 8)
        res = propagate(StandardMap(2.0), x0, N, k=2, seed=4, reducers=ind, reducers2=(qr,), check=False, keep_states=True)
        ch, allr = make_chain(ind, (qr,))
        W0 = O.frames(4, 1, m, 2, 2)
        orc = O.run(StandardMap(2.0).descriptor(), ch, x0, W0, N, 1, 2, W2=W0, math=O.MATH_CL)
        compare(res, orc, allr, 2, 2)
        a = propagate(StandardMap(2.0), x0, N, k=2, seed=4, reducers=ind, check=False)
        b = propagate(StandardMap(2.0), x0, N, k=2, seed=4, reducers=(qr,), check=False)
        assert np.array_equal(res.W, a.W) and np.array_equal(res.W2, b.W)
        for f, g in zip(res.finals[:4], a.finals):
            for key in f:
                assert np.array_equal(np.asarray(f[key]), np.asarray(g[key])), key
        assert np.array_equal(res.finals[4]["lambda"], b.finals[0]["lambda"])


def test_launch_split_invariance():
    """Launches of L steps that carry x, W and the states give the single-launch result bitwise (D44,
    src/Tangent/propagate.jl:92-109)."""
    m, N = 300, 2400
    chain = (MEGNO(replace=False), FLI(replace=False), SALI())
    say("launch split", "m", m, "N", N)
    x0 = starts(StandardMap, m, 2)
    a = propagate(StandardMap(3.0), x0, N, k=2, tau=2, seed=1, reducers=chain, check=False, keep_states=True)
    for L_ in (2, 300, 1000):
        b = propagate(StandardMap(3.0), x0, N, k=2, tau=2, seed=1, reducers=chain, check=False, keep_states=True,
                      launch_steps=L_)
        assert b.provenance["launches_per_chunk"] == -(-N // L_)
        assert np.array_equal(a.states, b.states) and np.array_equal(a.W, b.W) and np.array_equal(a.x, b.x)


def test_host_chunk_entry_point():
    """cl_propagate_chunk on host buffers in the reference layouts, split over launches with n0 > 0, equals the
    device entry point bitwise."""
    m, N = 200, 1200
    chain = (QRSpectrum(MGS()),)
    say("host chunk API", "m", m, "N", N)
    lib, ctx = _context()
    x0 = starts(StandardMap, m, 6)
    ref = propagate(StandardMap(6.0), x0, N, k=2, seed=12, reducers=chain, check=False, keep_states=True)
    sd = StandardMap(6.0).descriptor()
    ch, allr = make_chain(chain)
    nb = lib.cl_state_bytes(ctypes.byref(sd), ctypes.byref(ch), 1, 2, 2)
    x = np.ascontiguousarray(x0.copy())
    W = np.ascontiguousarray(O.frames(12, 1, m, 2, 2).transpose(0, 2, 1))
    st = np.zeros((m, nb), np.uint8)
    code = np.zeros(m, np.int8)
    h = np.zeros(m, np.int64)
    for n0 in range(0, N, 400):
        rc = lib.cl_propagate_chunk(ctx, ctypes.byref(sd), ctypes.byref(ch), 1, m, 2, 2, 2, n0, 400, N, 1,
                                    x.ctypes.data, W.ctypes.data, None, st.ctypes.data, code.ctypes.data,
                                    h.ctypes.data, None)
        assert rc == 0, lib.cl_last_error(ctx)
    assert np.array_equal(st, ref.states)
    assert np.array_equal(x, ref.x) and np.array_equal(W.transpose(0, 2, 1), ref.W)


def test_device_math_and_frames_bitwise():
    """sin/cos/log and the seeded frames: device bits equal host bits."""
    import torch
    lib, ctx = _context()
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-20, 20, 200000), 10 ** rng.uniform(-300, 300, 200000),
                         -(10 ** rng.uniform(0, 18, 100000)), rng.uniform(0, 1e9, 100000)])
    say("device math", len(xs))
    hs, hc, hl = (np.empty_like(xs) for _ in range(3))
    lib.cl_math_host(len(xs), xs.ctypes.data, hs.ctypes.data, hc.ctypes.data, hl.ctypes.data)
    dx = torch.from_numpy(xs).cuda()
    ds, dc, dl = (torch.empty_like(dx) for _ in range(3))
    assert lib.cl_math_device(ctx, len(xs), dx.data_ptr(), ds.data_ptr(), dc.data_ptr(), dl.data_ptr(), None) == 0
    torch.cuda.synchronize()
    for h, d in ((hs, ds), (hc, dc), (hl, dl)):
        assert np.array_equal(h.view(np.uint64), d.cpu().numpy().view(np.uint64))
    for d_, k in ((2, 1), (2, 2)):
        say("device frames", d_, k)
        m = 100000
        dW = torch.empty((d_ * k, m), dtype=torch.float64, device="cuda")
        assert lib.cl_initial_deviations_device(ctx, 77, 1, m, d_, k, dW.data_ptr(), None) == 0
        torch.cuda.synchronize()
        host = np.empty((m, k, d_))
        lib.cl_initial_deviations(77, 1, m, d_, k, host.ctypes.data)
        assert np.array_equal(dW.cpu().numpy().T.reshape(m, k, d_), host)


def test_device_sqrt12_is_correctly_rounded():
    """The 2-entry norms' square root on [1, 2] (no range scaling, no special values) equals the host's IEEE square
    root bit for bit: 2^24 uniform arguments, the end points and their neighbours, the squares of 26-bit numbers
    (exact roots) and their neighbours, and arguments whose root lies within 2^-30 ulp of a rounding midpoint
    (x = RN(m^2) for the midpoints m = (2a + 1) 2^-53 of [1, sqrt 2): the hardest cases for a rounding decision)."""
    import torch
    lib, ctx = _context()
    rng = np.random.default_rng(12)
    u = 1.0 + rng.random(1 << 24)
    ends = np.array([1.0, 2.0, np.nextafter(1.0, 2.0), np.nextafter(2.0, 1.0), np.nextafter(np.nextafter(1.0, 2.0), 2.0)])
    r26 = 1.0 + rng.integers(0, 1 << 26, 200000) * 2.0 ** -26
    r26 = r26[r26 * r26 <= 2.0]
    sq = r26 * r26  # exact (26 + 26 bits)
    near = np.concatenate([sq, np.nextafter(sq, 0.0), np.nextafter(sq, 4.0)])
    # midpoints (2a + 1) 2^-53 of [1, sqrt 2) and x = RN(midpoint^2) by exact integer arithmetic
    odd = 2 * rng.integers(1 << 52, int((1 << 53) * 0.7071), 200000, dtype=np.int64) + 1
    hard = np.array([math.ldexp(float(int(o) * int(o)), -106) for o in odd])
    xs = np.concatenate([u, ends, near[(near >= 1.0) & (near <= 2.0)], hard[(hard >= 1.0) & (hard <= 2.0)]])
    say("device sqrt12", len(xs))
    dx = torch.from_numpy(xs).cuda()
    dy = torch.empty_like(dx)
    assert lib.cl_sqrt12_device(ctx, len(xs), dx.data_ptr(), dy.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(np.sqrt(xs).view(np.uint64), dy.cpu().numpy().view(np.uint64))


def test_k0_shear_closed_form_on_gpu():
    """tc_ac01_04.jl:52-64 on the device: K = 0, W0 = I: W_n = [[1,n],[0,1]], x_n = x0 + n p0 bitwise."""
    for n in (1, 1000, 10 ** 6):
        say("shear", n)
        r = propagate(StandardMap(0.0), [[0.25, 0.5]], n, k=2, seed=1, W0=[np.eye(2)])
        assert np.array_equal(r.W[0], [[1.0, float(n)], [0.0, 1.0]])
        assert np.array_equal(r.x[0], [0.25 + 0.5 * n, 0.5])


def test_health_errors_match_reference_rules():
    """A non-finite start raises NumericalHealthError('state') at the first handoff; a zero frame column stops a QR
    spectrum with an orthonormalisation breakdown at its pivot (src/Tangent/orbit.jl:53-57,
    src/Lyapunov/qr_spectrum.jl:44-60); codes and steps equal the oracle's."""
    say("health: non-finite start")
    x0 = np.array([[0.1, 0.2], [np.inf, 0.0], [0.3, 0.4]])
    with pytest.raises(NumericalHealthError) as e:
        propagate(StandardMap(1.0), x0, 10, k=2, seed=1, reducers=(QRSpectrum(MGS()),))
    assert e.value.what == "state" and e.value.orbit == 2 and e.value.step == 1
    say("health: zero column")
    W0 = np.array([np.eye(2), [[1.0, 0.0], [0.0, 0.0]], np.eye(2)])
    with pytest.raises(NumericalHealthError) as e:
        propagate(StandardMap(0.0), np.zeros((3, 2)), 10, k=2, seed=1, W0=W0, reducers=(QRSpectrum(MGS()),))
    assert e.value.what == "orthonormalisation breakdown" and e.value.orbit == 2 and e.value.column == 2
    r = propagate(StandardMap(0.0), np.zeros((3, 2)), 10, k=2, seed=1, W0=W0, reducers=(QRSpectrum(MGS()),),
                  check=False, keep_states=True)
    ch, allr = make_chain((QRSpectrum(MGS()),))
    orc = O.run(StandardMap(0.0).descriptor(), ch, np.zeros((3, 2)), W0, 10, 1, 2)
    compare(r, orc, allr, 2, 2)


def test_config1_bitwise_full():
    """The reference's CPU-runnable case at full size: standard map K = 6, 4096 starts uniform on the torus from seed
    1, 1e5 steps, k = 2, tau = 1, QRSpectrum(MGS()). GPU and oracle agree bit for bit on every orbit."""
    m, N = 4096, 100000
    say("config 1 full", m, N)
    x0 = torus_uniform(1, m)
    chain = (QRSpectrum(MGS()),)
    res = propagate(StandardMap(6.0), x0, N, k=2, tau=1, seed=1, reducers=chain, check=False, keep_states=True)
    ch, allr = make_chain(chain)
    orc = O.run(StandardMap(6.0).descriptor(), ch, x0, O.frames(1, 1, m, 2, 2), N, 1, 2, math=O.MATH_CL)
    compare(res, orc, allr, 2, 2)
    lam = res.finals[0]["lambda"]
    assert np.all(res.code == 0)
    # chaotic sea at K = 6: lambda_1 near the large-K law ln(K/2) (Chirikov), pairing lambda_1 + lambda_2 = 0
    assert abs(np.median(lam[:, 0]) - math.log(3.0)) < 0.1
    assert np.max(np.abs(lam[:, 0] + lam[:, 1])) < 1e-9


HH_CHAINS = [((), (), 4), ((QRSpectrum(GivensQR()),), (), 4), ((GALI(),), (), 4), ((MEGNO(replace=False),), (), 4),
             ((OFLI(),), (), 4), ((MEGNO(replace=False), GALI()), (QRSpectrum(GivensQR()),), 4),
             ((MEGNO(replace=False), FLI(replace=False), OFLI(replace=False), SALI()), (), 2), ((Benettin(),), (), 1)]


@pytest.mark.parametrize("chains", HH_CHAINS, ids=lambda c: cname(c[0] + c[1]) + f"-k{c[2]}")
def test_henon_heiles_gauss_bitwise(chains):
    """Henon-Heiles under the s = 3 Gauss method: the stage iteration, increment, step Jacobian and reducers
    on the GPU equal the oracle bit for bit."""
    from chaos_amd import HenonHeilesGauss
    from chaos_amd.inputs import hh_section_starts
    red, red2, k = chains
    m, N = 256, 400
    for E, tol in ((1 / 12, 1e-15), (1 / 6, None)):
        say("henon-heiles", cname(red + red2), "k", k, "E", E, "tol", tol)
        sysm = HenonHeilesGauss(h=0.01, tol=tol)
        x0 = hh_section_starts(5, m, E)
        res = propagate(sysm, x0, N, k=k, tau=10, seed=21, reducers=red, reducers2=red2, check=False,
                        keep_states=True)
        ch, allr = make_chain(red, red2)
        W0 = O.frames(21, 1, m, 4, k)
        orc = O.run(sysm.descriptor(), ch, x0, W0, N, 10, k, W2=W0 if red2 else None, math=O.MATH_CL)
        compare(res, orc, allr, 4, k)


@pytest.mark.parametrize("block", [0, 128])
def test_henon_heiles_large_groups_bitwise(block):
    """The Gauss kernel parks its frames in LDS slots sized by the group: at 65536 orbits the launch takes 256-lane
    groups (block 0: the built-in rule), or the 128 lanes a tuned setting asks for. The two-frame C4 chain (MEGNO +
    GALI | Givens QR spectrum) equals the oracle bit for bit on every 256th orbit."""
    from chaos_amd import HenonHeilesGauss
    from chaos_amd.autotune import set_tuning
    from chaos_amd.inputs import hh_section_starts
    red, red2 = (MEGNO(replace=False), GALI()), (QRSpectrum(GivensQR()),)
    m, N = 65536, 40
    say("henon-heiles large groups", "block", block, "m", m)
    sysm = HenonHeilesGauss(h=0.01, tol=1e-15)
    x0 = hh_section_starts(7, m, 1 / 8)
    set_tuning(block=block)
    try:
        res = propagate(sysm, x0, N, k=4, tau=10, seed=8, reducers=red, reducers2=red2, check=False, keep_states=True)
    finally:
        set_tuning()
    sel = np.arange(0, m, 256)
    ch, allr = make_chain(red, red2)
    W0 = O.frames(8, 1, m, 4, 4)[sel]
    orc = O.run(sysm.descriptor(), ch, x0[sel], W0, N, 10, 4, W2=W0, math=O.MATH_CL)
    assert np.array_equal(res.x[sel].view(np.uint64), orc["x"].view(np.uint64))
    assert np.array_equal(res.W[sel].view(np.uint64), orc["W"].view(np.uint64))
    assert np.array_equal(res.W2[sel].view(np.uint64), orc["W2"].view(np.uint64))
    assert np.array_equal(res.states[sel], orc["states"])
    assert np.array_equal(res.code[sel], orc["code"]) and np.array_equal(res.hstep[sel], orc["hstep"])


def test_split_passes_with_failures_equal_fused_and_oracle():
    """A two-frame chain runs as two single-frame passes; orbits that fail in either pass are rerun by the fused
    kernel. With failures planted in one frame at a time (a zero column on either frame, overflow on either frame,
    a non-finite start), the result equals the fused kernel's and the oracle's bit for bit, launch split or not."""
    import ctypes
    import os
    import torch
    from chaos_amd import _lib as L
    from chaos_amd.tangent import _context
    say("split with failures")
    ind = (MEGNO(replace=False), FLI(replace=False), OFLI(replace=False), SALI())
    qr = (QRSpectrum(HouseholderQR()),)
    ch, allr = make_chain(ind, qr)
    m, N = 6, 60
    x0 = starts(StandardMap, m, 21)
    x0[5] = [np.inf, 0.1]
    W0 = O.frames(21, 1, m, 2, 2).copy()
    W2 = W0.copy()
    W0[1][:, 1] = 0.0
    W2[2][:, 1] = 0.0
    W2[3] *= 1.5e308
    W0[4] *= 1.5e308
    sd = StandardMap(2.0).descriptor()
    lib, ctx = _context()
    words = lib.cl_state_bytes(ctypes.byref(sd), ctypes.byref(ch), 1, 2, 2) // 8

    def run(split, L_):
        os.environ["CL_SPLIT"] = "1" if split else "0"
        try:
            dx = torch.from_numpy(np.ascontiguousarray(x0.T)).cuda()
            dW = torch.from_numpy(np.ascontiguousarray(W0.transpose(2, 1, 0).reshape(4, m))).cuda()
            dW2 = torch.from_numpy(np.ascontiguousarray(W2.transpose(2, 1, 0).reshape(4, m))).cuda()
            dst = torch.zeros((words, m), dtype=torch.int64, device="cuda")
            dc = torch.zeros(m, dtype=torch.int8, device="cuda")
            dh = torch.zeros(m, dtype=torch.int64, device="cuda")
            n0 = 0
            while n0 < N:
                st = min(L_, N - n0)
                rc = lib.cl_propagate_device(ctx, ctypes.byref(sd), ctypes.byref(ch), 1, m, 2, 2, 2, n0, st, N, 1,
                                             dx.data_ptr(), dW.data_ptr(), dW2.data_ptr(), dst.data_ptr(),
                                             dc.data_ptr(), dh.data_ptr(), None)
                assert rc == 0
                n0 += st
            torch.cuda.synchronize()
            return dict(x=dx.cpu().numpy().T.copy(), W=dW.cpu().numpy().reshape(2, 2, m).transpose(2, 1, 0),
                        W2=dW2.cpu().numpy().reshape(2, 2, m).transpose(2, 1, 0),
                        states=dst.cpu().numpy().T.copy().view(np.uint8).reshape(m, -1), code=dc.cpu().numpy(),
                        hstep=dh.cpu().numpy())
        finally:
            os.environ.pop("CL_SPLIT", None)

    orc = O.run(sd, ch, x0, W0, N, 1, 2, W2=W2, math=O.MATH_CL)
    assert sorted(set(orc["code"].tolist())) == [0, 1, 2, 3]
    for split in (True, False):
        for L_ in (N, 7):
            g = run(split, L_)
            for key in ("x", "W", "W2", "states", "code", "hstep"):
                a, b = np.ascontiguousarray(g[key]), np.ascontiguousarray(orc[key])
                assert a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8)), (split, L_, key)


def test_split_and_fused_equal_for_the_gauss_flow():
    """The Henon-Heiles two-frame chain runs fused by default (its step dominates); forcing the split gives the
    same bits."""
    import os
    from chaos_amd import HenonHeilesGauss, GALI, GivensQR
    from chaos_amd.inputs import hh_section_starts
    say("gauss split vs fused")
    x0 = hh_section_starts(3, 64, 1 / 8)
    out = []
    for flag in ("0", "1"):
        os.environ["CL_SPLIT"] = flag
        try:
            out.append(propagate(HenonHeilesGauss(h=0.01, tol=1e-15), x0, 200, k=4, tau=10, seed=3,
                                 reducers=(MEGNO(replace=False), GALI()), reducers2=(QRSpectrum(GivensQR()),),
                                 check=False, keep_states=True))
        finally:
            os.environ.pop("CL_SPLIT", None)
    a, b = out
    assert np.array_equal(a.states, b.states) and np.array_equal(a.W, b.W) and np.array_equal(a.W2, b.W2)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.code, b.code)



MON_MAP_CASES = [(2, False, lambda K: (64.0, 1e-12)), (2, True, lambda K: (64.0, 1.75e-14 if K > 1 else 1.3e-15)),
                 (2, True, lambda K: (0.3, 1.0)), (1, False, lambda K: (64.0, 0.0))]


@pytest.mark.parametrize("k,qr,tols", MON_MAP_CASES, ids=["mon-k2", "mgs+mon-orth", "mgs+mon-det", "mon-k1"])
def test_residual_monitor_bitwise_maps(k, qr, tols):
    """ResidualMonitor on the standard map: the residuals of the one-step tangent, the frame orthogonality and the
    first exceedance equal the oracle's bit for bit. The tolerances put the first orthogonality (about half the
    orbits) or det (K = 6, every orbit) exceedance at varied steps; alone on a k = 2 frame the monitor trips at the
    first handoff, and with k = 1 the frame overflows at K = 6 (status 2)."""
    m, N = 384, 1200
    for K in (0.97, 6.0):
        s = StandardMap(K)
        dt, ot = tols(K)
        chain = ((QRSpectrum(MGS()),) if qr else ()) + (ResidualMonitor(s, dt, ot),)
        say("monitor", cname(chain), "K", K, "k", k, "tols", dt, ot, "m", m, "N", N)
        x0 = starts(StandardMap, m, 23)
        res = propagate(s, x0, N, k=k, tau=1, seed=7, reducers=chain, check=False, keep_states=True)
        ch, allr = make_chain(chain)
        orc = O.run(s.descriptor(), ch, x0, O.frames(7, 1, m, 2, k), N, 1, k, math=O.MATH_CL)
        compare(res, orc, allr, 2, k)


def test_residual_monitor_bitwise_gauss():
    """ResidualMonitor on Henon-Heiles under the Gauss method, alone and after a Givens QR spectrum: the 4x4 step
    Jacobian's determinant and symplectic residuals and the first exceedance (orthogonality or det for part of the
    orbits at these tolerances) equal the oracle's bit for bit."""
    from chaos_amd import HenonHeilesGauss
    from chaos_amd.inputs import hh_section_starts
    m, N = 256, 400
    hh = HenonHeilesGauss(h=0.01, tol=1e-15)
    x0 = hh_section_starts(5, m, 1 / 8)
    for qr, dt, ot in ((False, 64.0, 1e-12), (True, 64.0, 8.8e-16), (True, 1.0, 1.0)):
        chain = ((QRSpectrum(GivensQR()),) if qr else ()) + (ResidualMonitor(hh, dt, ot),)
        say("monitor gauss", cname(chain), "tols", dt, ot, "m", m, "N", N)
        res = propagate(hh, x0, N, k=4, tau=10, seed=21, reducers=chain, check=False, keep_states=True)
        ch, allr = make_chain(chain)
        orc = O.run(hh.descriptor(), ch, x0, O.frames(21, 1, m, 4, 4), N, 10, 4, math=O.MATH_CL)
        compare(res, orc, allr, 4, 4)
