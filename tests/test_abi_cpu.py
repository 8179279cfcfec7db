"""The C-ABI library loads, exports every symbol include/chaoslab_hip.h declares, and its host-side
pieces (sizes, frames, elementary functions, validation) agree with the oracle. No GPU is used."""
import ctypes
import os
import re

import mpmath as mp
import numpy as np
import pytest

from chaos_amd import _lib as L
from chaos_amd import StandardMap, QRSpectrum, MGS, MEGNO, SALI, FLI, OFLI, HouseholderQR, propagate
from chaos_amd.tangent import make_chain, state_layout
from chaos_amd.reducers import (CGS, RGS, AugmentedQR, GivensQR, Benettin, GALI)
from tests import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "chaoslab_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cl_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = L.load()
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SIGNATURES), "ctypes signatures must cover exactly the header"


def test_probe_without_device():
    lib = L.load()
    p = L.cl_probe_t()
    rc = lib.cl_probe(L.CL_ELT_F64, ctypes.byref(p))
    assert p.loadable == 1
    if rc != 0:
        assert rc == L.CL_ERR_NO_DEVICE and p.functional == 0


CHAINS = [
    (), (QRSpectrum(CGS()),), (QRSpectrum(MGS()),), (QRSpectrum(RGS()),), (QRSpectrum(HouseholderQR()),),
    (QRSpectrum(AugmentedQR()),), (QRSpectrum(GivensQR()),), (MEGNO(),), (SALI(),), (FLI(),), (OFLI(),), (GALI(),),
    (MEGNO(replace=False), SALI()), (FLI(replace=False), SALI()), (MEGNO(replace=False), FLI(replace=False), SALI()),
    (MEGNO(replace=False), FLI(replace=False), OFLI(replace=False), SALI()),
]


@pytest.mark.parametrize("chain", CHAINS, ids=lambda c: "+".join(type(r).__name__ for r in c) or "none")
def test_state_bytes_agree(chain):
    """The kernel's state tuple, the oracle's Julia-layout writer and the host decoder have one size."""
    lib = L.load()
    sysd = StandardMap(1.0).descriptor()
    ch, allr = make_chain(chain)
    nb = lib.cl_state_bytes(ctypes.byref(sysd), ctypes.byref(ch), L.CL_ELT_F64, 2, 2)
    assert nb >= 0
    assert nb == O.load().or_state_bytes(ctypes.byref(ch), 2, 2)
    assert nb == state_layout(allr, 2, 2)[1]


def test_state_bytes_reference_layouts():
    """Sizes of the reference's isbits states (Julia C-alignment): QRState{Float64,2} 104 B, MEGNOState 200 B,
    SALIState 72 B, FLIState 96 B, BenettinState 48 B (src/Lyapunov/qr_spectrum.jl:22-33 and the indicator files)."""
    d = {QRSpectrum(MGS()): 104, MEGNO(): 200, SALI(): 72, FLI(): 96}
    for r, b in d.items():
        assert r.dtype(2, 2).itemsize == b
    assert Benettin().dtype(2, 1).itemsize == 48


def test_unsupported_chain_reports_reason_3():
    lib = L.load()
    sysd = StandardMap(1.0).descriptor()
    ch, _ = make_chain((SALI(), MEGNO()))
    assert lib.cl_state_bytes(ctypes.byref(sysd), ctypes.byref(ch), L.CL_ELT_F64, 2, 2) == -L.CL_ERR_UNSUPPORTED


def test_host_frames_equal_oracle():
    lib = L.load()
    for d, k in ((2, 1), (2, 2)):
        out = np.empty((300, k, d))
        assert lib.cl_initial_deviations(123456789, 1, 300, d, k, out.ctypes.data) == 0
        assert np.array_equal(out.transpose(0, 2, 1), O.frames(123456789, 1, 300, d, k))


def test_host_math_accuracy():
    """The kernels' sin/cos/log (host build): within 1 ulp of a 200-bit reference over small, large (Payne-Hanek)
    and subnormal arguments, and just below 2^50, where the quadrant chosen from the rounded x * 2/pi can be off
    by one and the reduced argument reaches about 0.9 (the polynomials hold on |r| <= 1)."""
    lib = L.load()
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(-10, 10, 400), 10 ** rng.uniform(-300, 300, 400), -(10 ** rng.uniform(5, 20, 200)),
                         rng.uniform(4e14, 2.0 ** 50, 300),
                         np.array([np.pi / 4, np.pi / 2, 2.0 ** 50, 1e22, 5e-324, 1e308, 6381956970095103.0 * 2 ** 797,
                                   907878090244995.4, 788081350329621.0])])
    s, c, l = (np.empty_like(xs) for _ in range(3))
    lib.cl_math_host(len(xs), xs.ctypes.data, s.ctypes.data, c.ctypes.data, l.ctypes.data)
    mp.mp.prec = 200

    def ulps(ref, got):
        if ref == 0:
            return abs(got)
        e = mp.floor(mp.log(abs(ref), 2))
        return float(abs(mp.mpf(got) - ref) / mp.mpf(2) ** (e - 52))
    worst = 0.0
    for i, x in enumerate(xs):
        xm = mp.mpf(float(x))
        worst = max(worst, ulps(mp.sin(xm), s[i]), ulps(mp.cos(xm), c[i]))
        if x > 0:
            worst = max(worst, ulps(mp.log(xm), l[i]))
    assert worst <= 1.0
    assert np.isnan(l[xs < 0]).all()


def test_chunk_validation_errors():
    """Argument errors come back as code 10 before any device work (src/Tangent/orbit.jl:70-74)."""
    lib = L.load()
    sysd = StandardMap(1.0).descriptor()
    ch, _ = make_chain(())
    x = np.zeros((1, 2))
    W = np.zeros((1, 2, 2))
    code = np.zeros(1, np.int8)
    h = np.zeros(1, np.int64)
    # null context
    rc = lib.cl_propagate_chunk(None, ctypes.byref(sysd), ctypes.byref(ch), 1, 1, 2, 2, 2, 0, 10, 10, 1,
                                x.ctypes.data, W.ctypes.data, None, None, code.ctypes.data, h.ctypes.data, None)
    assert rc == L.CL_ERR_ARGUMENT


def test_propagate_validates_before_device():
    """propagate raises ArgumentError-like ValueErrors before touching a device (src/Tangent/propagate.jl:279-289)."""
    sm = StandardMap(1.0)
    x0 = np.zeros((3, 2))
    with pytest.raises(ValueError, match="multiple of tau"):
        propagate(sm, x0, 10, k=2, tau=3, seed=1)
    with pytest.raises(ValueError, match="k must be in"):
        propagate(sm, x0, 10, k=3, seed=1)
    with pytest.raises(ValueError, match="one matrix per orbit"):
        propagate(sm, x0, 10, k=2, seed=1, W0=np.zeros((2, 2, 2)))
    with pytest.raises(ValueError, match="needs k = 2"):
        propagate(sm, x0, 10, k=1, seed=1, reducers=(MEGNO(),))
    with pytest.raises(ValueError, match="state length"):
        propagate(sm, np.zeros((3, 3)), 10, k=2, seed=1)


def test_chain_state_bytes_and_frames():
    """The chain's QR state is QRState{Float64,16} (16 k + 72 bytes); kernels exist for N in {8..128} powers of two
    and k = 16; host frames of the chain sizes equal the oracle's."""
    from chaos_amd import StandardMapChain, QRSpectrum, HouseholderQR, MGS
    from chaos_amd.tangent import make_chain
    lib = L.load()
    qr = make_chain((QRSpectrum(HouseholderQR()),))[0]
    for N in (8, 16, 32, 64, 128):
        sd = StandardMapChain(N, 1.5, 0.1, "ring").descriptor()
        assert lib.cl_state_bytes(ctypes.byref(sd), ctypes.byref(qr), 1, 2 * N, 16) == 16 * 16 + 72
        assert lib.cl_state_bytes(ctypes.byref(sd), ctypes.byref(make_chain(())[0]), 1, 2 * N, 16) == 0
        assert lib.cl_state_bytes(ctypes.byref(sd), ctypes.byref(qr), 1, 2 * N, 4) == -3
    sd = StandardMapChain(100, 1.5, 0.1, "ring").descriptor()
    assert lib.cl_state_bytes(ctypes.byref(sd), ctypes.byref(qr), 1, 200, 16) == -3
    sd = StandardMapChain(16, 1.5, 0.1, "ring").descriptor()
    mgs = make_chain((QRSpectrum(MGS()),))[0]
    assert lib.cl_state_bytes(ctypes.byref(sd), ctypes.byref(mgs), 1, 32, 16) == -3
    for d in (16, 256):
        W = np.empty((3, 16, d))
        assert lib.cl_initial_deviations(9, 1, 3, d, 16, W.ctypes.data) == 0
        assert np.array_equal(W.transpose(0, 2, 1), O.frames(9, 1, 3, d, 16))


def test_mlp_train_step_validation_errors():
    """cl_mlp_train_step rejects a null context and a malformed network descriptor with code 10 before it touches the
    context or the device (the context below is a zeroed stand-in that the checks never reach)."""
    lib = L.load()
    X = np.zeros((4, 8), np.uint16)
    v = np.zeros(8, np.float32)
    s = np.zeros(2, np.float64)
    ptrs = [np.zeros(64, np.float32) for _ in range(7)]

    def net(nlayers, sizes, null_slot=None):
        n = L.cl_mlp_net()
        n.nlayers = nlayers
        for i, z in enumerate(sizes):
            n.sizes[i] = z
        for l in range(L.CL_MLP_MAX_LAYERS):
            for j, nm in enumerate(("W", "Wb", "b", "mW", "vW", "mb", "vb")):
                getattr(n, nm)[l] = None if (l, nm) == null_slot else ptrs[j].ctypes.data
        return n

    fake = ctypes.create_string_buffer(4096)

    def call(ctx, n, Bn=8, ldx=8, step=1):
        return lib.cl_mlp_train_step(ctx, ctypes.byref(n), X.ctypes.data, ldx, Bn, v.ctypes.data, v.ctypes.data,
                                     v.ctypes.data, 0.1, step, 1e-3, 0.9, 0.999, 1e-8, s.ctypes.data, None)

    ok = net(2, (4, 3, 2))
    assert call(None, ok) == L.CL_ERR_ARGUMENT
    assert call(fake, net(0, (4,))) == L.CL_ERR_ARGUMENT
    assert call(fake, net(L.CL_MLP_MAX_LAYERS + 1, (4,) * 9)) == L.CL_ERR_ARGUMENT
    assert call(fake, net(2, (4, 3, 3))) == L.CL_ERR_ARGUMENT            # the loss needs two outputs
    assert call(fake, net(2, (4, 0, 2))) == L.CL_ERR_ARGUMENT
    assert call(fake, net(2, (4, 3, 2), null_slot=(1, "vb"))) == L.CL_ERR_ARGUMENT
    assert call(fake, ok, Bn=0) == L.CL_ERR_ARGUMENT
    assert call(fake, ok, ldx=4) == L.CL_ERR_ARGUMENT                     # row stride below the batch
    assert call(fake, ok, step=0) == L.CL_ERR_ARGUMENT
