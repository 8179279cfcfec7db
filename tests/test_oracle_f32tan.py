"""The oracle's FP32-tangent policy (CL_ELT_F64_F32TAN) on the CPU: FP64 orbit and Jacobian, Float32 frames between
handoffs, FP64 reducers. Its Lyapunov exponents stay within 1e-6 of the same-orbit __float128 recomputation (the
tangent products, the QR and the log sums in 113-bit arithmetic along the same FP64 orbit, the reference's same-orbit
method of test/tangent-core/tc_ac01_04.jl:66-88), on regular and chaotic orbits; the frames it hands back are Float32
numbers; the orbit is the FP64 one bit for bit; and each indicator's difference from the Float64 run is stated."""
import numpy as np

from chaos_amd import StandardMap, WrappedStandardMap, _lib as L
from chaos_amd.reducers import QRSpectrum, MGS, HouseholderQR, MEGNO, FLI, SALI, OFLI
from chaos_amd.tangent import make_chain, decode_states, finals_of
from chaos_amd.inputs import grid_centres
from tests import oracle_lib as O

F32 = L.CL_ELT_F64_F32TAN


def _finals(out, red, red2=()):
    _, allr = make_chain(red, red2)
    return finals_of(decode_states(out["states"], allr, 2, 2), allr, 2, 2)


def test_f32_tangent_lambda_against_quad_same_orbit():
    """QR(MGS), k = 2, tau = 1, 4000 steps on 256 grid orbits at K = 0.97, 2, 6, against the same-orbit __float128
    recomputation. Each step rounds the frame to Float32, a direction error of about eps32 = 2^-24. On a chaotic orbit
    such an error decays, so the exponent error stays near eps32. On a
    regular orbit the tangent product shears: a direction error at step j changes log ||P v|| by up to (n - j) times
    itself, so the error grows like sqrt(n) eps32 (independent roundings) and at most like n eps32 (correlated ones
    along a quasi-periodic orbit). Stated: the median error below 1e-7 at every K; every orbit within n eps32; at most
    8% / 3% / 2% of the orbits above 1e-6 at K = 0.97 / 2 / 6 (K = 6 keeps small islands). The Float64 run is within
    1e-12 on every orbit."""
    red = (QRSpectrum(MGS()),)
    ch, _ = make_chain(red)
    x0 = grid_centres(16)
    W0 = O.frames(1, 1, len(x0), 2, 2)
    n = 4000
    for K, frac in ((0.97, 0.08), (2.0, 0.03), (6.0, 0.02)):
        sysd = StandardMap(K).descriptor()
        f = _finals(O.run(sysd, ch, x0, W0, n, 1, 2, elt=F32), red)[0]["lambda"]
        d = _finals(O.run(sysd, ch, x0, W0, n, 1, 2), red)[0]["lambda"]
        q = O.same_orbit_quad(K, ch, x0, W0, n, 1, 2)
        assert np.max(np.abs(d - q)) < 1e-12
        err = np.max(np.abs(f - q), axis=1)
        assert np.median(err) < 1e-7, (K, np.median(err))
        assert np.mean(err > 1e-6) <= frac, (K, np.mean(err > 1e-6))
        assert 0 < np.max(err) <= n * 2.0 ** -24


def test_f32_tangent_orbit_bits_and_float_frames():
    """The orbit is the FP64 one bit for bit (the tangent precision does not touch it); the returned frames are Float32
    numbers; the wrapped map too."""
    red = (MEGNO(replace=False), FLI(replace=False), OFLI(replace=False), SALI())
    red2 = (QRSpectrum(HouseholderQR()),)
    ch, _ = make_chain(red, red2)
    x0 = grid_centres(8)
    W0 = O.frames(2, 1, len(x0), 2, 2)
    for sys in (StandardMap(2.0), WrappedStandardMap(2.0)):
        sysd = sys.descriptor()
        x = x0 if sys.state_length == 2 else np.concatenate([x0, np.zeros_like(x0)], axis=1)
        a = O.run(sysd, ch, x, W0, 500, 1, 2, W2=W0, elt=F32)
        b = O.run(sysd, ch, x, W0, 500, 1, 2, W2=W0)
        assert np.array_equal(a["x"], b["x"])
        for key in ("W", "W2"):
            assert np.array_equal(a[key], a[key].astype(np.float32).astype(np.float64))


def test_f32_tangent_indicator_differences_stated():
    """Each indicator of the fused pass against the Float64 run on the same orbits, 64 grid orbits, 2000 steps, K = 0.97,
    2, 6 (the sizes below are the measured ones with a margin of about two): MEGNO Y within 3e-3 relative (its weights
    n a_n carry the frame errors of the late handoffs), Ybar 5e-4, FLI and OFLI 2e-4, SALI 1e-4 absolute, the QR
    exponents 5e-6; at K = 6 (chaotic: the errors decay) every difference is below 1e-6."""
    red = (MEGNO(replace=False), FLI(replace=False), OFLI(replace=False), SALI())
    red2 = (QRSpectrum(HouseholderQR()),)
    ch, _ = make_chain(red, red2)
    x0 = grid_centres(8)
    W0 = O.frames(2, 1, len(x0), 2, 2)
    rel = lambda u, v: np.max(np.abs(u - v) / np.maximum(1.0, np.abs(v)))
    for K in (0.97, 2.0, 6.0):
        sysd = StandardMap(K).descriptor()
        a = _finals(O.run(sysd, ch, x0, W0, 2000, 1, 2, W2=W0, elt=F32), red, red2)
        b = _finals(O.run(sysd, ch, x0, W0, 2000, 1, 2, W2=W0), red, red2)
        diffs = [rel(a[0]["Y"], b[0]["Y"]), rel(a[0]["Ybar"], b[0]["Ybar"]), rel(a[1]["fli"], b[1]["fli"]),
                 rel(a[2]["ofli"], b[2]["ofli"]), np.max(np.abs(a[3]["sali"] - b[3]["sali"])),
                 np.max(np.abs(a[4]["lambda"] - b[4]["lambda"]))]
        bounds = [1e-6] * 6 if K == 6.0 else [3e-3, 5e-4, 2e-4, 2e-4, 1e-4, 5e-6]
        assert all(x <= y for x, y in zip(diffs, bounds)), (K, diffs)
