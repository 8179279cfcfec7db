"""ResidualMonitor in the oracle (src/Diagnostics/monitor.jl:44-54), pinned by what the mathematics fixes: the
one-step tangents of the standard map and of the Gauss method are symplectic with det 1, so both residuals stay at
rounding level; max|M| of the map is max(1 + K cos x, |K cos x|, 1) <= 1 + |K|; a tolerance of zero trips the
orthogonality check of a frame that is not exactly orthonormal. The bounds catch the plausible mistakes: with the
form transposed, M' Omega' M - Omega = -2 Omega for a symplectic M (residual 2), and a dropped or sign-flipped term
of the determinant moves it away from 1 by O(1)."""
import numpy as np

from chaos_amd import StandardMap, HenonHeilesGauss, QRSpectrum, MGS, GivensQR, ResidualMonitor
from chaos_amd.inputs import torus_uniform, hh_section_starts
from chaos_amd.tangent import make_chain, decode_states, finals_of
from tests import oracle_lib as O

EPS = 2.0 ** -52


def run_map(K, mon, N=400, k=2):
    ch, allr = make_chain((QRSpectrum(MGS()), mon) if k == 2 else (mon,))
    x0 = torus_uniform(4, 32)
    r = O.run(StandardMap(K).descriptor(), ch, x0, O.frames(4, 1, 32, 2, k), N, 1, k)
    return r, finals_of(decode_states(r["states"], allr, 2, k), allr, 2, k), allr


def test_map_residuals_are_rounding_level_and_M_max_bounded():
    K = 6.0
    r, fin, allr = run_map(K, ResidualMonitor(StandardMap(K), 64.0, 1e-12))
    f = fin[1]
    assert np.all(r["code"] == 0)
    assert np.all(f["det_max"] <= 4 * EPS * (1 + K)) and np.all(f["symp_max"] <= 4 * EPS * (1 + K) ** 2)
    # MGS loses orthogonality as eps * cond(W), and one step has cond(M) <= (1 + K)^2
    assert np.all(f["orth_max"] <= 4 * EPS * (1 + K) ** 2) and np.all(f["orth_max"] > 0)
    assert np.all(f["M_max"] <= 1 + K) and np.all(f["M_max"] >= 1)
    st = decode_states(r["states"], allr, 2, 2)[1]
    assert np.all(st["step"] == 400) and np.all(st["sum_max"] == 0) and np.all(st["G_max"] == 0)


def test_zero_orth_tolerance_trips_at_first_handoff():
    r, fin, allr = run_map(2.0, ResidualMonitor(StandardMap(2.0), 64.0, 0.0))
    st = decode_states(r["states"], allr, 2, 2)[1]
    tripped = st["code"] == 3
    assert tripped.sum() > 0
    assert np.all(r["code"][tripped] == 3) and np.all(r["hstep"][tripped] == 1)
    assert np.all(st["value"][tripped] > 0) and np.all(st["threshold"][tripped] == 0)


def test_k1_has_no_orthogonality_check():
    r, fin, allr = run_map(1.0, ResidualMonitor(StandardMap(1.0), 64.0, 0.0), k=1)
    assert np.all(r["code"] == 0) and np.all(fin[0]["orth_max"] == 0)


def test_gauss_step_jacobian_is_symplectic():
    hh = HenonHeilesGauss(h=0.01, tol=1e-15)
    ch, allr = make_chain((QRSpectrum(GivensQR()), ResidualMonitor(hh, 64.0, 1e-12)))
    x0 = hh_section_starts(2, 16, 1 / 8)
    r = O.run(hh.descriptor(), ch, x0, O.frames(2, 1, 16, 4, 4), 300, 10, 4)
    f = finals_of(decode_states(r["states"], allr, 4, 4), allr, 4, 4)[1]
    assert np.all(r["code"] == 0)
    assert np.all(f["det_max"] <= 64 * EPS * (1 + f["M_max"])) and np.all(f["det_max"] < 1e-13)
    assert np.all(f["symp_max"] < 1e-13) and np.all(f["M_max"] > 0.9)

