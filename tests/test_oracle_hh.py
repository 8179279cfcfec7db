"""Pin the oracle's Henon-Heiles / Gauss s = 3 step (the C4 system) on the CPU.

The Jacobian is checked against central differences of the step computed in __float128, the method's
symplecticity and energy conservation against the reference's Henon-Heiles checks
(test/henon-heiles/hh_ac13_17.jl:16-33: energy, pairing, a zero pair), and the FP64 step against the
__float128 step."""
import numpy as np
import pytest

from chaos_amd import HenonHeilesGauss, hh_energy, QRSpectrum, GivensQR, MEGNO, GALI
from chaos_amd.inputs import hh_section_starts
from chaos_amd.tangent import make_chain, decode_states, finals_of
from tests import oracle_lib as O

SYS = HenonHeilesGauss(h=0.01, tol=1e-15)


def step(z, quad=False, W=None):
    ch, _ = make_chain(())
    z = np.atleast_2d(z)
    W0 = np.repeat(np.eye(4)[None], len(z), axis=0) if W is None else W
    return O.run(SYS.descriptor(), ch, z, W0, 1, 1, 4, quad=quad, nthreads=1)


def test_step_jacobian_against_quad_differences():
    z = hh_section_starts(3, 5, 1 / 8)
    J = step(z)["W"]
    delta = 1e-9
    for c in range(4):
        e = np.zeros(4)
        e[c] = delta
        fp = step(z + e, quad=True)["x"]
        fm = step(z - e, quad=True)["x"]
        col = (fp - fm) / (2 * delta)
        assert np.max(np.abs(J[:, :, c] - col)) < 1e-6  # limited by the doubles the quad run returns
    # sharper: the Jacobian in double against the quad Jacobian of the same step
    Jq = step(z, quad=True)["W"]
    assert np.max(np.abs(J - Jq)) < 1e-13


def test_step_symplectic_and_unimodular():
    Om = np.block([[np.zeros((2, 2)), np.eye(2)], [-np.eye(2), np.zeros((2, 2))]])
    for E in (1 / 12, 1 / 8, 1 / 6):
        z = hh_section_starts(7, 20, E)
        J = step(z)["W"]
        for j in J:
            assert np.max(np.abs(j.T @ Om @ j - Om)) < 1e-13
            assert abs(np.linalg.det(j) - 1) < 1e-13


def test_state_step_against_quad():
    z = hh_section_starts(11, 20, 1 / 6)
    a, b = step(z)["x"], step(z, quad=True)["x"]
    assert np.all(np.abs(a - b) <= np.spacing(np.abs(b)))  # within one ulp: the final rounding of x + inc


def test_energy_drift_and_spectrum_pairing():
    """1e4 steps at h = 0.01: energy drift small (order 6 method, compensated increments), the QR spectrum pairs
    (lambda_i + lambda_{5-i} ~ 0) and has a zero pair (the flow direction)."""
    z = hh_section_starts(1, 8, 1 / 12)
    red = (QRSpectrum(GivensQR()),)
    ch, allr = make_chain(red)
    W0 = O.frames(1, 1, 8, 4, 4)
    out = O.run(SYS.descriptor(), ch, z, W0, 10000, 10, 4, nthreads=4)
    assert np.all(out["code"] == 0)
    assert np.max(np.abs(hh_energy(out["x"]) - hh_energy(z))) < 1e-13
    lam = finals_of(decode_states(out["states"], allr, 4, 4), allr, 4, 4)[0]["lambda"]
    # volume preservation: the exponents sum to zero (to rounding); regular orbits at E = 1/12 have every finite-time
    # exponent decaying like ln(t)/t, so pairing holds only up to that allowance
    assert np.max(np.abs(lam.sum(axis=1))) < 1e-13
    assert np.max(np.abs(lam)) < 1e-3
    assert np.max(np.abs(lam[:, 0] + lam[:, 3])) < 5e-4


def test_c4_chain_runs_and_two_frames_match():
    """MEGNO (columns 1-2) + GALI on frame 1 and QR(Givens) on frame 2 equal the single-chain runs."""
    z = hh_section_starts(2, 6, 1 / 8)
    W0 = O.frames(2, 1, 6, 4, 4)
    ind, qr = (MEGNO(replace=False), GALI()), (QRSpectrum(GivensQR()),)
    ch, allr = make_chain(ind, qr)
    both = O.run(SYS.descriptor(), ch, z, W0, 2000, 10, 4, W2=W0, nthreads=2)
    one = O.run(SYS.descriptor(), make_chain(ind)[0], z, W0, 2000, 10, 4, nthreads=2)
    two = O.run(SYS.descriptor(), make_chain(qr)[0], z, W0, 2000, 10, 4, nthreads=2)
    nb1 = one["states"].shape[1]
    assert np.array_equal(both["states"][:, :nb1], one["states"])
    assert np.array_equal(both["states"][:, nb1:], two["states"])
    fin = finals_of(decode_states(both["states"], allr, 4, 4), allr, 4, 4)
    g = fin[1]["gali"]
    assert np.all((g[:, 0] >= g[:, 1]) & (g[:, 1] >= g[:, 2]))  # GALI_2 >= GALI_3 >= GALI_4 for unit columns
