"""The double-double (compensated) Henon-Heiles orbit of the oracle, CL_ELT_DD: the state carried as a value and a
correction, the correction entering the stage values and the update z, e = TwoSum(z, inc + e), the reference's
compensated summation (src/Integrators/gauss.jl:117-158, src/Integrators/stepper.jl:50-58). Pinned against the same
method in __float128 (the oracle's Quad policy): the compensated orbit z + e tracks it to a few ulp where the plain
Float64 orbit wanders 100 times further; and the energy error carries no rounding drift (the reference's AC15d,
test/integrator-catalog/ic_ac13_16.jl:105-138: the mean over the starts within 3 standard errors)."""
import numpy as np

from chaos_amd import HenonHeilesGauss, hh_energy, _lib as L
from chaos_amd.inputs import hh_section_starts
from chaos_amd.tangent import make_chain
from tests import oracle_lib as O

SYS = HenonHeilesGauss(h=0.01, tol=1e-15)
DD = L.CL_ELT_DD


def _runs(z, N, quad=False):
    ch, _ = make_chain(())
    W0 = np.repeat(np.eye(4)[None], len(z), axis=0)
    zz = np.concatenate([z, np.zeros_like(z)], axis=1)
    a = O.run(SYS.descriptor(), ch, z, W0, N, 10, 4, nthreads=8)
    b = O.run(SYS.descriptor(), ch, zz, W0, N, 10, 4, elt=DD, nthreads=8)
    q = O.run(SYS.descriptor(), ch, zz, W0, N, 10, 4, elt=DD, quad=True, nthreads=8) if quad else None
    return a, b, q


def test_dd_orbit_tracks_the_quad_method():
    """2000 steps at E = 1/8 on 4 section starts: |(z + e)_dd - (z + e)_quad| <= 1e-14 and at least 20 times below the
    plain Float64 orbit's distance from the same __float128 trajectory."""
    z = hh_section_starts(3, 4, 1 / 8)
    a, b, q = _runs(z, 2000, quad=True)
    ref = q["x"][:, :4] + q["x"][:, 4:]
    d_dd = np.max(np.abs((b["x"][:, :4] + b["x"][:, 4:]) - ref))
    d_64 = np.max(np.abs(a["x"] - ref))
    assert d_dd <= 1e-14 and d_dd * 20 <= d_64, (d_dd, d_64)


def test_dd_energy_without_drift():
    """1e5 steps on 16 starts at E = 1/12 and 1/8: the rms energy error of the compensated orbit stays below 1e-16
    (below one ulp of H) and 10 times below the plain orbit's; its mean is within 3 standard errors of zero."""
    for E in (1 / 12, 1 / 8):
        z = hh_section_starts(5, 16, E)
        a, b, _ = _runs(z, 100000)
        E0 = np.array([hh_energy(v) for v in z])
        ea = np.array([hh_energy(v) for v in a["x"]]) - E0
        eb = np.array([hh_energy(v[:4]) for v in b["x"]]) - E0
        rms_a, rms_b = np.sqrt(np.mean(ea ** 2)), np.sqrt(np.mean(eb ** 2))
        assert rms_b <= 1e-16 and 10 * rms_b <= rms_a, (E, rms_a, rms_b)
        se = np.std(eb, ddof=1) / np.sqrt(len(eb))
        assert abs(np.mean(eb)) <= 3 * se + 1e-20
