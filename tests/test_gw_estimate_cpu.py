"""The one-division G_w estimate of the pipeline kernel's checks (gw_quick2) against window_Gw (the conditioning
estimate of orthonormalize.jl:112-127) on the host: the kernels' window_Gw equals the oracle's bit for bit, and
wherever the estimate is defined it stays within GW_QUICK_REL = 1e-13 of window_Gw. The pipeline settles a check
or skips a running-maximum update from the estimate only with that margin, so this bound is what keeps its
results window_Gw's. Rounding analysis puts the gap below about 16 units in the last place; the test asserts 64
(the margin itself is about 900)."""
import numpy as np

from chaos_amd import _lib as L
from tests import oracle_lib as O

U = 2.0 ** -53
REL = 1e-13


def gw_host(R, G_max=None):
    R = np.ascontiguousarray(R, dtype=np.float64)
    n = R.shape[0]
    Gm = np.zeros(n) if G_max is None else np.ascontiguousarray(np.broadcast_to(G_max, (n,)), dtype=np.float64)
    q = np.empty(n)
    e = np.empty(n)
    st = np.empty(n, dtype=np.uint8)
    L.load().cl_gw_host(n, R.ctypes.data, Gm.ctypes.data, q.ctypes.data, e.ctypes.data, st.ctypes.data)
    return (q, e) if G_max is None else (q, e, st.astype(bool))


def G_of(gw):
    """G from G_w at k = 2 in IEEE double (orthonormalize.jl:112-127: (G_w + sqrt(max(G_w^2 - 4, 0))) / 2)."""
    return (gw + np.sqrt(np.maximum(gw * gw - 4.0, 0.0))) / 2.0


def logu(rng, lo, hi, n):
    return 10.0 ** rng.uniform(lo, hi, n)


def factors(seed, n):
    """Upper-triangular 2 x 2 factors over the estimate's whole range: diagonals log-uniform in [1e-100, 1e100],
    off-diagonals of either sign from 1e-300 to 1e100 and exact zeros, near-singular and near-orthogonal ones."""
    rng = np.random.default_rng(seed)
    R = np.zeros((n, 4))
    R[:, 0] = logu(rng, -100, 100, n)
    R[:, 3] = logu(rng, -100, 100, n)
    R[:, 2] = rng.choice([-1.0, 1.0], n) * logu(rng, -300, 100, n)
    R[rng.random(n) < 0.05, 2] = 0.0
    # factors of well-conditioned frames (the common case in a run): entries of order one
    k = n // 4
    R[:k, 0] = rng.uniform(0.1, 10.0, k)
    R[:k, 3] = rng.uniform(0.1, 10.0, k)
    R[:k, 2] = rng.uniform(-10.0, 10.0, k)
    # strongly ill-conditioned ones: G_w from 1e8 to beyond 2^52, many near the conditioning gate (2^51)
    R[k:2 * k, 3] = R[k:2 * k, 0] * logu(rng, -20, -8, k)
    R[k:2 * k, 2] = R[k:2 * k, 0] * rng.uniform(-1.0, 1.0, k)
    R[2 * k:2 * k + k // 4, 0] = 1.0
    R[2 * k:2 * k + k // 4, 2] = 0.0
    R[2 * k:2 * k + k // 4, 3] = 1.0 / rng.uniform(1.5e15, 2.5e15, k // 4)
    return R


def test_window_gw_equals_oracle():
    R = factors(1, 50_000)
    _, e = gw_host(R)
    ref = O.window_gw2(R)
    assert np.array_equal(e.view(np.uint64), ref.view(np.uint64))


def test_estimate_within_bound():
    R = factors(2, 200_000)
    q, e = gw_host(R)
    defined = q >= 0.0
    assert defined.mean() > 0.95
    fin = defined & np.isfinite(e) & np.isfinite(q)
    # where the pipeline trusts the estimate (below GW_QUICK_GATE = 2e15) window_Gw is finite and passes the
    # conditioning check G_w * 2^-52 < 0.5
    trusted = defined & (q < 2.0e15)
    assert np.all(fin[trusted])
    gap = np.abs(e[fin] - q[fin]) / q[fin]
    print("largest gap", gap.max() / U, "units of 2^-53 over", fin.sum(), "factors")
    assert gap.max() <= 64 * U < REL
    assert np.all(e[trusted] * (2.0 ** -52) < 0.5)
    assert np.sum(trusted & (q > 1e15)) > 10  # the gate's neighbourhood is sampled


def test_estimate_undefined_outside_range():
    cases = np.array([
        [1e-101, 0.0, 1.0, 1.0],     # a diagonal below 1e-100
        [1.0, 0.0, 1.0, 2e100],      # a diagonal above 1e100
        [1.0, 0.0, -2e100, 1.0],     # an off-diagonal above 1e100
        [1.0, 1e-300, 1.0, 1.0],     # not upper-triangular
        [np.nan, 0.0, 1.0, 1.0],
        [1.0, 0.0, np.inf, 1.0],
        [1.0, 0.0, np.nan, 1.0],
        [0.0, 0.0, 1.0, 1.0],
    ])
    q, _ = gw_host(cases)
    assert np.all(q == -1.0)
    # the range's ends are inside
    q, e = gw_host(np.array([[1e-100, 0.0, 1e100, 1e100], [1e100, 0.0, -1e100, 1e-100]]))
    assert np.all(q > 0.0) and np.all(np.abs(e - q) <= 64 * U * q)


def test_fused_step_skip_never_raises_the_maximum():
    """The fused QR step skips window_Gw when the estimate shows that G of the handoff cannot exceed the running
    G_max. Adversarial maxima around G of the exact G_w (equal, one ulp either side, relative offsets from 1e-15
    to 1e-9): wherever the step skips, G(window_Gw) <= G_max, so jmax would have kept G_max. Skips happen once
    G_max clears G by the margin (2e-12), and never below G_max = 2 or above 1e150."""
    rng = np.random.default_rng(5)
    R = factors(3, 40_000)
    _, e = gw_host(R)
    keep = np.isfinite(e) & (e > 2.0) & (e < 1e140)
    R = R[keep]
    G = G_of(e[keep])
    offs = np.array([0.0, -1e-15, 1e-15, -1e-13, 1e-13, 1e-12, 2e-12, 4e-12, 1e-11, 1e-9])
    skipped_total = 0
    for o in offs:
        for Gm in (G * (1.0 + o), np.nextafter(G, np.inf), np.nextafter(G, -np.inf)):
            q, e2, st = gw_host(R, Gm)
            assert np.array_equal(e2, e[keep])
            bad = st & (G > Gm)
            assert not bad.any(), (o, np.flatnonzero(bad)[:5])
            assert not (st & ((Gm < 2.0) | (Gm >= 1e150))).any()
            skipped_total += int(st.sum())
    # well clear of the margin the estimate settles every factor whose G_max is in range
    Gm = G * (1.0 + 1e-10)
    q, _, st = gw_host(R, Gm)
    inrange = (q >= 0.0) & (Gm >= 2.0) & (Gm < 1e150)
    assert st[inrange].all() and skipped_total > 0
