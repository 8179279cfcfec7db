"""Reducer types of the tangent stream and their host-side summaries.

Each reducer mirrors the reference type of the same name (constructor arguments, the state's
byte layout, and `reducer_final`), so that a chain's per-orbit states written by the kernel decode
to the same named fields the reference returns. States are read from the raw bytes of the chain's
state tuple (one row per orbit), vectorised over orbits with numpy.
"""
import numpy as np

from . import _lib as L

EPS64 = 2.0 ** -52


class Orthonormalizer:
    code = 0
    name = ""


class CGS(Orthonormalizer):
    """Classical Gram-Schmidt (src/Lyapunov/orthonormalize.jl:8,38-47)."""
    code, name = L.CL_QR_CGS, "CGS"


class MGS(Orthonormalizer):
    """Modified Gram-Schmidt (src/Lyapunov/orthonormalize.jl:10,49-58)."""
    code, name = L.CL_QR_MGS, "MGS"


class RGS(Orthonormalizer):
    """Repeated Gram-Schmidt, eta = 1/sqrt 2, at most 4 passes (src/Lyapunov/orthonormalize.jl:12,21,60-86)."""
    code, name = L.CL_QR_RGS, "RGS"


class HouseholderQR(Orthonormalizer):
    """Householder reflectors, V-H (src/Lyapunov/householder.jl:6)."""
    code, name = L.CL_QR_HOUSEHOLDER, "V-H"


class AugmentedQR(Orthonormalizer):
    """Householder on [W, I], V-A (src/Lyapunov/householder.jl:8)."""
    code, name = L.CL_QR_AUGMENTED, "V-A"


class GivensQR(Orthonormalizer):
    """Givens rotations on the rows of [W, I], V-G (src/Lyapunov/householder.jl:10)."""
    code, name = L.CL_QR_GIVENS, "V-G"


def _acc(a):
    """total() of CompensatedSum columns a[..., 0] (s) and a[..., 1] (c)."""
    return a[..., 0] + a[..., 1]


class Reducer:
    kind = 0

    def descriptor(self, frame):
        r = L.cl_reducer()
        r.kind = self.kind
        r.frame = frame
        if hasattr(self, "replace"):
            r.replace = int(self.replace)
        return r

    def nbytes(self, d, k):
        return self.dtype(d, k).itemsize

    def check(self, d, k):
        pass

    def provenance(self):
        return {}


class QRSpectrum(Reducer):
    """QRSpectrum(alg; orthogonality_tol=8.0) (src/Lyapunov/qr_spectrum.jl:14-18)."""
    kind = L.CL_RED_QR

    def __init__(self, alg, orthogonality_tol=8.0):
        if isinstance(alg, type):
            alg = alg()
        self.alg = alg
        self.orthogonality_tol = float(orthogonality_tol)

    def descriptor(self, frame):
        r = super().descriptor(frame)
        r.alg = self.alg.code
        r.tol = self.orthogonality_tol
        return r

    def dtype(self, d, k):
        return np.dtype([("sums", "<f8", (k, 2)), ("n", "<i8"), ("G_max", "<f8"), ("orth_max", "<f8"),
                         ("count_max", "<i8"), ("code", "i1"), ("pad", "u1", 7), ("column", "<i8"),
                         ("value", "<f8"), ("threshold", "<f8"), ("step", "<i8")])

    def final(self, st):
        """(n, sum, comp, lambda, G_max, orth_max, count_max), qr_spectrum.jl:88-93."""
        n = st["n"]
        tot = _acc(st["sums"])
        lam = np.where(n[:, None] == 0, 0.0, tot / np.where(n == 0, 1, n)[:, None])
        return dict(n=n, sum=st["sums"][..., 0].copy(), comp=st["sums"][..., 1].copy(), **{"lambda": lam},
                    G_max=st["G_max"].copy(), orth_max=st["orth_max"].copy(), count_max=st["count_max"].copy())

    def error(self, row):
        kind = "orthogonality" if row["code"] == 4 else "orthonormalisation breakdown"
        return kind, float(row["value"]), float(row["threshold"]), int(row["step"]), int(row["column"])

    def provenance(self):
        return dict(estimator="qr", orthonormalizer=self.alg.name, orthogonality_tol=self.orthogonality_tol)


class Benettin(Reducer):
    """Benettin() for k = 1 (src/Lyapunov/benettin.jl:11-53)."""
    kind = L.CL_RED_BENETTIN

    def check(self, d, k):
        if k != 1:
            raise ValueError(f"Benettin needs k = 1, got k = {k}")

    def dtype(self, d, k):
        return np.dtype([("sum", "<f8", 2), ("n", "<i8"), ("code", "i1"), ("pad", "u1", 7), ("value", "<f8"),
                         ("step", "<i8")])

    def final(self, st):
        n = st["n"]
        lam = np.where(n == 0, 0.0, _acc(st["sum"]) / np.where(n == 0, 1, n))
        return dict(n=n, sum=st["sum"][:, :1].copy(), comp=st["sum"][:, 1:].copy(), **{"lambda": lam[:, None]})

    def error(self, row):
        v = float(row["value"])
        return "deviation matrix", v, 0.0 if np.isfinite(v) else np.inf, int(row["step"]), 1

    def provenance(self):
        return dict(estimator="benettin")


def _k2(name, k):
    if k != 2:
        raise ValueError(f"{name} needs k = 2, got k = {k}")


class MEGNO(Reducer):
    """MEGNO(; replace=True) for k = 2 (src/Indicators/megno.jl:29-113)."""
    kind = L.CL_RED_MEGNO

    def __init__(self, replace=True):
        self.replace = bool(replace)

    def check(self, d, k):
        # k = 2 as the reference; a wider frame is read (columns 1 and 2) only when MEGNO does not replace it
        if k != 2 and (self.replace or k < 2):
            raise ValueError(f"MEGNO needs k = 2, got k = {k} (k > 2 only with replace=False: columns 1-2 are read)")

    def dtype(self, d, k):
        return np.dtype([("c", "<f8", (2, 5, 2)), ("value", "<f8"), ("n", "<i8"), ("K", "<i8"), ("step", "<i8"),
                         ("code", "i1"), ("column", "i1"), ("pad", "u1", 6)])

    def final(self, st):
        """(n, handoffs, Y, Ybar, Ybar20, sum, comp, R1, R2, Q, Q2), megno.jl:97-111."""
        c = st["c"]  # [orbit, column, sum(s,R1,R2,Q,Q2), (s,c)]
        n = st["n"].astype(np.float64)[:, None]
        K = st["K"][:, None]
        tot = _acc(c)
        R1, R2, Q, Q2 = tot[:, :, 1], tot[:, :, 2], tot[:, :, 3], tot[:, :, 4]
        zero = K == 0
        nn = np.where(zero, 1.0, n)
        tau = (st["n"][:, None] // np.where(zero, 1, K)).astype(np.float64)
        Y = np.where(zero, 0.0, 2 * R1 / nn)
        Ybar = np.where(zero, 0.0, Q / np.where(zero, 1, K))
        Ybar20 = np.where(zero, 0.0, tau * Q2 / (nn * nn * nn))  # tn^3 is (tn tn) tn, not pow
        return dict(n=st["n"].copy(), handoffs=st["K"].copy(), Y=Y, Ybar=Ybar, Ybar20=Ybar20,
                    sum=c[:, :, 0, 0].copy(), comp=c[:, :, 0, 1].copy(), R1=R1, R2=R2, Q=Q, Q2=Q2)

    def error(self, row):
        v = float(row["value"])
        return "deviation matrix", v, 0.0 if np.isfinite(v) else np.inf, int(row["step"]), int(row["column"])

    def provenance(self):
        return dict(indicator="megno", convention="weights at the end of the window, mean over the handoffs 1..K")


class RLI(Reducer):
    """RLI(; replace=True) for k = 2 on a pair system (src/Indicators/rli.jl:1-140): the two columns' log sums (A, B),
    DL = |S_B - S_A| / n and the compensated sum Q of the DL series; RLI = Q / K."""
    kind = L.CL_RED_RLI

    def __init__(self, replace=True):
        self.replace = bool(replace)

    def check(self, d, k):
        if k != 2 or d % 2:
            raise ValueError(f"RLI needs k = 2 on an even dimension, got d = {d}, k = {k}")

    def dtype(self, d, k):
        return np.dtype([("sA", "<f8", 2), ("sB", "<f8", 2), ("Q", "<f8", 2), ("value", "<f8"), ("n", "<i8"),
                         ("K", "<i8"), ("step", "<i8"), ("code", "i1"), ("column", "i1"), ("pad", "u1", 6)])

    def final(self, st):
        """(n, handoffs, DL, RLI, sum, comp, Q), rli.jl:116-119."""
        sA, sB = st["sA"], st["sB"]
        K = st["K"]
        zero = K == 0
        n = np.where(zero, 1, st["n"]).astype(np.float64)
        dl = np.where(zero, 0.0, np.abs((sB[:, 0] - sA[:, 0]) + (sB[:, 1] - sA[:, 1])) / n)
        Q = st["Q"][:, 0] + st["Q"][:, 1]
        rli = np.where(zero, 0.0, Q / np.where(zero, 1, K))
        return dict(n=st["n"].copy(), handoffs=K.copy(), DL=dl, RLI=rli,
                    sum=np.stack([sA[:, 0], sB[:, 0]], axis=1), comp=np.stack([sA[:, 1], sB[:, 1]], axis=1), Q=Q)

    def error(self, row):
        v = float(row["value"])
        what = "pairing" if int(row["code"]) == 2 else "deviation matrix"
        return what, v, 0.0 if np.isfinite(v) else np.inf, int(row["step"]), int(row["column"])

    def provenance(self):
        return dict(indicator="rli", convention="same xi_0, shadow along xi_0, mean over the handoffs 1..K")


class SALI(Reducer):
    """SALI() for k = 2, replaces W by unit columns (src/Indicators/sali.jl:14-68)."""
    kind = L.CL_RED_SALI

    def check(self, d, k):
        _k2("SALI", k)

    def dtype(self, d, k):
        return np.dtype([("s", "<f8", (2, 2)), ("sali", "<f8"), ("value", "<f8"), ("n", "<i8"), ("step", "<i8"),
                         ("code", "i1"), ("column", "i1"), ("pad", "u1", 6)])

    def final(self, st):
        s, c = st["s"][:, :, 0], st["s"][:, :, 1]
        L_ = np.where(st["n"] == 0, 0.0, -((s[:, 0] + s[:, 1]) + (c[:, 0] + c[:, 1])))
        return dict(n=st["n"].copy(), sali=st["sali"].copy(), L=L_, sum=s.copy(), comp=c.copy(),
                    floor=4 * EPS64)

    error = MEGNO.error

    def provenance(self):
        return dict(indicator="sali")


class FLI(Reducer):
    """FLI(; replace=True) for k = 2 (src/Indicators/fli.jl:26-96)."""
    kind = L.CL_RED_FLI

    def __init__(self, replace=True):
        self.replace = bool(replace)

    def check(self, d, k):
        _k2("FLI", k)

    def dtype(self, d, k):
        return np.dtype([("s", "<f8", (2, 2)), ("f1", "<f8"), ("fB", "<f8"), ("value", "<f8"), ("n", "<i8"),
                         ("step", "<i8"), ("step1", "<i8"), ("stepB", "<i8"), ("code", "i1"), ("column", "i1"),
                         ("colB", "i1"), ("pad", "u1", 5)])

    def final(self, st):
        return dict(n=st["n"].copy(), fli=st["fB"].copy(), fli1=st["f1"].copy(), sum=st["s"][:, :, 0].copy(),
                    comp=st["s"][:, :, 1].copy(), stepB=st["stepB"].copy(), colB=st["colB"].astype(np.int64),
                    step1=st["step1"].copy())

    error = MEGNO.error

    def provenance(self):
        return dict(indicator="fli")


class OFLI(Reducer):
    """OFLI(field; replace=True): the FLI of the components orthogonal to a direction field
    (src/Indicators/ofli.jl:23-126). The reference defines the field for flows only (their vector
    field); on a map this build uses the map displacement F(x) - x."""
    kind = L.CL_RED_OFLI

    def __init__(self, flow=None, replace=True):
        self.flow = flow
        self.replace = bool(replace)

    def dtype(self, d, k):
        return np.dtype([("acc", "<f8", (k, 2)), ("g1", "<f8"), ("gB", "<f8"), ("inst", "<f8"), ("f1", "<f8"),
                         ("fB", "<f8"), ("sin1", "<f8"), ("sinB", "<f8"), ("sinmin", "<f8"), ("value", "<f8"),
                         ("nminf", "<i8"), ("n", "<i8"), ("step", "<i8"), ("step1", "<i8"), ("stepB", "<i8"),
                         ("colB", "i1"), ("code", "i1"), ("column", "i1"), ("pad", "u1", 5)])

    def final(self, st):
        return dict(n=st["n"].copy(), ofli=st["gB"].copy(), ofli1=st["g1"].copy(), inst=st["inst"].copy(),
                    fli=st["fB"].copy(), fli1=st["f1"].copy(), sum=st["acc"][..., 0].copy(),
                    comp=st["acc"][..., 1].copy(), stepB=st["stepB"].copy(), colB=st["colB"].astype(np.int64),
                    step1=st["step1"].copy(), sinB=st["sinB"].copy(), sin1=st["sin1"].copy(),
                    sinmin=st["sinmin"].copy(), nminf=st["nminf"].copy())

    def error(self, row):
        v = float(row["value"])
        if row["code"] == 2:
            return "vector field", v, 2.0 ** -1022, int(row["step"]), 0
        return "deviation matrix", v, 0.0 if np.isfinite(v) else np.inf, int(row["step"]), int(row["column"])

    def provenance(self):
        return dict(indicator="ofli", ofli_flow=str(self.flow), ofli_replace=self.replace)


class GALI(Reducer):
    """GALI() for 2 <= k <= d (src/Indicators/gali.jl:17-201)."""
    kind = L.CL_RED_GALI

    def check(self, d, k):
        if not (2 <= k <= d):
            raise ValueError("GALI needs 2 <= k <= d")

    def dtype(self, d, k):
        return np.dtype([("s", "<f8", k), ("c", "<f8", k), ("gali", "<f8", k - 1), ("value", "<f8"),
                         ("n", "<i8"), ("step", "<i8"), ("code", "i1"), ("pad", "u1", 7), ("column", "<i8")])

    def final(self, st, d=None):
        k = st["s"].shape[1]
        s, c = st["s"], st["c"]
        Ls = s[:, 0].copy()
        Lc = c[:, 0].copy()
        for i in range(1, k):
            Ls = Ls + s[:, i]
            Lc = Lc + c[:, i]
        L_ = np.where(st["n"] == 0, 0.0, -(Ls + Lc))
        return dict(n=st["n"].copy(), gali=st["gali"].copy(), L=L_ if d == k else None, sum=s.copy(), comp=c.copy(),
                    floor=np.array([2 * (i + 2) * EPS64 for i in range(k - 1)]))

    def error(self, row):
        v = float(row["value"])
        thr = np.inf if not np.isfinite(v) else (0.0 if v == 0 else 2.0 ** -1022)
        return "deviation matrix", v, thr, int(row["step"]), int(row["column"])

    def provenance(self):
        return dict(indicator="gali")


class ResidualMonitor(Reducer):
    """ResidualMonitor(sys, det_tol, orth_tol) (src/Diagnostics/monitor.jl:9-54): at each handoff the det and
    symplecticity residuals of the one-step tangent at the current point (det_tol times eps, scaled by
    1 + max|M| and N^2 (1 + max|M|)^2, src/Diagnostics/residuals.jl:25-33) and, for k >= 2, the orthogonality
    residual of the frame against orth_tol; running maxima. The system is the one being propagated."""
    kind = L.CL_RED_MONITOR
    NAMES = ("det", "symplecticity", "orthogonality")

    def __init__(self, sys, det_tol, orth_tol):
        self.sys, self.det_tol, self.orth_tol = sys, float(det_tol), float(orth_tol)

    def descriptor(self, frame):
        r = super().descriptor(frame)
        r.tol = self.det_tol
        r.tol2 = self.orth_tol
        return r

    def dtype(self, d, k):
        return np.dtype([("det_max", "<f8"), ("symp_max", "<f8"), ("orth_max", "<f8"), ("sum_max", "<f8"),
                         ("pair_max", "<f8"), ("G_max", "<f8"), ("M_max", "<f8"), ("code", "i1"), ("pad", "u1", 7),
                         ("value", "<f8"), ("threshold", "<f8"), ("step", "<i8")])

    def final(self, st):
        return dict(det_max=st["det_max"].copy(), symp_max=st["symp_max"].copy(), orth_max=st["orth_max"].copy(),
                    sum_max=st["sum_max"].copy(), pair_max=st["pair_max"].copy(), M_max=st["M_max"].copy(),
                    G_max=st["G_max"].copy())

    def error(self, row):
        v = float(row["value"])
        thr = float(row["threshold"]) if np.isfinite(v) else float("inf")
        return self.NAMES[int(row["code"]) - 1], v, thr, int(row["step"]), 0

    def provenance(self):
        return dict(monitor_det_tol=self.det_tol, monitor_orth_tol=self.orth_tol)


# ---------------------------------------------------------------- residual definitions and thresholds
# (src/Diagnostics/residuals.jl:4-44), host-side for the end-of-run checks of a monitor's state.

def det_threshold(tol, M):
    """tol * eps * (1 + max|M|) (residuals.jl:29)."""
    return (float(tol) * EPS64) * (1.0 + float(np.max(np.abs(M))))


def symp_threshold(tol, M):
    """tol * eps * N^2 * (1 + max|M|)^2 with 2N = size(M, 1) (residuals.jl:32-35)."""
    n = np.shape(M)[0] // 2
    a = 1.0 + float(np.max(np.abs(M)))
    return ((float(tol) * EPS64) * float(n * n)) * (a * a)


def sum_threshold(tol, G_max, tau, M_max):
    """tol * eps * (G_max / tau + M_max), independent of the run length (residuals.jl:38); vectorised."""
    return (float(tol) * EPS64) * (np.asarray(G_max, dtype=np.float64) / tau + np.asarray(M_max, dtype=np.float64))


def eps_sum(lam):
    """|sum_i lambda_i| over the last axis, summed left to right (residuals.jl:13)."""
    lam = np.asarray(lam, dtype=np.float64)
    s = lam[..., 0].copy()
    for i in range(1, lam.shape[-1]):
        s = s + lam[..., i]
    return np.abs(s)


def eps_pair(lam):
    """max_i |lambda_i + lambda_{M+1-i}| in estimator order, no sorting (residuals.jl:16-18)."""
    lam = np.asarray(lam, dtype=np.float64)
    M = lam.shape[-1]
    out = np.abs(lam[..., 0] + lam[..., M - 1])
    for i in range(1, M // 2):
        out = np.maximum(out, np.abs(lam[..., i] + lam[..., M - 1 - i]))
    return out


def check_residual(what, value, threshold, *, step, orbit, backend="cpu"):
    """Raise NumericalHealthError when `value` is not finite or exceeds `threshold`, reporting an infinite
    threshold for a non-finite value (src/Diagnostics/residuals.jl:36-44)."""
    from .tangent import NumericalHealthError
    value, threshold = float(value), float(threshold)
    if np.isfinite(value) and value <= threshold:
        return None
    raise NumericalHealthError(what, value, threshold if np.isfinite(value) else float("inf"), int(step), int(orbit),
                               backend)


def monitor_exponents(st, lam, G, n, *, pairing_tol, tau, M_max, step, orbit0=1):
    """Feed each orbit's exponents `lam` (m, 2N, estimator order) and window growth `G` (m,) to its monitor
    state `st` (the decoded ResidualMonitor states, one row per orbit) and return the new states
    (src/Diagnostics/monitor.jl:66-87): a recorded det, symplecticity or orthogonality failure is raised first;
    then G_max and M_max take the running maxima, the sum residual is checked against sum_threshold
    (raising for the first orbit above it) and sum_max and pair_max take the running maxima. `n` is the run
    length the exponents are quoted at (not used by the checks)."""
    from .tangent import NumericalHealthError
    st = np.array(st, copy=True)
    lam = np.asarray(lam, dtype=np.float64)
    if lam.ndim == 1:
        lam = lam[None, :]
    m = len(st)
    bad = np.flatnonzero(st["code"] != 0)
    if len(bad):
        i = int(bad[0])
        v = float(st["value"][i])
        raise NumericalHealthError(ResidualMonitor.NAMES[int(st["code"][i]) - 1], v,
                                   float(st["threshold"][i]) if np.isfinite(v) else float("inf"), int(st["step"][i]),
                                   orbit0 + i, "cpu")
    G_max = np.maximum(st["G_max"], np.broadcast_to(np.asarray(G, dtype=np.float64), (m,)))
    M_hi = np.maximum(st["M_max"], np.broadcast_to(np.asarray(M_max, dtype=np.float64), (m,)))
    s = eps_sum(lam)
    thr = sum_threshold(pairing_tol, G_max, tau, M_hi)
    for i in range(m):
        check_residual("sum", s[i], thr[i], step=step, orbit=orbit0 + i)
    st["sum_max"] = np.maximum(st["sum_max"], s)
    st["pair_max"] = np.maximum(st["pair_max"], eps_pair(lam))
    st["G_max"] = G_max
    st["M_max"] = M_hi
    return st


class RunMonitor(Reducer):
    """RunMonitor(sys) (src/Experiments/run_record.jl:25-43): the recorder's non-raising monitor, the running
    maxima of |det M - 1|, max |M' Omega M - Omega| and max |M| of the one-step tangent at each handoff. It never
    fails."""
    kind = L.CL_RED_RUNMONITOR

    def __init__(self, sys):
        self.sys = sys

    def dtype(self, d, k):
        return np.dtype([("det_max", "<f8"), ("symp_max", "<f8"), ("M_max", "<f8")])

    def final(self, st):
        return dict(det_max=st["det_max"].copy(), symp_max=st["symp_max"].copy(), M_max=st["M_max"].copy())


# ------------------------------------------------------------------------------ StretchingNumbers
SN_MAX_HANDOFFS = 10 ** 9
SN_CONVENTION = ("cumulative spectra, D at every interval, columns renormalised separately, no orthogonalisation, "
                 "overflow cells of width h")


def spectrum_cell(z, lo, hi, B):
    """The cell of a stretching number (spectrum_cell, src/Indicators/stretching.jl:104-108): 0 below lo, B + 1 above
    hi, else min(floor((z - lo) / h) + 1, B) with h = (hi - lo) / B, in float64."""
    z, lo, hi = float(z), float(lo), float(hi)
    if z < lo:
        return 0
    if z > hi:
        return int(B) + 1
    return min(int(np.floor((z - lo) / ((hi - lo) / float(B)))) + 1, int(B))


def spectral_distance(N1, N2, h):
    """Spectral distance D = sqrt(S) / (k sqrt(h)) of two count vectors (spectral_distance, stretching.jl:189-204):
    S the exact integer sum of squared differences, k the common total; 0 for total 0."""
    a, b = np.asarray(N1, dtype=np.int64), np.asarray(N2, dtype=np.int64)
    if a.shape != b.shape:
        raise ValueError(f"count vectors of lengths {a.size} and {b.size}")
    k1, k2 = int(a.sum()), int(b.sum())
    if k1 != k2:
        raise ValueError(f"count vectors must have equal totals, got {k1} and {k2}")
    if k1 == 0:
        return 0.0
    S = int(((a - b) ** 2).sum())
    return float(np.sqrt(np.float64(S))) / (k1 * float(np.sqrt(np.float64(h))))


def spectrum_density(counts, h):
    """Normalised spectrum N_c / (k h) of cells 1..B of a count column of length B + 2 (stretching.jl:212-215)."""
    c = np.asarray(counts, dtype=np.int64)
    k = int(c.sum())
    return np.zeros(c.size - 2) if k == 0 else c[1:-1].astype(np.float64) / (k * h)


def stretching_range(sys):
    """Derived symmetric range (-hi, hi) for the standard maps (stretching_range, stretching.jl:235-241): hi just above
    L_sigma = ln sigma(K), sigma(K) = (sqrt(S2) + sqrt(S2 - 4)) / 2 with S2 = 2 K^2 + 2 |K| + 5, padded by 8 eps
    relative, computed at 256 bits and rounded once."""
    import mpmath as mp
    with mp.workprec(256):
        K = mp.mpf(float(sys.K))
        S2 = 2 * K * K + 2 * abs(K) + 5
        L = mp.log((mp.sqrt(S2) + mp.sqrt(S2 - 4)) / 2)
        hi = float(L * (1 + 8 * mp.mpf(2) ** -52))
    return (-hi, hi)


class StretchingNumbers(Reducer):
    """StretchingNumbers(; range, bins=1000, interval=100, min_per_cell=100, replace=True) for k = 2
    (src/Indicators/stretching.jl:34-147): cumulative spectra of the per-handoff log growth of the two columns in a
    UInt32 workspace of 2 (bins + 2) counts per orbit, and the compensated sums of a and a^2."""
    kind = L.CL_RED_STRETCHING

    def __init__(self, *, range, bins=1000, interval=100, min_per_cell=100, replace=True):
        lo, hi = (float(v) for v in range)
        if not (np.isfinite(lo) and np.isfinite(hi) and lo < hi):
            raise ValueError(f"range must be finite with s_lo < s_hi, got ({lo}, {hi})")
        if int(bins) < 1:
            raise ValueError(f"bins must be at least 1, got {bins}")
        if int(bins) > 2 ** 30:
            raise ValueError(f"bins must fit the descriptor (at most 2^30), got {bins}")
        if int(interval) < 1:
            raise ValueError(f"interval must be at least 1 map step, got {interval}")
        if int(min_per_cell) < 0:
            raise ValueError(f"min_per_cell must be nonnegative, got {min_per_cell}")
        self.lo, self.hi, self.bins = lo, hi, int(bins)
        self.interval, self.min_per_cell, self.replace = int(interval), int(min_per_cell), bool(replace)

    def descriptor(self, frame):
        r = super().descriptor(frame)
        r.lo, r.hi, r.bins = self.lo, self.hi, self.bins
        return r

    def check(self, d, k):
        _k2("StretchingNumbers", k)

    def workspace(self, d, k):
        """reducer_workspace (stretching.jl:93-96): 2 (bins + 2) UInt32 entries per orbit."""
        return 2 * (self.bins + 2)

    def check_run(self, tau, N):
        """check_run (stretching.jl:66-70): interval a multiple of tau, at most 10^9 handoffs."""
        if self.interval % int(tau):
            raise ValueError(f"interval {self.interval} must be a multiple of tau = {tau}")
        if int(N) // int(tau) > SN_MAX_HANDOFFS:
            raise ValueError(f"a run of {int(N) // int(tau)} handoffs exceeds the limit of 10^9 handoffs")

    def check_frame(self, W0, orbit=0):
        """check_frame (stretching.jl:258-275): both columns unit within 8 eps, then not parallel; returns the larger
        norm defect, else raises NumericalHealthError("initial frame") at step 0 naming the column."""
        from .tangent import NumericalHealthError
        W0 = np.asarray(W0, dtype=np.float64)
        if W0.ndim != 2 or W0.shape[1] != 2:
            raise ValueError("StretchingNumbers needs a frame with 2 columns")
        thr = 8 * EPS64
        defect = 0.0
        for i in range(2):
            dv = abs(_snorm(W0[:, i]) - 1)
            if not (np.isfinite(dv) and dv <= thr):
                raise NumericalHealthError("initial frame", float(dv), thr, 0, orbit, "cpu", i + 1)
            defect = max(defect, dv)
        para = min(_snorm(W0[:, 0] - W0[:, 1]), _snorm(W0[:, 0] + W0[:, 1]))
        if not para > thr:
            raise NumericalHealthError("initial frame", float(para), thr, 0, orbit, "cpu", 2)
        return defect

    def dtype(self, d, k):
        return np.dtype([("s1", "<f8", 2), ("s2", "<f8", 2), ("q1", "<f8", 2), ("q2", "<f8", 2), ("value", "<f8"),
                         ("n", "<i8"), ("K", "<i8"), ("step", "<i8"), ("code", "i1"), ("column", "i1"),
                         ("pad", "u1", 6)])

    def final(self, st, ws=None):
        """reducer_final (stretching.jl:165-185): (n, handoffs, sum, comp, sum2, comp2, mean, std) and, with the
        workspace rows ws (m, 2 (bins + 2)), D, undersampled, counts ((bins + 2) x 2 per orbit), below, above."""
        n, K = st["n"], st["K"]
        zero = K == 0
        nn = np.where(zero, 1, n).astype(np.float64)
        tau = (n // np.where(zero, 1, K)).astype(np.float64)
        out = dict(n=n.copy(), handoffs=K.copy(),
                   sum=np.stack([st["s1"][:, 0], st["s2"][:, 0]], 1), comp=np.stack([st["s1"][:, 1], st["s2"][:, 1]], 1),
                   sum2=np.stack([st["q1"][:, 0], st["q2"][:, 0]], 1),
                   comp2=np.stack([st["q1"][:, 1], st["q2"][:, 1]], 1))
        mean, std = [], []
        for a, q in (("s1", "q1"), ("s2", "q2")):
            m_ = np.where(zero, 0.0, _acc(st[a]) / nn)
            v = _acc(st[q]) / (np.where(zero, 1.0, tau) * nn) - m_ * m_
            mean.append(m_)
            std.append(np.where(zero, 0.0, np.sqrt(np.maximum(v, 0.0))))
        out["mean"], out["std"] = np.stack(mean, 1), np.stack(std, 1)
        if ws is not None:
            B = self.bins
            ws = np.asarray(ws, dtype=np.uint32).reshape(len(n), 2, B + 2)
            counts = ws.transpose(0, 2, 1).copy()
            h = (self.hi - self.lo) / B
            out["counts"] = counts
            # spectral_distance per orbit, vectorised: both columns count one entry per handoff, so the totals agree
            c64 = counts.astype(np.int64)
            k1 = c64[:, :, 0].sum(axis=1)
            S = ((c64[:, :, 0] - c64[:, :, 1]) ** 2).sum(axis=1)
            with np.errstate(invalid="ignore", divide="ignore"):
                D = np.sqrt(S.astype(np.float64)) / (k1.astype(np.float64) * float(np.sqrt(np.float64(h))))
            out["D"] = np.where(k1 == 0, 0.0, D)
            out["undersampled"] = K < self.min_per_cell * B
            out["below"] = counts[:, 0, :].astype(np.int64)
            out["above"] = counts[:, B + 1, :].astype(np.int64)
        return out

    def error(self, row):
        v = float(row["value"])
        if int(row["code"]) == 2:
            return "handoffs", v, float(SN_MAX_HANDOFFS), int(row["step"]), 0
        return "deviation matrix", v, 0.0 if np.isfinite(v) else np.inf, int(row["step"]), int(row["column"])

    def provenance(self):
        return dict(indicator="stretching", convention=SN_CONVENTION, spectrum_bins=self.bins, spectrum_lo=self.lo,
                    spectrum_hi=self.hi, spectrum_cell_width=(self.hi - self.lo) / self.bins,
                    spectrum_interval=self.interval, spectrum_min_per_cell=self.min_per_cell, replace=self.replace)


def _snorm(v):
    """safe_norm of a float64 vector (src/Foundation/norm.jl:8-12), NaN propagating."""
    v = np.asarray(v, dtype=np.float64)
    a = np.abs(v)
    if np.any(np.isnan(a)):
        return float("nan")
    m = float(a.max())
    if not (np.isfinite(m) and m > 0):
        return m
    s = 0.0
    for x in v:
        s += (x / m) * (x / m)
    return m * float(np.sqrt(s))
