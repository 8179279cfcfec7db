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
    """Classical Gram-Schmidt (src/Lyapunov/orthonormalize.jl:250-251)."""
    code, name = L.CL_QR_CGS, "CGS"


class MGS(Orthonormalizer):
    """Modified Gram-Schmidt (src/Lyapunov/orthonormalize.jl:252-253)."""
    code, name = L.CL_QR_MGS, "MGS"


class RGS(Orthonormalizer):
    """Repeated Gram-Schmidt, eta = 1/sqrt 2, at most 4 passes (src/Lyapunov/orthonormalize.jl:254-255)."""
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
    """QRSpectrum(alg; orthogonality_tol=8.0) (src/Lyapunov/qr_spectrum.jl:161-165)."""
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
        """(n, sum, comp, lambda, G_max, orth_max, count_max), qr_spectrum.jl:235-240."""
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
    """Benettin() for k = 1 (src/Lyapunov/benettin.jl:143-185)."""
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
        Ybar20 = np.where(zero, 0.0, tau * Q2 / nn ** 3)
        return dict(n=st["n"].copy(), handoffs=st["K"].copy(), Y=Y, Ybar=Ybar, Ybar20=Ybar20,
                    sum=c[:, :, 0, 0].copy(), comp=c[:, :, 0, 1].copy(), R1=R1, R2=R2, Q=Q, Q2=Q2)

    def error(self, row):
        v = float(row["value"])
        return "deviation matrix", v, 0.0 if np.isfinite(v) else np.inf, int(row["step"]), int(row["column"])

    def provenance(self):
        return dict(indicator="megno", convention="weights at the end of the window, mean over the handoffs 1..K")


class SALI(Reducer):
    """SALI() for k = 2, replaces W by unit columns (src/Indicators/sali.jl:147-201)."""
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
    """FLI(; replace=True) for k = 2 (src/Indicators/fli.jl:292-362)."""
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
    """GALI() for 2 <= k <= d (src/Indicators/gali.jl:190-374)."""
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
    threshold for a non-finite value (residuals.jl:46-50)."""
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
