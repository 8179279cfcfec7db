"""Weighted Birkhoff averages (src/Indicators/wbirkhoff.jl, wbirkhoff_reducer.jl): the reducer `WeightedBirkhoff`
for `propagate`, its observables, the pure function `weighted_birkhoff` on a stored series and the index
`birkhoff_index`.

The reducer reads K observables of the orbit point at every handoff and keeps, per horizon N, the weighted sums of
the first N samples (window 1) and of the next N (window 2) with the weight w(jj) = exp(-1 / (t^p (1 - t)^p)),
t = min(jj, N - jj) / N, and the largest |f|. `reducer_final` gives avg = A / B per window and horizon and the
index of the pair. The weights come from the library (`cl_birkhoff_weights`), the same arithmetic the kernels use,
so the pure function on the sampled series equals the reducer bit for bit.
"""
import ctypes
import math

import numpy as np

from . import _lib as L
from .reducers import Reducer

WB_MIN_WINDOW = 16
WB_FLOOR_MULTIPLIER = 32
WB_MAX_ORDER = 4
WB_MAX_OBSERVABLES = L.CL_WB_MAX_OBSERVABLES   # the reference allows 31; beyond 8 there is no kernel
EPS64 = 2.0 ** -52
FLOATMIN = 2.0 ** -1022
TWO_PI = 2 * (3.141592653589793 + 1.2246467991473532e-16)


class CosObservable:
    """cos(k x[i]) of the orbit point, i 1-based (CosObservable, wbirkhoff.jl:17-28)."""

    def __init__(self, i, k=1):
        self.i, self.k = int(i), int(k)

    def __call__(self, x):
        return math.cos(float(self.k) * float(x[self.i - 1]))

    def name(self):
        return f"cos({self.k} x[{self.i}])"

    def code(self):
        return L.CL_OBS_COS, self.i - 1, self.k


class SinObservable(CosObservable):
    """sin(k x[i]) of the orbit point (SinObservable, wbirkhoff.jl:30-40)."""

    def __call__(self, x):
        return math.sin(float(self.k) * float(x[self.i - 1]))

    def name(self):
        return f"sin({self.k} x[{self.i}])"

    def code(self):
        return L.CL_OBS_SIN, self.i - 1, self.k


class LiftedObservable:
    """x[iq] + 2 pi x[im], the coordinate rebuilt from its winding count (LiftedObservable, wbirkhoff.jl:42-56)."""

    def __init__(self, iq, im):
        self.iq, self.im = int(iq), int(im)

    def __call__(self, x):
        return float(x[self.iq - 1]) + TWO_PI * float(x[self.im - 1])

    def name(self):
        return f"lifted(x[{self.iq}], x[{self.im}])"

    def code(self):
        return L.CL_OBS_LIFTED, self.iq - 1, self.im - 1


def observable_name(o):
    return o.name() if hasattr(o, "name") else "function"


def _windows(window):
    ws = (int(window),) if isinstance(window, (int, np.integer)) else tuple(int(w) for w in window)
    if not ws or any(w < WB_MIN_WINDOW for w in ws):
        raise ValueError(f"window must be at least {WB_MIN_WINDOW}, got {window}")
    return ws


def _order(order):
    if not (isinstance(order, (int, np.integer)) and 1 <= int(order) <= WB_MAX_ORDER):
        raise ValueError(f"birkhoff_order must be an integer from 1 to {WB_MAX_ORDER} (the centre weight exp(-4^p) "
                         f"underflows above), got {order}")
    return int(order)


def weights(N, order=1):
    """The unnormalised weights w(jj), jj = 0..N-1, of a horizon N (_wb, src/Indicators/naff_reducer.jl:17), from the
    library's host code (the kernels' arithmetic)."""
    w = np.empty(int(N), dtype=np.float64)
    rc = L.load().cl_birkhoff_weights(int(N), int(order), w.ctypes.data)
    if rc:
        raise ValueError(f"no weights for N = {N}, order = {order}")
    return w


def _csum_add(s, c, y):
    """Neumaier's compensated addition (add, src/Foundation/Foundation.jl:48-58) on float64 scalars."""
    t = s + y
    c = c + (((s - t) + y) if abs(s) >= abs(y) else ((y - t) + s))
    return t, c


def weighted_birkhoff(series, window, *, order=1, weights_=None):
    """Weighted Birkhoff averages of a stored series (weighted_birkhoff, wbirkhoff.jl:139-162): series (n,) or (n, K),
    window N or a tuple of N. Returns dict(avg1, avg2, S (K x H), B (H,)): avg = A / B with A = sum w f and B = sum w
    (Neumaier sums, window 2 restarting the weight index, one B for both windows), S the largest |f| of the 2N samples.
    `weights_` (length N, one window only) replaces the default weights."""
    ws = _windows(window)
    p = _order(order)
    f = np.asarray(series, dtype=np.float64)
    if f.ndim == 1:
        f = f[:, None]
    n, K = f.shape
    if n < 2 * max(ws):
        raise ValueError(f"series: {n} samples given, 2 window = {2 * max(ws)} needed")
    if not np.all(np.isfinite(f[:2 * max(ws)])):
        raise ValueError(f"series has a non-finite value among its first {2 * max(ws)} samples")
    if weights_ is not None and not (len(ws) == 1 and len(weights_) == ws[0]):
        raise ValueError(f"weights: a vector of length window = {ws[0]} is needed, with a single window")
    H = len(ws)
    avg1, avg2, S, B = np.empty((K, H)), np.empty((K, H)), np.empty((K, H)), np.empty(H)
    for h, N in enumerate(ws):
        w = weights(N, p) if weights_ is None else np.asarray(weights_, dtype=np.float64)
        bs = bc = 0.0
        for x in w:
            if x != 0.0:
                bs, bc = _csum_add(bs, bc, float(x))
        B[h] = bs + bc
        if not B[h] > 0:
            raise ValueError(f"birkhoff_order = {p}: the summed weight is 0" if weights_ is None
                             else "weights sum to 0 (or less): the weight total must be positive")
        for k in range(K):
            a = [[0.0, 0.0], [0.0, 0.0]]
            s_ = 0.0
            for j in range(2 * N):
                i = 1 if j >= N else 0
                wj = float(w[j - i * N])
                fj = float(f[j, k])
                s_ = max(s_, abs(fj))
                if wj != 0.0:
                    a[i][0], a[i][1] = _csum_add(a[i][0], a[i][1], wj * fj)
            avg1[k, h] = (a[0][0] + a[0][1]) / B[h]
            avg2[k, h] = (a[1][0] + a[1][1]) / B[h]
            S[k, h] = s_
    return dict(avg1=avg1, avg2=avg2, S=S, B=B)


def birkhoff_index(avg1, avg2, S, floor_multiplier=WB_FLOOR_MULTIPLIER):
    """The index of an average pair (birkhoff_index, wbirkhoff.jl:144-171), elementwise: delta = avg1 - avg2,
    phi = max(32 eps S, floatmin), absdig = -log10 max(|delta|, phi), reldig = -log10(max(|delta|, phi) / m) with
    m = |avg1| / 2 + |avg2| / 2 where m > phi, dig = max(absdig, reldig) (absdig where reldig is NaN), below_floor =
    |delta| <= phi; S = 0 gives NaN digits and `zero`; a non-finite delta gives NaN digits and no flags."""
    if not floor_multiplier > 0:
        raise ValueError(f"floor_multiplier must be positive, got {floor_multiplier}")
    a1, a2, S = (np.asarray(v, dtype=np.float64) for v in (avg1, avg2, S))
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        delta = a1 - a2
        phi = np.maximum(float(floor_multiplier) * EPS64 * S, FLOATMIN)
        a = np.abs(delta)
        below = a <= phi
        e = np.maximum(a, phi)
        m = np.abs(a1) / 2 + np.abs(a2) / 2
        absdig = -np.log10(e)
        reldig = np.where(m > phi, -np.log10(e / m), np.nan)
        dig = np.where(np.isnan(reldig), absdig, np.maximum(absdig, reldig))
    zero = S == 0
    bad = ~np.isfinite(a)
    nan = np.full(np.shape(delta), np.nan)
    absdig = np.where(zero | bad, nan, absdig)
    reldig = np.where(zero | bad, nan, reldig)
    dig = np.where(zero | bad, nan, dig)
    below = np.where(bad & ~zero, False, below)
    return dict(delta=delta, absdig=absdig, reldig=reldig, dig=dig, below_floor=below, zero=zero)


class WeightedBirkhoff(Reducer):
    """WeightedBirkhoff(; window, observables, birkhoff_order=1) (src/Indicators/wbirkhoff_reducer.jl:4-116): a reader
    of the orbit point for any k, the last reducer of its chain; window N or a tuple of up to 16 N (each >= 16); up to 8
    observables (CosObservable, SinObservable, LiftedObservable). A run has 2 max(window) handoffs (check_run)."""
    kind = L.CL_RED_BIRKHOFF

    def __init__(self, *, window, observables, birkhoff_order=1):
        self.window = _windows(window)
        if len(self.window) > L.CL_WB_MAX_WINDOWS:
            raise ValueError(f"window: at most {L.CL_WB_MAX_WINDOWS} horizons")
        self.birkhoff_order = _order(birkhoff_order)
        obs = tuple(observables)
        if not (1 <= len(obs) <= 31) or not all(hasattr(o, "code") for o in obs):
            raise ValueError("observables must be a nonempty tuple of at most 31 observables of the module's kinds")
        if len(obs) > WB_MAX_OBSERVABLES:
            raise ValueError(f"observables: the kernels take at most {WB_MAX_OBSERVABLES}")
        self.observables = obs
        self._desc = L.cl_birkhoff()
        self._desc.H, self._desc.K, self._desc.order = len(self.window), len(obs), self.birkhoff_order
        for h, N in enumerate(self.window):
            self._desc.window[h] = N
        for k, o in enumerate(obs):
            self._desc.kind[k], self._desc.a[k], self._desc.b[k] = o.code()

    def descriptor(self, frame):
        r = super().descriptor(frame)
        r.birkhoff = ctypes.pointer(self._desc)
        return r

    def check_run(self, tau, N, x0=None):
        """check_run (wbirkhoff_reducer.jl:102-116): tau >= 1, and N / tau = 2 max(window) handoffs; with x0 every
        observable is evaluated there and must be a finite real number."""
        if int(tau) < 1:
            raise ValueError(f"tau must be at least 1, got {tau}")
        need = 2 * max(self.window)
        if not (int(N) % int(tau) == 0 and int(N) // int(tau) == need):
            raise ValueError(f"the run of {N} steps at tau = {tau} does not have the {need} handoffs (2 window) that "
                             f"window = {self.window} needs")
        if x0 is not None:
            for k, o in enumerate(self.observables):
                v = o(x0)
                if not np.isfinite(v):
                    raise ValueError(f"observables[{k + 1}] is not a finite real number at the initial point: {v}")

    def dtype(self, d, k):
        H, K = len(self.window), len(self.observables)
        cell = np.dtype([("A1", "<f8", 2), ("A2", "<f8", 2), ("S", "<f8")])
        hor = np.dtype([("B", "<f8", 2), ("cells", cell, (K,))])
        return np.dtype([("hor", hor, (H,)), ("n", "<i8"), ("handoffs", "<i4"), ("code", "<i2"), ("pad", "u1", 2)])

    def final(self, st):
        """reducer_final (wbirkhoff_reducer.jl:76-86): n, handoffs, complete, avg1, avg2, S (m x K x H), B (m x H),
        delta, absdig, reldig, dig, below_floor, zero. avg1 from handoff N, avg2 and the index from 2N (NaN before)."""
        H, K = len(self.window), len(self.observables)
        s = st["handoffs"].astype(np.int64)
        hor = st["hor"]
        B = hor["B"][..., 0] + hor["B"][..., 1]                        # (m, H)
        A1 = hor["cells"]["A1"][..., 0] + hor["cells"]["A1"][..., 1]   # (m, H, K)
        A2 = hor["cells"]["A2"][..., 0] + hor["cells"]["A2"][..., 1]
        S = hor["cells"]["S"].transpose(0, 2, 1).copy()                 # (m, K, H)
        Nw = np.array(self.window)
        with np.errstate(invalid="ignore", divide="ignore"):
            avg1 = np.where((s[:, None] >= Nw)[:, :, None], A1 / B[:, :, None], np.nan).transpose(0, 2, 1)
            avg2 = np.where((s[:, None] >= 2 * Nw)[:, :, None], A2 / B[:, :, None], np.nan).transpose(0, 2, 1)
        done = np.broadcast_to((s[:, None] >= 2 * Nw)[:, None, :], avg1.shape)
        ix = birkhoff_index(avg1, avg2, S)
        out = dict(n=st["n"].copy(), handoffs=s, complete=s >= 2 * max(self.window), avg1=avg1, avg2=avg2, S=S, B=B)
        for key in ("delta", "absdig", "reldig", "dig"):
            out[key] = np.where(done, ix[key], np.nan)
        out["below_floor"] = np.where(done, ix["below_floor"], False)
        out["zero"] = np.where(done, ix["zero"], False)
        return out

    def error(self, row):
        return "observable", float("nan"), math.inf, int(row["n"]), int(row["code"])

    def provenance(self):
        return dict(indicator="weighted-birkhoff", estimator="weighted-birkhoff-average", weight="wb_exp",
                    birkhoff_order=self.birkhoff_order, window=self.window,
                    observables=tuple(observable_name(o) for o in self.observables),
                    handoffs=2 * max(self.window), floor_multiplier=WB_FLOOR_MULTIPLIER, min_window=WB_MIN_WINDOW)


# ---------------------------------------------------------------- WeightedPhaseAdvance
class PlanarObservable:
    """PlanarObservable(iq, ip) (naff_reducer.jl:88-97): the sample x[iq] + i x[ip] (1-based indices), its increment the
    principal argument of z' conj(z)."""

    def __init__(self, iq, ip):
        self.iq, self.ip = int(iq), int(ip)

    def code(self):
        return L.CL_OBS_PLANAR, self.iq - 1, self.ip - 1

    def name(self):
        return "planar"


class LiftedAngleObservable:
    """LiftedAngleObservable(iq, im=0) (naff_reducer.jl:99-112): a wrapped coordinate x[iq] and its winding count x[im]
    (0: none), its increment (q' - q) + 2 pi (m' - m)."""

    def __init__(self, iq, im=0):
        self.iq, self.im = int(iq), int(im)

    def code(self):
        return L.CL_OBS_LIFTED_ANGLE, self.iq - 1, self.im - 1

    def name(self):
        return "lifted-angle"


EPS64 = 2.0 ** -52


def circle_distance(a, b):
    """abs(d - round(d)), d = a - b (naff_map.jl:8), round half to even."""
    d = np.asarray(a) - np.asarray(b)
    return np.abs(d - np.round(d))


def frequency_index(delta, floor_multiplier=32):
    """(D, below_floor) of NF10 (naff_map.jl:26-29): log10(max(delta, 32 eps)), delta <= 32 eps."""
    fl = float(floor_multiplier) * EPS64
    delta = np.asarray(delta, dtype=np.float64)
    return np.log10(np.maximum(delta, fl)), delta <= fl


class WeightedPhaseAdvance(Reducer):
    """WeightedPhaseAdvance(; window, observable, birkhoff_order=1) (src/Indicators/naff_reducer.jl:142-254): a reader of
    the orbit point for any k; window N or a tuple of 1 or 2 N (each >= 16; the kernels take 2 at most), a run has
    2 max(window) + 1 handoffs (check_run)."""
    kind = L.CL_RED_PHASE_ADVANCE

    def __init__(self, *, window, observable, birkhoff_order=1):
        self.window = _windows(window)
        if len(self.window) > 2:
            raise ValueError("window: the kernels take 1 or 2 horizons")
        if not isinstance(observable, (PlanarObservable, LiftedAngleObservable)):
            raise ValueError("observable must be a PlanarObservable or a LiftedAngleObservable")
        self.observable = observable
        self.birkhoff_order = _order(birkhoff_order)
        self._desc = L.cl_birkhoff()
        self._desc.H, self._desc.K, self._desc.order = len(self.window), 1, self.birkhoff_order
        for h, N in enumerate(self.window):
            self._desc.window[h] = N
        self._desc.kind[0], self._desc.a[0], self._desc.b[0] = observable.code()

    def descriptor(self, frame):
        r = super().descriptor(frame)
        r.alg = len(self.window)
        r.birkhoff = ctypes.pointer(self._desc)
        return r

    def check_run(self, tau, N, d=None):
        """check_run (naff_reducer.jl:236-242): N / tau = 2 max(window) + 1 handoffs."""
        need = 2 * max(self.window) + 1
        if not (int(N) % int(tau) == 0 and int(N) // int(tau) == need):
            raise ValueError(f"the run of {N} steps at tau = {tau} does not have the {need} handoffs (2 window + 1) "
                             f"that window = {self.window} needs")

    def dtype(self, d, k):
        hor = np.dtype([("A1", "<f8", 2), ("B1", "<f8", 2), ("A2", "<f8", 2), ("B2", "<f8", 2)])
        return np.dtype([("hor", hor, (len(self.window),)), ("prev1", "<f8"), ("prev2", "<f8"), ("n", "<i8"),
                         ("handoffs", "<i4"), ("nwrap", "<i2"), ("code", "i1"), ("pad", "u1")])

    def final(self, st):
        """reducer_final (naff_reducer.jl:220-230): nu1 once handoffs >= N + 1, nu2 from 2N + 1 (NaN before), delta, D,
        below_floor per horizon; complete, near_pi."""
        tp = 2.0 * math.pi
        s = st["handoffs"].astype(np.int64)
        hor = st["hor"]
        tot = lambda a: a[..., 0] + a[..., 1]
        nu1 = np.stack([np.where(s >= N + 1, tot(hor["A1"][:, h]) / (tp * tot(hor["B1"][:, h])), np.nan)
                        for h, N in enumerate(self.window)], axis=1)
        nu2 = np.stack([np.where(s >= 2 * N + 1, tot(hor["A2"][:, h]) / (tp * tot(hor["B2"][:, h])), np.nan)
                        for h, N in enumerate(self.window)], axis=1)
        with np.errstate(invalid="ignore"):
            delta = circle_distance(nu1, nu2)
            D, below = frequency_index(delta)
        D = np.where(np.isnan(delta), np.nan, D)
        below = np.where(np.isnan(delta), False, below)
        return dict(n=st["n"].copy(), handoffs=s, complete=s >= 2 * max(self.window) + 1,
                    near_pi=st["nwrap"].astype(np.int64), nu1=nu1, nu2=nu2, delta=delta, D=D, below_floor=below)

    def error(self, row):
        return "series", float("nan"), float("inf"), int(row["n"]), 1
