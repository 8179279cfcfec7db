"""System descriptors: the maps and flows the kernels integrate, named as in the reference."""
import math

import numpy as np

from . import _lib as L


class StandardMap:
    """StandardMap(K): p' = p + K sin x, x' = x + p' on lifted coordinates (src/Systems/standard_map.jl:8-43)."""
    dimension = 2
    state_length = 2

    def __init__(self, K):
        self.K = float(K)

    def descriptor(self):
        s = L.cl_system()
        s.kind = L.CL_SYS_STANDARD_MAP
        s.K = self.K
        return s

    def __repr__(self):
        return f"StandardMap({self.K!r})"


class StandardMapChain:
    """StandardMapChain(N, K, eps; boundary): N coupled standard maps, state (q_1..q_N, p_1..p_N) on lifted coordinates,
    p_j' = p_j + K sin q_j + eps sum_nbr sin(q_j - q_k), q_j' = q_j + p_j' (src/Systems/standard_map_chain.jl:1-53).
    `boundary` is "open" or "ring" (ring needs N >= 3). The reference bounds N <= 100; this build accepts N <= 128
    and its device kernels take N in {8, 16, 32, 64, 128}."""

    def __init__(self, N, K, eps, boundary="open"):
        N = int(N)
        if boundary not in ("open", "ring"):
            raise ValueError(f"boundary must be 'open' or 'ring', got {boundary!r}")
        if not 2 <= N <= 128:
            raise ValueError(f"number of sites N must be in 2:128, got {N}")
        if boundary == "ring" and N < 3:
            raise ValueError(f"a ring needs N >= 3, got {N}")
        self.N, self.K, self.eps, self.boundary = N, float(K), float(eps), boundary
        self.dimension = 2 * N
        self.state_length = 2 * N

    def descriptor(self):
        s = L.cl_system()
        s.kind = L.CL_SYS_STANDARD_MAP_CHAIN
        s.N = self.N
        s.ring = 1 if self.boundary == "ring" else 0
        s.K = self.K
        s.eps = self.eps
        return s

    def __repr__(self):
        return f"StandardMapChain({self.N}, {self.K!r}, {self.eps!r}; boundary={self.boundary})"


class WrappedStandardMap:
    """WrappedStandardMap(K): the standard map on [0, 2pi)^2 with exact winding counts, state (x, p, m, n)
    (src/Systems/wrapped_standard_map.jl:9-81)."""
    dimension = 2
    state_length = 4

    def __init__(self, K):
        self.K = float(K)

    def descriptor(self):
        s = L.cl_system()
        s.kind = L.CL_SYS_WRAPPED_STANDARD_MAP
        s.K = self.K
        return s

    def __repr__(self):
        return f"WrappedStandardMap({self.K!r})"


class StandardMapK:
    """The standard map with its kick strength carried per orbit: state (x, p, K), each orbit stepping exactly as
    StandardMap(K) (src/Systems/standard_map.jl:18-43). No reference counterpart (the reference holds one K per
    system); it lets one launch cover an ensemble over K."""
    dimension = 2
    state_length = 3

    def descriptor(self):
        s = L.cl_system()
        s.kind = L.CL_SYS_STANDARD_MAP_K
        return s

    def __repr__(self):
        return "StandardMapK()"


class HenonHeilesGauss:
    """IntegratorMap(Gauss(Float64, 3; tol, maxiter), HenonHeilesFlow(), h): one step of the s = 3 Gauss-Legendre
    method on H = |p|^2/2 + |q|^2/2 + q1^2 q2 - q2^3/3, state (q1, q2, p1, p2)
    (src/Integrators/integrator_map.jl:12-23, src/Integrators/gauss.jl:49-69, src/Systems/flow_models.jl:30,101-103).
    `tol=None` is the reference default acceptance 8 eps max(1, |Y|)."""
    dimension = 4
    state_length = 4

    def __init__(self, h=0.01, tol=None, maxiter=50, stages=3):
        if stages != 3:
            raise ValueError("the kernels are built for the s = 3 Gauss method")
        if maxiter < 2:
            raise ValueError("maxiter must be at least 2")
        self.h = float(h)
        self.tol = float("nan") if tol is None else float(tol)
        self.maxiter = int(maxiter)
        self.stages = stages

    def descriptor(self):
        s = L.cl_system()
        s.kind = L.CL_SYS_HENON_HEILES_GAUSS
        s.stages = self.stages
        s.maxiter = self.maxiter
        s.h = self.h
        s.tol = self.tol
        return s

    def __repr__(self):
        return f"HenonHeilesGauss(h={self.h!r}, tol={self.tol!r}, maxiter={self.maxiter})"


def hh_energy(z):
    """H(z) of the Henon-Heiles flow in the reference's operation order (src/Systems/flow_models.jl:101)."""
    z = np.asarray(z, dtype=np.float64)
    q1, q2, p1, p2 = z[..., 0], z[..., 1], z[..., 2], z[..., 3]
    return ((p1 * p1 + p2 * p2) / 2 + (q1 * q1 + q2 * q2) / 2 + (q1 * q1) * q2) - (q2 * q2) * q2 / 3


TWO_PI = 2.0 * (3.141592653589793 + 1.2246467991473532e-16)


def _reduce(u, tp):
    """(u mod tp in [0, tp), the count) as the reference's _reduce (wrapped_standard_map.jl:16-29). A non-finite u
    follows Julia's arithmetic, where floor keeps NaN and +-Inf: the remainder is NaN and the count u / tp, so a
    non-finite start gives a NaN state instead of a conversion error."""
    if not math.isfinite(u):
        a = u / tp
        return u - tp * a, a
    a = math.floor(u / tp)
    r = u - tp * a
    if r < 0:
        r += tp
        a -= 1
    if r >= tp:
        r -= tp
        a += 1
    return r, float(a)


def torus_state(z):
    """State (x, p, m, n) of WrappedStandardMap for the lifted point z = (X, P)
    (src/Systems/wrapped_standard_map.jl:59-70)."""
    x, m = _reduce(float(z[0]), TWO_PI)
    p, n = _reduce(float(z[1]), TWO_PI)
    return (x, p, m, n)


class PairMap:
    """PairMap(sys): the product of a standard map with itself (src/Systems/pair.jl:1-68), state (state A; state B),
    dimension 4; each half steps by the base map and the tangent acts block by block. RunMonitor reads orbit A."""
    dimension = 4

    def __init__(self, sys):
        if not isinstance(sys, (StandardMap, WrappedStandardMap)):
            raise ValueError("the pair kernels run the standard map, lifted or wrapped")
        self.sys = sys
        self.K = sys.K
        self.state_length = 2 * sys.state_length

    def descriptor(self):
        s = L.cl_system()
        s.kind = L.CL_SYS_PAIR_WRAPPED_STANDARD_MAP if isinstance(self.sys, WrappedStandardMap) \
            else L.CL_SYS_PAIR_STANDARD_MAP
        s.K = self.K
        return s

    def __repr__(self):
        return f"PairMap({self.sys!r})"


def _wrap_diff(u, tp):
    """Minimal image of a coordinate difference (rli.jl:145): u - tp round(u / tp), ties to even; a non-finite
    difference stays non-finite (NaN), as Julia's round keeps NaN and Inf."""
    if not math.isfinite(u):
        return u - tp * (u / tp)
    return u - tp * round(u / tp)


def _safe_norm2(a, b):
    if math.isnan(a) or math.isnan(b):  # Julia's max propagates NaN
        return math.nan
    m = max(abs(a), abs(b))
    if not (math.isfinite(m) and m > 0):
        return m
    o = b if abs(a) >= abs(b) else a
    r = o / m
    return m * math.sqrt(1.0 + r * r)


def pair_distance(sys, z):
    """safe_norm of the minimal-image separation of the two halves of a pair state (rli.jl:156-166)."""
    h = len(z) // 2
    return _safe_norm2(_wrap_diff(float(z[h]) - float(z[0]), TWO_PI), _wrap_diff(float(z[h + 1]) - float(z[1]), TWO_PI))


class SeparationError(ValueError):
    """NumericalHealthError("separation", value, threshold, 0, orbit) of pair_start (rli.jl:189-202)."""

    def __init__(self, value, threshold, orbit=0):
        self.value, self.threshold, self.orbit = value, threshold, orbit
        super().__init__(f"separation: value {value!r}, threshold {threshold!r}, step 0, orbit {orbit}")


def pair_start(sys, x0, u, Dx, orbit=0):
    """Start of a pair on the base map `sys` (pair_start, rli.jl:180-202): state A = x0 (torus_state(x0) for the
    wrapped map), B = A + Dx u (reduced with A's counts when wrapped); the realized separation must be finite,
    positive and within [Dx/2, 2 Dx]. Returns (z, Dx, Dx_r, mode)."""
    dx = float(Dx)
    if not (math.isfinite(dx) and dx > 0):
        raise SeparationError(dx, 0.0 if math.isfinite(dx) else math.inf, orbit)
    wrapped = isinstance(sys, WrappedStandardMap)
    zA = torus_state(x0) if wrapped else (float(x0[0]), float(x0[1]))
    d0, d1 = dx * float(u[0]), dx * float(u[1])
    if wrapped:
        x, a = _reduce(zA[0] + d0, TWO_PI)
        p, b = _reduce(zA[1] + d1, TWO_PI)
        zB = (x, p, zA[2] + a, zA[3] + b)
    else:
        zB = (zA[0] + d0, zA[1] + d1)
    z = np.array(tuple(zA) + tuple(zB), dtype=np.float64)
    dr = pair_distance(sys, z)
    if not (math.isfinite(dr) and dr > 0):
        raise SeparationError(dr, 0.0 if math.isfinite(dr) else math.inf, orbit)
    if not (dx / 2 <= dr <= 2 * dx):
        raise SeparationError(dr, dx, orbit)
    return z, dx, dr, "wrapped" if wrapped else "lifted"


def pair_frame(xi):
    """The block frame of RLI (rli.jl:12-14): column 1 (xi; 0) for orbit A, column 2 (0; xi) for orbit B."""
    W = np.zeros((4, 2))
    W[0:2, 0] = xi
    W[2:4, 1] = xi
    return W

