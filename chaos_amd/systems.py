"""System descriptors: the maps and flows the kernels integrate, named as in the reference."""
import math

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


class WrappedStandardMap:
    """WrappedStandardMap(K): the standard map on [0, 2pi)^2 with exact winding counts, state (x, p, m, n)
    (src/Systems/wrapped_standard_map.jl:75-123)."""
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


TWO_PI = 2.0 * (3.141592653589793 + 1.2246467991473532e-16)


def _reduce(u, tp):
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
    (src/Systems/wrapped_standard_map.jl:131-136)."""
    x, m = _reduce(float(z[0]), TWO_PI)
    p, n = _reduce(float(z[1]), TWO_PI)
    return (x, p, m, n)
