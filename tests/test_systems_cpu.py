"""Host helpers of the systems on non-finite input, as the reference's arithmetic gives them: torus_state of a NaN or
infinite point is a NaN state (Julia's floor keeps NaN and Inf, src/Systems/wrapped_standard_map.jl:16-29,59-70) and
pair_start of such a point raises the separation error (src/Indicators/rli.jl:180-202)."""
import math

import numpy as np
import pytest

from chaos_amd import StandardMap, WrappedStandardMap, torus_state, pair_start, SeparationError


@pytest.mark.parametrize("bad", [math.nan, math.inf, -math.inf])
def test_torus_state_non_finite(bad):
    x, p, m, n = torus_state((bad, 0.5))
    assert math.isnan(x) and p == 0.5 and n == 0.0
    assert math.isnan(m) or math.isinf(m)
    x, p, m, n = torus_state((1.0, bad))
    assert x == 1.0 and math.isnan(p)


@pytest.mark.parametrize("sys", [StandardMap(1.0), WrappedStandardMap(1.0)])
@pytest.mark.parametrize("bad", [math.nan, math.inf])
def test_pair_start_non_finite_raises_separation(sys, bad):
    with pytest.raises(SeparationError):
        pair_start(sys, (bad, 0.3), (1.0, 0.0), 1e-8, orbit=4)


def test_torus_state_finite_unchanged():
    x, p, m, n = torus_state((7.0, -1.0))
    tp = 2.0 * (math.pi + 1.2246467991473532e-16)
    assert np.isclose(x, 7.0 - tp) and m == 1.0 and np.isclose(p, tp - 1.0) and n == -1.0
