"""Synthetic initial conditions of the benchmark workloads (numpy, host side).

Per-orbit draws are keyed like the frames: key = mix(mix(seed) + orbit) with the splitmix64
finaliser (src/Tangent/initial.jl:1-7), xored with a constant so that the initial-condition stream
and the frame stream of an orbit are distinct; u(c) = ((mix(key + (c+1) phi) >> 12) + 1/2) 2^-52.
"""
import numpy as np

PHI = np.uint64(0x9E3779B97F4A7C15)
IC_SALT = np.uint64(0x243F6A8885A308D3)
TWO_PI = 6.283185307179586


def mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + PHI
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def orbit_keys(seed, m, orbit0=1):
    ids = np.arange(orbit0, orbit0 + m, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return mix64(mix64(np.uint64(seed)) + ids)


def uniforms(keys, c):
    with np.errstate(over="ignore"):
        v = mix64(keys + np.uint64(c + 1) * PHI)
    return ((v >> np.uint64(12)).astype(np.float64) + 0.5) * 2.0 ** -52


def torus_uniform(seed, m, orbit0=1):
    """(m, 2) points 2 pi u with u uniform on [0,1)^2: the unit torus mapped onto the map's 2 pi torus."""
    keys = orbit_keys(seed, m, orbit0) ^ IC_SALT
    return np.stack([TWO_PI * uniforms(keys, 0), TWO_PI * uniforms(keys, 1)], axis=1)


def grid_centres(n):
    """(n*n, 2) cell centres 2 pi (i + 1/2)/n of an n x n grid; orbit i*n + j has x index i, p index j."""
    c = TWO_PI * ((np.arange(n, dtype=np.float64) + 0.5) / n)
    X, P = np.meshgrid(c, c, indexing="ij")
    return np.stack([X.ravel(), P.ravel()], axis=1)


def uniform_k(seed, m, lo, hi, orbit0=1):
    """Per-orbit parameter drawn uniform on [lo, hi) from the orbit's initial-condition stream."""
    keys = orbit_keys(seed, m, orbit0) ^ IC_SALT
    return lo + (hi - lo) * uniforms(keys, 2)


def hh_section_starts(seed, m, E, section="q2", orbit0=1, max_tries=64):
    """(m, 4) starts on the energy surface H = E of the Henon-Heiles flow, uniform on a Poincare section.

    section="q2": the literal reading of the workload (y = q2 = 0, p_y = p2 > 0 solved from E): (q1, p1) drawn
    uniform in the box |q1|, |p1| <= sqrt(2E) and accepted where 2E - q1^2 - p1^2 > 0, then p2 = sqrt(that).
    section="q1": the reference's convention (src/Systems/henon_heiles.jl:62-68): q1 = 0, (q2, p2) drawn in the
    box [a, b] x [-sqrt(2E), sqrt(2E)] (b, a the roots of q2^2/2 - q2^3/3 = E around 0) and p1 = sqrt(2(E - p2^2/2
    - q2^2/2 + q2^3/3)) > 0. Rejection draws use the orbit's initial-condition stream, attempt t drawing
    uniforms 2t+3 and 2t+4."""
    keys = orbit_keys(seed, m, orbit0) ^ IC_SALT
    out = np.full((m, 4), np.nan)
    todo = np.ones(m, dtype=bool)
    r = np.sqrt(2 * E)
    if section == "q1":
        from scipy.optimize import brentq
        u = lambda q: q * q / 2 - q * q * q / 3 - E  # q^3 as the reference's literal power: (q q) q
        a, b = (brentq(u, -1.0, 0.0), brentq(u, 0.0, 1.0)) if E < 1 / 6 else (-0.5, 1.0)
    for t in range(max_tries):
        if not todo.any():
            break
        u1, u2 = uniforms(keys, 2 * t + 3), uniforms(keys, 2 * t + 4)
        if section == "q2":
            q1, p1 = -r + 2 * r * u1, -r + 2 * r * u2
            rem = 2 * E - q1 * q1 - p1 * p1
            ok = todo & (rem > 0)
            out[ok] = np.stack([q1, np.zeros(m), p1, np.sqrt(np.where(rem > 0, rem, 0))], axis=1)[ok]
        else:
            q2, p2 = a + (b - a) * u1, -r + 2 * r * u2
            rem = 2 * (E - p2 * p2 / 2 - (q2 * q2 / 2 - q2 * q2 * q2 / 3))
            ok = todo & (rem > 0) & (np.abs(q2) <= 1)
            out[ok] = np.stack([np.zeros(m), q2, np.sqrt(np.where(rem > 0, rem, 0)), p2], axis=1)[ok]
        todo &= ~ok
    if todo.any():
        raise RuntimeError(f"{int(todo.sum())} starts found no admissible point in {max_tries} draws")
    return out


def chain_starts(seed, m, N, orbit0=1):
    """(m, 2N) chain states 2 pi u, u uniform on [0,1)^{2N}: positions and momenta of every site on the torus."""
    keys = orbit_keys(seed, m, orbit0) ^ IC_SALT
    return np.stack([TWO_PI * uniforms(keys, c) for c in range(2 * N)], axis=1)


def chain_step_counts(seed, m, lo, hi, tau, orbit0=1):
    """Per-orbit step counts uniform on the multiples of tau in [lo, hi], from draw 2N+... of the orbit's stream
    (an index past any state draw: 1000)."""
    keys = orbit_keys(seed, m, orbit0) ^ IC_SALT
    nlo, nhi = -(-lo // tau), hi // tau
    return tau * (nlo + np.floor(uniforms(keys, 1000) * (nhi - nlo + 1)).astype(np.int64))
