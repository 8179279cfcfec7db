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
