"""Seeded synthetic inputs: SplitMix64, counter-based and vectorised.

Draw i (0-based) of seed s is  mix(s + (i + 1) * 0x9E3779B97F4A7C15)  with the
standard SplitMix64 finaliser; u = (z >> 11) * 2**-53 lies in [0, 1).
Normals use Box-Muller on consecutive draws (u1, u2):
  r = sqrt(-2 ln(1 - u1)),  z0 = r cos(2 pi u2),  z1 = r sin(2 pi u2).
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(n, seed=0, start=0):
    i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(n, seed=0, start=0):
    return (splitmix64(n, seed, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform(shape, lo=-1.0, hi=1.0, seed=0, dtype=np.float64):
    n = int(np.prod(shape))
    return (lo + (hi - lo) * uniform01(n, seed)).reshape(shape).astype(dtype, copy=False)


def normal(shape, seed=0, sigma=1.0, dtype=np.float64):
    n = int(np.prod(shape))
    m = (n + 1) // 2
    u = uniform01(2 * m, seed).reshape(m, 2)
    r = np.sqrt(-2.0 * np.log1p(-u[:, 0]))
    th = 2.0 * np.pi * u[:, 1]
    z = np.empty((m, 2))
    z[:, 0] = r * np.cos(th)
    z[:, 1] = r * np.sin(th)
    return (sigma * z.reshape(-1)[:n]).reshape(shape).astype(dtype, copy=False)


def complex_normal(shape, seed=0, dtype=np.complex128):
    """re, im iid N(0, 1), interleaved from one stream."""
    z = normal(tuple(shape) + (2,), seed)
    return (z[..., 0] + 1j * z[..., 1]).astype(dtype, copy=False)
