"""Seeded synthetic inputs: SplitMix64, counter-based and vectorised.

Draw i (0-based) of seed s is  mix(s + (i + 1) * 0x9E3779B97F4A7C15)  with the
standard SplitMix64 finaliser; u = (z >> 11) * 2**-53 lies in [0, 1).
Normals use Box-Muller on consecutive draws (u1, u2):
  r = sqrt(-2 ln(1 - u1)),  z0 = r cos(2 pi u2),  z1 = r sin(2 pi u2).
"""
import math

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


def uniform_device(n, lo=-1.0, hi=1.0, seed=0, dtype=None, device="cuda", chunk=1 << 26, start=0):
    """uniform(n, lo, hi, seed) computed on a torch device, bit-identical to the
    numpy version (SplitMix64 in wrapping int64 arithmetic, logical shifts by
    masking), in chunks of `chunk` draws."""
    import torch
    dtype = dtype or torch.float64
    out = torch.empty(n, dtype=dtype, device=device)

    def c64(v):   # uint64 constant as the int64 with the same bits
        return v - (1 << 64) if v >= 1 << 63 else v

    gamma, m1, m2 = c64(0x9E3779B97F4A7C15), c64(0xBF58476D1CE4E5B9), c64(0x94D049BB133111EB)

    def lsr(z, k):
        return (z >> k) & ((1 << (64 - k)) - 1)

    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        i = torch.arange(start + s + 1, start + s + m + 1, dtype=torch.int64, device=device)
        z = c64(int(seed) % (1 << 64)) + i * gamma
        z = (z ^ lsr(z, 30)) * m1
        z = (z ^ lsr(z, 27)) * m2
        z = z ^ lsr(z, 31)
        u = lsr(z, 11).to(torch.float64) * (2.0 ** -53)
        out[s:s + m] = (lo + (hi - lo) * u).to(dtype)
    return out


def _smooth7(lo, hi):
    vals = []
    p2 = 1
    while p2 <= hi:
        p3 = p2
        while p3 <= hi:
            p5 = p3
            while p5 <= hi:
                p7 = p5
                while p7 <= hi:
                    if p7 >= lo:
                        vals.append(p7)
                    p7 *= 7
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return np.array(sorted(vals), dtype=np.int64)


def _primes_upto(n):
    sieve = np.ones(n + 1, dtype=bool)
    sieve[:2] = False
    for p in range(2, int(n ** 0.5) + 1):
        if sieve[p]:
            sieve[p * p::p] = False
    return np.nonzero(sieve)[0]


def ragged_lengths(count, lo=64, hi=262144, smooth_frac=0.7, seed=0, prime_min=1009):
    """Signal lengths of a ragged batch.

    Draws (u0, u1, u2) per signal from the SplitMix64 stream of `seed`:
      * u0 < smooth_frac: the 2^a 3^b 5^c 7^d number in [lo, hi] nearest in
        log scale to T = lo (hi/lo)^u1;
      * otherwise p m with a prime p >= prime_min (> 1000): T = prime_min
        (hi/prime_min)^u1 (no shorter length has such a factor),
        m = 1 + floor(u2 min(16, T / prime_min)), p = the largest prime
        <= T / m, so that prime_min <= p and p m <= T <= hi.
    """
    u = uniform01(3 * count, seed).reshape(count, 3)
    sm = _smooth7(lo, hi)
    logs = np.log(sm)
    primes = _primes_upto(hi)
    out = np.empty(count, dtype=np.int64)
    for i in range(count):
        u0, u1, u2 = u[i]
        if u0 < smooth_frac:
            t = np.log(lo) + u1 * (np.log(hi) - np.log(lo))
            j = int(np.searchsorted(logs, t))
            cands = [k for k in (j - 1, j) if 0 <= k < len(sm)]
            out[i] = sm[min(cands, key=lambda k: abs(logs[k] - t))]
        else:
            T = prime_min * (hi / prime_min) ** u1
            m = 1 + int(u2 * min(16.0, T / prime_min))
            m = max(1, min(m, int(T // prime_min)))
            k = int(np.searchsorted(primes, int(T // m), side="right")) - 1
            out[i] = int(primes[k]) * m
    return out


def complex_normal_device(n, seed=0, device="cuda", dtype=None, chunk=1 << 25):
    """complex_normal((n,), seed) computed on a torch device: value i is the
    Box-Muller pair of draws 2i, 2i+1 (re = r cos, im = r sin), as in
    complex_normal (equal to it up to the device's rounding of log, sqrt, cos
    and sin)."""
    import torch
    dtype = dtype or torch.complex128
    out = torch.empty(n, dtype=dtype, device=device)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        u = uniform_device(2 * m, 0.0, 1.0, seed=seed, device=device, start=2 * s).view(m, 2)
        r = torch.sqrt(-2.0 * torch.log1p(-u[:, 0]))
        th = 2.0 * math.pi * u[:, 1]
        out[s:s + m] = torch.complex(r * torch.cos(th), r * torch.sin(th)).to(dtype)
    return out
