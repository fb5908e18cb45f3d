"""Resampling from impulse/signal.nim over the C ABI: ``upfirdn`` runs on the GPU
(impulse_fft_upfirdn_*); ``resample`` and the filter design it needs
(``kaiser``, ``firls``, ``generate_resampling_filter``,
``reduce_resampling_rates``) are host code with the reference's names, argument
meanings and errors.

Arrays are numpy arrays (host one-shot calls) or torch device tensors (device
calls on the current stream), rank 1.  Integer inputs are filtered in FP64 and
rounded back, complex inputs are filtered as their real and imaginary parts.
"""
import ctypes
import math
from collections import OrderedDict

import numpy as np

from . import _lib
from ._lib import check, lib


def _is_torch(x):
    return type(x).__module__.startswith("torch")


class UpfirdnPlan:
    """One filter, rates and signal length (impulse_fft_upfirdn_create); reusable."""

    def __init__(self, nt, h, up=1, down=1, batch=1, algorithm=_lib.UPFIRDN_AUTO):
        hh = np.ascontiguousarray(h, dtype=np.float64).ravel()
        p = ctypes.c_void_p()
        check(lib().impulse_fft_upfirdn_create(ctypes.byref(p), int(batch), int(nt), hh.ctypes.data, hh.size,
                                               int(up), int(down), int(algorithm)))
        self.handle = p
        self._destroy = lib().impulse_fft_upfirdn_destroy
        self.batch, self.nt, self.up, self.down = int(batch), int(nt), int(up), int(down)
        n, a, L, b = ctypes.c_size_t(), ctypes.c_int(), ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().impulse_fft_upfirdn_info(p, ctypes.byref(n), ctypes.byref(a), ctypes.byref(L), ctypes.byref(b)))
        self.nout, self.algorithm, self.fft_length, self.blocks = n.value, a.value, L.value, b.value

    def __del__(self):
        # bound at creation: at interpreter exit the module globals may be gone
        h, destroy = getattr(self, "handle", None), getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self.handle = None

    def execute(self, t, y=None):
        """t: [batch, nt] float64 (numpy or torch device tensor); returns y [batch, nout]."""
        if _is_torch(t):
            import torch
            if not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
                raise ValueError("device upfirdn takes a contiguous float64 device tensor")
            if t.numel() != self.batch * self.nt:
                raise ValueError("signal length does not match the plan")
            if y is None:
                y = torch.empty((self.batch, self.nout), dtype=torch.float64, device=t.device)
            check(lib().impulse_fft_upfirdn_execute(self.handle, t.data_ptr(), y.data_ptr(),
                                                    torch.cuda.current_stream().cuda_stream))
            return y
        t = np.ascontiguousarray(t, dtype=np.float64)
        if t.size != self.batch * self.nt:
            raise ValueError("signal length does not match the plan")
        if y is None:
            y = np.empty((self.batch, self.nout), np.float64)
        check(lib().impulse_fft_upfirdn_execute_host(self.handle, t.ctypes.data, y.ctypes.data))
        return y


_plans = OrderedDict()


def _plan(nt, h, up, down, batch, algorithm):
    key = (nt, h.tobytes(), up, down, batch, algorithm)
    p = _plans.get(key)
    if p is None:
        p = UpfirdnPlan(nt, h, up, down, batch, algorithm)
        _plans[key] = p
        while len(_plans) > 16:
            _plans.popitem(last=False)
    else:
        _plans.move_to_end(key)
    return p


def upfirdn(t, h, up=1, down=1, algorithm=_lib.UPFIRDN_AUTO):
    """Upsample, FIR filter and downsample a rank-1 signal (impulse/signal.nim:545-594):
    ``t.upsample(up).convolve(h, mode = full, down = down)``.  The input comes first, the filter
    second (the Matlab order the reference documents)."""
    if _is_torch(t):
        import torch
        if t.dim() != 1:
            raise ValueError("upfirdn takes a rank-1 signal")
        hn = h.detach().cpu().numpy() if _is_torch(h) else np.asarray(h)
        if np.iscomplexobj(hn):
            # a complex filter runs as its real and imaginary taps, as on the host path
            hr, hi = np.ascontiguousarray(hn.real, np.float64), np.ascontiguousarray(hn.imag, np.float64)
            if np.any(hi):
                return upfirdn(t, hr, up, down, algorithm) + 1j * upfirdn(t, hi, up, down, algorithm)
            hn = hr
        hh = np.ascontiguousarray(hn, dtype=np.float64).ravel()
        if t.is_complex():
            parts = torch.stack([t.real, t.imag]).to(torch.float64).contiguous()
            y = _plan(t.shape[0], hh, up, down, 2, algorithm).execute(parts)
            yc = torch.complex(y[0], y[1])
            return yc if t.dtype == torch.complex128 else yc.to(t.dtype)
        y = _plan(t.shape[0], hh, up, down, 1, algorithm).execute(t.to(torch.float64).contiguous())[0]
        if not t.dtype.is_floating_point:
            # integer signal: an integer filter gives integer outputs (rounded, not truncated);
            # a real filter gives float64, as on the host path
            if np.issubdtype(hn.dtype, np.integer):
                return torch.round(y).to(torch.promote_types(t.dtype, torch.from_numpy(np.zeros(1, hn.dtype)).dtype))
            return y
        return y if t.dtype == torch.float64 else y.to(t.dtype)
    t = np.asarray(t)
    h = np.asarray(h)
    if t.ndim != 1:
        raise ValueError("upfirdn takes a rank-1 signal")
    if np.iscomplexobj(h):
        hr, hi = np.ascontiguousarray(h.real, np.float64), np.ascontiguousarray(h.imag, np.float64)
        if np.any(hi):
            a = upfirdn(t, hr, up, down, algorithm)
            b = upfirdn(t, hi, up, down, algorithm)
            return a + 1j * b
        h = hr
    hh = np.ascontiguousarray(h, dtype=np.float64).ravel()
    if np.iscomplexobj(t):
        y = _plan(t.size, hh, up, down, 2, algorithm).execute(np.stack([t.real, t.imag]))
        return y[0] + 1j * y[1]
    y = _plan(t.size, hh, up, down, 1, algorithm).execute(t)[0]
    if np.issubdtype(t.dtype, np.integer) and np.issubdtype(h.dtype, np.integer):
        return np.rint(y).astype(np.result_type(t.dtype, h.dtype))
    return y


# ---------------------------------------------------------------------------
# filter design (host)
# ---------------------------------------------------------------------------

def i0(x):
    """Modified Bessel function of the first kind, order 0 (numpy's Clenshaw-Chebyshev routine, the
    algorithm impulse/signal.nim:92-133 ports)."""
    return np.i0(x)


def kaiser(size=10, beta=5.0):
    """Kaiser window of ``size`` points (impulse/signal.nim:135-227):
    w[n] = I0(beta sqrt(1 - ((n - a) / a)^2)) / I0(beta), a = (size - 1) / 2."""
    if size < 0:
        raise AssertionError("The size of the Kaiser window must be non-negative")
    if size == 1:
        return np.ones(1)
    n = np.arange(float(size))
    a = (size - 1.0) / 2.0
    return i0(beta * np.sqrt(1.0 - ((n - a) / a) ** 2)) / i0(beta)


def _band_integrals(k, lo, hi, slope, icpt, cosine):
    """Integral over [lo, hi] of (slope f + icpt) times cos(2 pi k f) (or sin), k > 0."""
    w = 2.0 * np.pi * k
    if cosine:
        # [(s f + b) sin(w f) / w + s cos(w f) / w^2]
        return (slope * (np.cos(w * hi) - np.cos(w * lo)) / (w * w)
                + hi * (slope * hi + icpt) * np.sinc(2.0 * k * hi) - lo * (slope * lo + icpt) * np.sinc(2.0 * k * lo))
    # [-(s f + b) cos(w f) / w + s sin(w f) / w^2]
    return (slope * (np.sin(w * hi) - np.sin(w * lo)) / (w * w)
            + ((slope * lo + icpt) * np.cos(w * lo) - (slope * hi + icpt) * np.cos(w * hi)) / w)


def _gram(k, lo, hi, cosine):
    """Integral over [lo, hi] of cos(2 pi k f) cos(2 pi l f) (sin sin when not cosine), all k, l."""
    kp = 2.0 * (k[:, None] + k[None, :])
    km = 2.0 * (k[:, None] - k[None, :])
    sgn = 1.0 if cosine else -1.0
    return 0.5 * (hi * (np.sinc(km * hi) + sgn * np.sinc(kp * hi)) - lo * (np.sinc(km * lo) + sgn * np.sinc(kp * lo)))


def firls(fir_order, bands, desired, weights=None, symmetric=True, fs=2.0):
    """Least-squares linear-phase FIR design of ``fir_order + 1`` taps (impulse/signal.nim:229-543).

    ``bands`` / ``desired``: band edges (0 .. fs/2) and the gains at those edges, as [k, 2] or flat
    even-length arrays; ``weights``: one per band.  Type I / II (symmetric) or III / IV (anti-symmetric)
    by the order's parity.  The amplitude response A(f) = sum_k a_k cos(2 pi k f) (sin for the
    anti-symmetric types) minimises sum_bands w_b int (A - D)^2 df with D linear in each band: the
    normal equations G a = r are solved, or, as the reference does when the weights are all equal and the
    bands leave no gaps, taken as diagonal (a = 4 w r, the constant term halved)."""
    if fs <= 0:
        raise ValueError(f"Sampling frequency fs must be positive but got {fs}")
    bands = np.asarray(bands, dtype=np.float64).ravel()
    desired = np.asarray(desired, dtype=np.float64).ravel()
    nyq = fs / 2.0
    if bands.max() > nyq or bands.min() < 0.0:
        raise ValueError(f"Some frequency values are outside of the valid [0.0, {nyq}] range:\n{bands}")
    if bands.size % 2:
        raise ValueError(f"Frequency band tensor length ({bands.size}) is not even")
    if bands.size != desired.size:
        raise ValueError(f"Frequency and desired gain tensors must have the same length "
                         f"({bands.size}!={desired.size})")
    even = fir_order % 2 == 0
    if not even and bands[-1] == 1.0 and desired[-1] != 0.0:
        raise ValueError(f"Filter order ({fir_order}) must be even when the last frequency is 1.0 and its "
                         f"desired gain is not 0.0")
    w = np.ones(bands.size // 2) if weights is None or len(weights) == 0 else np.asarray(weights, np.float64)
    if bands.size != 2 * w.size:
        raise ValueError(f"Weigth tensor length ({w.size}) must be half the length of the frequency band and "
                         f"desired gain pair tensor lenghts ({bands.size})")
    df = np.diff(bands)
    if np.any(df < 0.0):
        raise ValueError("Frequency band tensor values must increasing monotonically")
    bands = bands / fs
    diagonal = bool(np.all(w == w[0])) and (df.size == 1 or bool(np.all(df[1::2] == 0.0)))
    half = fir_order // 2
    if symmetric:
        k = np.arange(half + 1.0) + (0.0 if even else 0.5)
    else:
        k = np.arange(1.0, half + 1.0) if even else np.arange(half + 1.0) + 0.5
    has_const = symmetric and even            # k = 0 term (cosine basis only)
    kk = k[1:] if has_const else k            # the k > 0 terms
    r = np.zeros(k.size)
    G = np.zeros((k.size, k.size))
    for b in range(w.size):
        lo, hi = bands[2 * b], bands[2 * b + 1]
        slope = (desired[2 * b + 1] - desired[2 * b]) / (hi - lo)
        icpt = desired[2 * b] - slope * lo
        wb = abs(w[b])
        if has_const:
            r[0] += wb * (icpt * (hi - lo) + slope / 2.0 * (hi * hi - lo * lo))
        r[1 if has_const else 0:] += wb * _band_integrals(kk, lo, hi, slope, icpt, symmetric)
        if not diagonal:
            G += wb * _gram(k, lo, hi, symmetric)
    if diagonal:
        a = 4.0 * w[0] * r
        if has_const:
            a[0] /= 2.0
    else:
        a = np.linalg.solve(G, r)
    if symmetric:
        if even:
            return np.concatenate([a[:0:-1] * 0.5, a[:1], a[1:] * 0.5])
        return 0.5 * np.concatenate([a[::-1], a])
    if even:
        return 0.5 * np.concatenate([-a[::-1], [0.0], a])
    return 0.5 * np.concatenate([-a[::-1], a])


def reduce_resampling_rates(up, down):
    """impulse/signal.nim:596-611: divide both rates by their greatest common divisor."""
    g = math.gcd(up, down)
    return up // g, down // g


def generate_resampling_filter(up, down, fir_order_factor=10, beta=5.0):
    """impulse/signal.nim:613-649: a firls low-pass of order 2 fir_order_factor max(up, down), cut-off
    1 / max(up, down), windowed by kaiser(order + 1, beta), scaled to sum to ``up``; ``fir_order_factor``
    0 gives ``up`` ones (step upsampling)."""
    if fir_order_factor == 0:
        return np.ones(up)
    m = max(up, down)
    fc = 1.0 / m
    order = 2 * fir_order_factor * m
    h = firls(order, [0.0, fc, fc, 1.0], [1.0, 1.0, 0.0, 0.0])
    h = h * kaiser(order + 1, beta)
    return h * (float(up) / h.sum())


def adjust_filter_position(h, input_len, result_len, up, down):
    """impulse/signal.nim:651-688: leading zeros put the filter's middle point on a kept sample,
    trailing zeros make upfirdn long enough; returns (filter, delay in output samples)."""
    mid = (len(h) - 1.0) / 2.0
    lead = int(math.floor(down - (mid - down * math.floor(mid / down))))
    mid += lead
    delay = int(math.floor(math.ceil(mid) / down))
    filtered_len = (input_len - 1) * up + len(h) + lead
    trail = 0
    while delay + result_len >= math.ceil((filtered_len + trail) / down):
        trail += 1
    return np.concatenate([np.zeros(lead), np.asarray(h, np.float64), np.zeros(trail)]), delay


def resample(t, h=None, up=1, down=1, fir_order_factor=10, beta=5.0, algorithm=_lib.UPFIRDN_AUTO):
    """Resample by up / down (impulse/signal.nim:690-809).

    With ``h`` given (signal.nim:690-744) the rates are used as they are; without it
    (signal.nim:746-809) they are reduced first and the filter comes from generate_resampling_filter.
    Keeps the ceil(len(t) up / down) samples aligned with the input."""
    n = int(t.shape[0])
    if h is None:
        if up == down:
            return t.clone() if _is_torch(t) else np.array(t, copy=True)
        up, down = reduce_resampling_rates(up, down)
        h = generate_resampling_filter(up, down, fir_order_factor, beta)
    else:
        if up == 1 and down == 1:
            return t.clone() if _is_torch(t) else np.array(t, copy=True)
        h = np.squeeze(h.detach().cpu().numpy() if _is_torch(h) else np.asarray(h))
        if h.ndim != 1:
            raise ValueError(f"Squeezed filter rank ({h.ndim}) must be 1")
        if np.iscomplexobj(h) and not np.any(h.imag):
            h = h.real
    result_len = int(math.ceil(n * up / down))
    h, delay = adjust_filter_position(h, n, result_len, up, down)
    y = upfirdn(t, h, up, down, algorithm)
    return y[delay:delay + result_len]
