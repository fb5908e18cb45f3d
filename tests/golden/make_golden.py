"""Writes tests/golden/golden.npz from the REFERENCE's own PocketFFT.

Run in the container that holds the reference tree, after
`make -C oracle ref` has built oracle/_ref/libpocketfft_ref.so from the
reference sources in place:

    python tests/golden/make_golden.py          # golden.npz (and golden_r2r.npz)
    python tests/golden/make_golden.py --r2r    # golden_r2r.npz only
    python tests/golden/make_golden.py --f32    # golden_f32.npz (T = float)

Inputs come from impulse_amd.synth (SplitMix64, seed recorded per case) and
are stored next to the outputs, so the fixture is self-contained data.
Outputs are what pocketfft::c2c / r2c / c2r (C++ path,
impulse/fft/cpp_pocketfft/pocketfft_hdronly.h:3272-3390) and rfft_forward
(C path, impulse/fft/c_pocketfft/pocketfft.c:2184) return.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from impulse_amd import synth  # noqa: E402
from oracle import ref  # noqa: E402

LENGTHS_FULL = list(range(1, 33)) + [48, 60, 64, 97, 100, 128, 243, 255, 256, 1009]
LENGTHS_FWD = [2310, 4095, 4096]


def main():
    if not ref.available():
        raise SystemExit("oracle/_ref/libpocketfft_ref.so is not built")
    g = {}
    for n in LENGTHS_FULL + LENGTHS_FWD:
        x = synth.complex_normal((n,), seed=1000 + n)
        g[f"c2c_{n}_in"] = x
        g[f"c2c_{n}_fwd"] = ref.c2c(x, [0], True)
        if n in LENGTHS_FULL:
            g[f"c2c_{n}_bwd"] = ref.c2c(x, [0], False, 0.5)
        r = synth.uniform((n,), -1.0, 1.0, seed=2000 + n)
        g[f"r2c_{n}_in"] = r
        g[f"r2c_{n}_out"] = ref.r2c(r, [0], True)
        g[f"hc_{n}_out"] = ref.rfft_packed(r, True)
        if n in LENGTHS_FULL:
            g[f"r2c_{n}_bwd"] = ref.r2c(r, [0], False, 2.0)
            c = synth.complex_normal((n // 2 + 1,), seed=3000 + n)
            g[f"c2r_{n}_in"] = c
            g[f"c2r_{n}_out"] = ref.c2r(c, n, [0], False, 1.0 / n)
            g[f"c2r_{n}_fwdflag"] = ref.c2r(c, n, [0], True)
    # n-D cases, including a strided (every other column) input view
    a = synth.complex_normal((12, 40), seed=7)
    g["nd_c2c_in"] = a
    g["nd_c2c_01"] = ref.c2c(a[:, ::2], [0, 1], True)
    g["nd_c2c_10_bwd"] = ref.c2c(a[:, ::2], [1, 0], False, 0.25)
    b = synth.uniform((6, 10, 9), -1.0, 1.0, seed=8)
    g["nd_r2c_in"] = b
    g["nd_r2c_012"] = ref.r2c(b, [0, 1, 2], True)
    g["nd_r2c_20"] = ref.r2c(b, [2, 0], True)
    g["nd_r2c_1_bwd"] = ref.r2c(b, [1], False)
    h = synth.complex_normal((6, 10, 5), seed=9)
    g["nd_c2r_in"] = h
    g["nd_c2r_012"] = ref.c2r(h, 9, [0, 1, 2], False, 1.0 / 540)
    g["nd_c2r_2"] = ref.c2r(h, 8, [2], False)
    # rows of the r2c [256, 4096] and [256, 4095] workloads (uniform[-1,1], seed 0)
    for n in (4096, 4095):
        rows = synth.uniform((256, n), -1.0, 1.0, seed=0)[:2]
        g[f"rows_{n}_in"] = rows
        g[f"rows_{n}_out"] = ref.r2c(rows, [1], True)
    out = os.path.join(HERE, "golden.npz")
    np.savez_compressed(out, **g)
    print(out, os.path.getsize(out), "bytes,", len(g), "arrays")
    r2r()


R2R_LENGTHS = [1, 2, 3, 4, 5, 6, 7, 8, 13, 16, 17, 31, 32, 33, 64, 97, 100, 128, 255, 256, 1009, 2048]
R2R_ORTHO = [2, 3, 5, 8, 17, 100]
FP_LENGTHS = [1, 2, 3, 4, 5, 6, 7, 8, 13, 16, 17, 31, 64, 97, 100, 255, 256, 1000]


def r2r():
    """pocketfft::dct / dst of every type on 1-D inputs (plain and ortho) and n-D views; r2r_fftpack (all
    flag pairs) and the separable / genuine Hartley transforms."""
    g = {}
    for n in R2R_LENGTHS:
        x = synth.uniform((n,), -1.0, 1.0, seed=4000 + n)
        g[f"r2r_{n}_in"] = x
        for name, fn in (("dct", ref.dct), ("dst", ref.dst)):
            for t in (1, 2, 3, 4):
                if name == "dct" and t == 1 and n == 1:
                    continue
                g[f"{name}{t}_{n}"] = fn(x, [0], t)
                if n in R2R_ORTHO:
                    g[f"{name}{t}_{n}_ortho"] = fn(x, [0], t, 0.5, True)
    # r2r_fftpack (all four flag pairs) and the separable Hartley transform, 1-D
    for n in FP_LENGTHS:
        x = synth.uniform((n,), -1.0, 1.0, seed=5000 + n)
        g[f"fp_{n}_in"] = x
        for r2h in (0, 1):
            for fw in (0, 1):
                g[f"fp{r2h}{fw}_{n}"] = ref.r2r_fftpack(x, [0], bool(r2h), bool(fw), 1.0 if fw else 0.5)
        g[f"hartley_{n}"] = ref.r2r_separable_hartley(x, [0])
    v = synth.uniform((6, 9, 10), -1.0, 1.0, seed=11)
    g["nd_r2r_in"] = v
    g["nd_fp11_20"] = ref.r2r_fftpack(v[:, ::2, :], [2, 0], True, True, 0.25)
    g["nd_fp00_1"] = ref.r2r_fftpack(v[::-1], [1], False, False)
    g["nd_fp01_02"] = ref.r2r_fftpack(v, [0, 2], False, True)
    g["nd_hartley_012"] = ref.r2r_separable_hartley(v, [0, 1, 2], 0.5)
    g["nd_genuine_12"] = ref.r2r_genuine_hartley(v, [1, 2])
    g["nd_genuine_201"] = ref.r2r_genuine_hartley(v[:, ::-1, :], [2, 0, 1], 2.0)
    for t in (1, 2, 3, 4):
        g[f"nd_dct{t}_20"] = ref.dct(v[:, ::2, :], [2, 0], t, 0.25, t == 2)
        g[f"nd_dst{t}_012"] = ref.dst(v[::-1], [0, 1, 2], t, 1.0, t == 3)
    out = os.path.join(HERE, "golden_r2r.npz")
    np.savez_compressed(out, **g)
    print(out, os.path.getsize(out), "bytes,", len(g), "arrays")


def f32():
    """pocketfft::c2c / r2c / c2r<float> (the reference's T = float instantiation, reached from impulse
    through FFTDesc[float32] / DataDesc[float32]) on float32 inputs: 1-D, n-D views and rows of the
    [256, 4096] / [256, 4095] workloads."""
    g = {}
    for n in LENGTHS_FULL + LENGTHS_FWD:
        x = synth.complex_normal((n,), seed=6000 + n).astype(np.complex64)
        g[f"c2c_{n}_in"] = x
        g[f"c2c_{n}_fwd"] = ref.c2c(x, [0], True)
        g[f"c2c_{n}_bwd"] = ref.c2c(x, [0], False, 0.5)
        r = synth.uniform((n,), -1.0, 1.0, seed=7000 + n).astype(np.float32)
        g[f"r2c_{n}_in"] = r
        g[f"r2c_{n}_out"] = ref.r2c(r, [0], True)
        g[f"r2c_{n}_bwd"] = ref.r2c(r, [0], False, 2.0)
        c = synth.complex_normal((n // 2 + 1,), seed=8000 + n).astype(np.complex64)
        g[f"c2r_{n}_in"] = c
        g[f"c2r_{n}_out"] = ref.c2r(c, n, [0], False, 1.0 / n)
    a = synth.complex_normal((12, 40), seed=17).astype(np.complex64)
    g["nd_c2c_in"] = a
    g["nd_c2c_01"] = ref.c2c(a[:, ::2], [0, 1], True)
    g["nd_c2c_10_bwd"] = ref.c2c(a[::-1, :], [1, 0], False, 0.25)
    b = synth.uniform((6, 10, 9), -1.0, 1.0, seed=18).astype(np.float32)
    g["nd_r2c_in"] = b
    g["nd_r2c_012"] = ref.r2c(b, [0, 1, 2], True)
    g["nd_r2c_20"] = ref.r2c(b, [2, 0], True)
    h = synth.complex_normal((6, 10, 5), seed=19).astype(np.complex64)
    g["nd_c2r_in"] = h
    g["nd_c2r_012"] = ref.c2r(h, 9, [0, 1, 2], False, 1.0 / 540)
    for n in (4096, 4095):
        rows = synth.uniform((256, n), -1.0, 1.0, seed=0)[:2].astype(np.float32)
        g[f"rows_{n}_in"] = rows
        g[f"rows_{n}_out"] = ref.r2c(rows, [1], True)
    out = os.path.join(HERE, "golden_f32.npz")
    np.savez_compressed(out, **g)
    print(out, os.path.getsize(out), "bytes,", len(g), "arrays")


if __name__ == "__main__":
    if "--r2r" in sys.argv:
        r2r()
    elif "--f32" in sys.argv:
        f32()
    else:
        main()
