"""Matrix-core complex FFTs (impulse_fft_mfma_*) against the FP64 oracle, over
the powers of two from 256 to 2^21.

Input is complex64; the reference is the oracle's c2c of the same values
widened to complex128.  Bars (relative L2 per transform), stated here:
  * "bf16x" (bf16 slices, exact FP32 accumulation, FP64 combine):  1e-11
  * "f32"   (FP32 MFMA, split-FP32 twiddles, complex64 output):    2e-6
The FP32 bar is FP32 rounding of about four stages (2^-24 ~ 6e-8 each, plus
the complex64 output rounding); the bf16x bar is set by the five-slice
truncation (~2^-40 per stage).
At the full size (8192 transforms of 65536, re/im N(0,1)), Parseval holds on
every transform and 16 sampled transforms match the FP64 drop-in plan.
"""
import numpy as np
import pytest
import torch

import oracle
from impulse_amd import _lib, synth
from impulse_amd import fft as F
from impulse_amd.mfma import MfmaFFT

pytestmark = pytest.mark.gpu

BAR = {"bf16x": 1e-11, "f32": 2e-6}


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.set_device(0)
    yield
    torch.cuda.synchronize()


def _rel_rows(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


# 256: one pass; 512 .. 4096 (and 8192 for F32): one pass with the length/256-point row DFTs in
# LDS (tiles of whole signals); 65536: both four-step passes on the matrix cores; the other
# powers of two: the matrix-core column pass and FP64 row transforms; odd batches leave a
# partial last tile
@pytest.mark.parametrize("prec", ["bf16x", "f32"])
@pytest.mark.parametrize("length,batch", [(256, 37), (256, 16), (65536, 3), (512, 7), (1024, 9), (2048, 5),
                                          (4096, 5), (8192, 3), (16384, 2), (131072, 2), (1 << 20, 1),
                                          (1 << 21, 1)])
@pytest.mark.parametrize("forward", [True, False])
def test_against_oracle(prec, length, batch, forward):
    print(f"case mfma {prec} L={length} batch={batch} forward={forward}", flush=True)
    x = synth.complex_normal((batch, length), seed=31).astype(np.complex64)
    fct = 1.0 if forward else 1.0 / length
    plan = MfmaFFT(length, batch, prec)
    y = plan(torch.from_numpy(x).cuda(), forward=forward, fct=fct)
    torch.cuda.synchronize()
    ref = oracle.c2c(x.astype(np.complex128), [1], forward, fct)
    err = _rel_rows(y.cpu().numpy().astype(np.complex128), ref)
    print(f"  max rel-L2 {err.max():.3e}  median {np.median(err):.3e}")
    assert err.max() <= BAR[prec]


@pytest.mark.parametrize("prec", ["bf16x", "f32"])
def test_closed_forms(prec):
    L = 256
    x = np.zeros((3, L), np.complex64)
    x[0, 0] = 1.0                  # impulse -> all ones
    x[1, :] = 2.0                  # constant -> 2L delta
    x[2, 5] = 1.0                  # shifted impulse -> exp(-2 pi i 5 k / L)
    y = MfmaFFT(L, 3, prec)(torch.from_numpy(x).cuda()).cpu().numpy()
    k = np.arange(L)
    want = np.stack([np.ones(L), 2.0 * L * (k == 0), np.exp(-2j * np.pi * 5 * k / L)])
    tol = 1e-13 if prec == "bf16x" else 1e-6
    assert np.max(np.abs(y - want)) <= tol * 2 * L


def test_zero_and_tiny_scale():
    # an all-zero transform stays zero; values near the bottom of the exponent range keep their accuracy
    x = np.zeros((2, 256), np.complex64)
    x[1] = (synth.complex_normal((256,), seed=4) * 1e-30).astype(np.complex64)
    y = MfmaFFT(256, 2, "bf16x")(torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.all(y[0] == 0)
    ref = oracle.c2c(x[1:].astype(np.complex128), [1], True, 1.0)
    assert _rel_rows(y[1:], ref).max() <= 1e-11


def test_rejects_bad_arguments():
    for bad in (128, 768, 1 << 22):   # below 256, not a power of two, above 2^21
        with pytest.raises(_lib.FFTError):
            MfmaFFT(bad, 4, "f32")
    with pytest.raises(_lib.FFTError):
        MfmaFFT(256, 0, "f32")
    p = MfmaFFT(256, 2, "f32")
    with pytest.raises(ValueError):
        p(torch.zeros(256, dtype=torch.complex64, device="cuda"))


def test_workspace_only_for_two_pass_lengths():
    # 256 and the one-pass lengths (512 .. 4096, and 8192 for F32, whose 32-column tile is one
    # whole signal) hold no intermediate; longer ones, one intermediate of the output type
    for L in (256, 512, 4096, 8192):
        assert MfmaFFT(L, 3, "f32").workspace_bytes == 0
    assert MfmaFFT(4096, 3, "bf16x").workspace_bytes == 0
    assert MfmaFFT(16384, 3, "f32").workspace_bytes == 16384 * 3 * 8
    assert MfmaFFT(8192, 3, "bf16x").workspace_bytes == 8192 * 3 * 16


@pytest.mark.parametrize("prec", ["bf16x", "f32"])
def test_full_size_8192x65536(prec):
    print(f"case mfma full size {prec}", flush=True)
    B, L = 8192, 65536
    x = synth.complex_normal_device(B * L, seed=0, dtype=torch.complex64)
    plan = MfmaFFT(L, B, prec)
    y = plan(x)
    torch.cuda.synchronize()
    xin = x.view(B, L)
    yv = y.view(B, L)
    ex = (xin.real.double() ** 2 + xin.imag.double() ** 2).sum(1)
    ey = (yv.real.double() ** 2 + yv.imag.double() ** 2).sum(1) / L
    parseval = float(((ey - ex).abs() / ex).max())
    # sampled transforms against the FP64 drop-in plan
    rng = np.random.default_rng(3)
    rows = rng.choice(B, 16, replace=False)
    xs = xin[rows].to(torch.complex128).contiguous()
    ref = torch.empty_like(xs)
    p = F.get_plan(_lib.C2C, [16, L], [L * 16, 16], [L * 16, 16], [1], True)
    p.execute(xs.data_ptr(), ref.data_ptr(), 1.0, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    err = (torch.linalg.vector_norm(yv[rows].to(torch.complex128) - ref, dim=1) /
           torch.linalg.vector_norm(ref, dim=1)).cpu().numpy()
    print(f"  Parseval max rel {parseval:.3e}; vs FP64 plan: max rel-L2 {err.max():.3e} median {np.median(err):.3e}")
    assert err.max() <= BAR[prec]
    assert parseval <= (1e-12 if prec == "bf16x" else 1e-5)
