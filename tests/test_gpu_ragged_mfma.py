"""Ragged batches with the matrix-core precision (impulse_fft_ragged_set_precision,
IMPULSE_FFT_RAGGED_MFMA_BF16X): the radix-16 stages of transforms that span whole
wavefronts run as exact bf16-slice products on the MFMA units.

Bar: 1e-10 relative L2 per signal against the oracle's r2c of the signal widened
to float64 (the matrix-core form's target); the slice truncation is ~2^-40 of
the largest value per stage, so the measured error sits near 1e-12.
* the coverage batch of every branch under all 18 configurations the autotuner
  may pick (radix order bit 1 is cleared by the precision: powers of two as
  radix-16 stages);
* that the matrix cores really ran: signals whose transforms have radix-16
  stages differ from the FP64 precision's bins (and stay within the bar),
  while signals too short to span a wavefront are bit-identical to FP64;
* odd pairs, scale and conjugation, switching back to FP64, graph replays;
* the 20,000-signal batch: Parseval on every signal and 48 sampled signals
  against the oracle at the bar.
"""
import numpy as np
import pytest
import torch

import oracle
from impulse_amd import fft as F
from impulse_amd import ragged, synth

from test_gpu_ragged import CONFIGS, COVER, PAIRS, _check_against_oracle, _inputs

pytestmark = pytest.mark.gpu

BAR = 1e-10


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.set_device(0)
    yield
    torch.cuda.synchronize()


def _run(b, xd, precision, fct=1.0):
    b.set_precision(precision)
    yd = torch.full((b.out_elems,), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    b.execute(xd, yd, fct=fct)
    torch.cuda.synchronize()
    return yd


@pytest.mark.parametrize("cfg", CONFIGS)
def test_cover_lengths_every_config(cfg):
    print(f"case ragged mfma cover cfg={cfg}", flush=True)
    lengths = np.array(COVER, dtype=np.int64)
    b = ragged.RaggedBatch(lengths)
    x = _inputs(lengths, seed=21)
    xd = torch.from_numpy(x).cuda()
    b.configure(*cfg)
    y64 = _run(b, xd, "fp64").cpu().numpy()
    ymf = _run(b, xd, "bf16x")
    worst = _check_against_oracle(b, x, ymf, range(len(lengths)), bar=BAR)
    print(f"  worst rel-L2 {worst:.2e}")
    ymf = ymf.cpu().numpy()
    changed = [int(n) for n, o in zip(b.lengths, b.out_offsets)
               if not np.array_equal(ymf[o:o + n // 2 + 1], y64[o:o + n // 2 + 1])]
    print(f"  signals whose bins differ from FP64: {changed}")
    # n = 4096: M = 2048 = 16 16 8, LDS-resident under every configuration (one transform over the
    # whole workgroup), so two of its stages ran on the matrix cores: its error is the slices'
    # (~1e-12), far above FP64 rounding (~1e-16).  n = 16 and 26 (M = 8, 13) have no radix-16 stage.
    i = COVER.index(4096)
    o = int(b.out_offsets[i])
    ref = oracle.r2c(x[int(b.in_offsets[i]):int(b.in_offsets[i]) + 4096].astype(np.float64), [0], True, 1.0)
    err4096 = float(np.linalg.norm(ymf[o:o + 2049] - ref) / np.linalg.norm(ref))
    print(f"  n = 4096: rel-L2 {err4096:.2e}")
    assert 1e-14 < err4096 <= BAR
    assert 16 not in changed and 26 not in changed


@pytest.mark.parametrize("cfg", [CONFIGS[0], CONFIGS[5], CONFIGS[-1]])
def test_paired_odd_lengths(cfg):
    lengths = np.array(PAIRS, dtype=np.int64)
    b = ragged.RaggedBatch(lengths)
    x = _inputs(lengths, seed=23)
    xd = torch.from_numpy(x).cuda()
    b.configure(*cfg)
    yd = _run(b, xd, "bf16x", fct=0.5)
    worst = _check_against_oracle(b, x, yd, range(len(lengths)), bar=BAR, fct=0.5)
    print(f"  worst rel-L2 {worst:.2e}")


def test_conjugate_replays_and_switch_back():
    lengths = synth.ragged_lengths(300, hi=70000, seed=4)
    b = ragged.RaggedBatch(lengths, forward=False)
    x = _inputs(lengths, seed=9)
    xd = torch.from_numpy(x).cuda()
    y64 = _run(b, xd, "fp64", fct=0.25)
    b.set_precision("bf16x")
    y1 = torch.zeros(b.out_elems, dtype=torch.complex128, device="cuda")
    b.execute(xd, y1, fct=0.25)                       # captures the matrix-core graph
    torch.cuda.synchronize()
    _check_against_oracle(b, x, y1, range(0, 300, 7), bar=BAR, forward=False, fct=0.25)
    for r in range(3):                               # replays, with eager work in between
        y2 = torch.zeros_like(y1)
        b.execute(xd, y2, fct=0.25)
        z = (xd * 2.0).sum()
        torch.cuda.synchronize()
        assert float(z) == float(z)
        assert torch.equal(y1, y2), r
    assert not torch.equal(y1, y64)
    y3 = _run(b, xd, "fp64", fct=0.25)               # back to FP64: the same bins as before
    assert torch.equal(y3, y64)


def test_rejects_bad_precision():
    b = ragged.RaggedBatch([16, 17])
    with pytest.raises(ValueError):
        b.set_precision("fp16")
    with pytest.raises(F._lib.FFTError):
        F._lib.check(F._lib.lib().impulse_fft_ragged_set_precision(b.handle, 2))


def test_full_batch_20000():
    print("case ragged mfma 20000 signals", flush=True)
    lengths = synth.ragged_lengths(20000, seed=0)
    b = ragged.RaggedBatch(lengths)
    b.set_precision("bf16x")
    total = int(lengths.sum())
    xd = synth.uniform_device(total, -1.0, 1.0, seed=0, dtype=torch.float32)
    yd = torch.empty(b.out_elems, dtype=torch.complex128, device="cuda")
    b.execute(xd, yd)
    torch.cuda.synchronize()
    n = torch.from_numpy(lengths).cuda()
    seg_in = torch.repeat_interleave(torch.arange(len(lengths), device="cuda"), n)
    e_in = torch.zeros(len(lengths), dtype=torch.float64, device="cuda")
    e_in.index_add_(0, seg_in, xd.double() ** 2)
    nb = n // 2 + 1
    seg_out = torch.repeat_interleave(torch.arange(len(lengths), device="cuda"), nb)
    pos = torch.arange(int(nb.sum()), device="cuda") - torch.repeat_interleave(torch.cumsum(nb, 0) - nb, nb)
    nn = torch.repeat_interleave(n, nb)
    w = torch.full_like(pos, 2, dtype=torch.float64)
    w[pos == 0] = 1.0
    w[(nn % 2 == 0) & (pos == nn // 2)] = 1.0
    e_out = torch.zeros(len(lengths), dtype=torch.float64, device="cuda")
    e_out.index_add_(0, seg_out, w * yd.abs() ** 2)
    rel = (e_out / n.double() - e_in).abs() / e_in
    print(f"Parseval over 20000 signals: max rel {float(rel.max()):.2e}")
    assert float(rel.max()) <= BAR
    rng = np.random.default_rng(1)
    order = np.argsort(lengths)
    idx = np.unique(np.r_[order[-8:], rng.choice(len(lengths), 40, replace=False)])
    x_host = xd.cpu().numpy()
    worst = _check_against_oracle(b, x_host, yd, idx, bar=BAR)
    print(f"48 sampled signals vs oracle: worst rel-L2 {worst:.2e}")
