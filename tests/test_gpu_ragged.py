"""Ragged batches (impulse_fft_ragged_*): parity with the oracle's r2c per signal.

* a coverage batch whose lengths reach every branch of the per-signal plan
  (n = 1, 2, 3; LDS-resident; four-step; Bluestein in LDS; Bluestein
  four-step; even and odd n), under every configuration the autotuner may
  pick, against the oracle at 1e-12 relative L2 per signal;
* the graph replay against a direct launch (bit-identical), repeated
  replays with eager work in between, scale and conjugation (forward=False);
* the 20,000-signal batch (lengths log-uniform in [64, 262144], 70 %
  7-smooth, 30 % with a prime factor > 1000, f32 values uniform[-1,1]):
  Parseval on every signal, 48 sampled signals against the oracle, and the
  same sample against the FP64 per-signal plan path.
"""
import numpy as np
import pytest
import torch

import oracle
from impulse_amd import fft as F
from impulse_amd import ragged, synth

pytestmark = pytest.mark.gpu

CONFIGS = [(1024, 8192, 0, 0), (1024, 8192, 1, 0), (1024, 4096, 0, 0), (512, 4096, 0, 0), (512, 4096, 1, 0),
           (512, 4096, 1, 4), (512, 2048, 1, 0), (512, 2048, 1, 4), (256, 2048, 0, 0), (256, 2048, 1, 0),
           (256, 2048, 0, 4), (256, 2048, 1, 4), (256, 2048, 0, 3), (256, 2048, 1, 3), (256, 2048, 3, 4),
           (256, 2048, 2, 4), (256, 2048, 3, 3), (512, 4096, 3, 4)]

# every branch: LDS-resident (smooth / Bluestein, odd / even), split smooth (even with the paired
# output, odd), split Bluestein with short columns (3001, 3022) and long ones, n = 1, 2, 3
COVER = [1, 2, 3, 4, 5, 6, 7, 8, 9, 16, 26, 64, 97, 127, 128, 194, 1000, 1009, 2018, 3001, 3022, 4095, 4096,
         8190, 8192, 15625, 16384, 16386, 17000, 19683, 20000, 26013, 65536, 100003, 131072, 200006, 262143,
         262144]


@pytest.fixture(scope="module", autouse=True)
def _device():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.set_device(0)
    yield
    torch.cuda.synchronize()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _inputs(lengths, seed):
    total = int(np.sum(lengths))
    x = synth.uniform(total, -1.0, 1.0, seed=seed).astype(np.float32)
    return x


def _check_against_oracle(batch, x_host, y_dev, idx, bar=1e-12, forward=True, fct=1.0):
    y = y_dev.cpu().numpy()
    worst = 0.0
    for i in idx:
        n, io, oo = int(batch.lengths[i]), int(batch.in_offsets[i]), int(batch.out_offsets[i])
        ref = oracle.r2c(x_host[io:io + n].astype(np.float64), [0], forward, fct)
        err = _rel(y[oo:oo + n // 2 + 1], ref)
        worst = max(worst, err)
        assert err <= bar, (n, err)
    return worst


@pytest.mark.parametrize("cfg", CONFIGS)
def test_cover_lengths_every_config(cfg):
    print(f"case ragged cover cfg={cfg}", flush=True)
    lengths = np.array(COVER, dtype=np.int64)
    b = ragged.RaggedBatch(lengths)
    x = _inputs(lengths, seed=21)
    xd = torch.from_numpy(x).cuda()
    yd = torch.full((b.out_elems,), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    b.configure(*cfg)
    b.execute(xd, yd)
    torch.cuda.synchronize()
    worst = _check_against_oracle(b, x, yd, range(len(lengths)))
    print(f"  worst rel-L2 {worst:.2e}")


# odd lengths that repeat share one complex transform per pair (z = x + i x2);
# an odd count leaves one signal unpaired
PAIRS = [3, 3, 5, 5, 5, 97, 97, 105, 105, 1009, 1009, 3001, 3001, 15625, 15625, 15625, 19683, 19683, 26013, 26013,
         100003, 100003, 262143, 262143, 64, 100]


@pytest.mark.parametrize("cfg", [CONFIGS[0], CONFIGS[5], CONFIGS[-1]])
def test_paired_odd_lengths(cfg):
    print(f"case ragged pairs cfg={cfg}", flush=True)
    lengths = np.array(PAIRS, dtype=np.int64)
    b = ragged.RaggedBatch(lengths)
    x = _inputs(lengths, seed=23)
    xd = torch.from_numpy(x).cuda()
    yd = torch.full((b.out_elems,), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    b.configure(*cfg)
    b.execute(xd, yd, fct=0.5)
    torch.cuda.synchronize()
    worst = _check_against_oracle(b, x, yd, range(len(lengths)), fct=0.5)
    print(f"  worst rel-L2 {worst:.2e}")


def test_gaps_offsets_scale_and_conjugate():
    # signals at scattered offsets with gaps; outputs in reverse order; fct and forward=False
    rng = np.random.default_rng(5)
    lengths = rng.choice([33, 64, 97, 1024, 4097, 9000, 30011], size=40)
    gaps = rng.integers(0, 7, size=40)
    ins = np.cumsum(np.r_[0, (lengths + gaps)[:-1]])
    nb = lengths // 2 + 1
    outs = np.cumsum(np.r_[0, nb[::-1][:-1]])[::-1]
    b = ragged.RaggedBatch(lengths, ins, outs, forward=False)
    x = synth.uniform(int(ins[-1] + lengths[-1]), -1, 1, seed=3).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    sentinel = complex(7.0, -7.0)
    yd = torch.full((b.out_elems + 3,), sentinel, dtype=torch.complex128, device="cuda")
    b.execute(xd, yd, fct=0.25)
    torch.cuda.synchronize()
    _check_against_oracle(b, x, yd, range(len(lengths)), forward=False, fct=0.25)
    assert bool(torch.all(yd[b.out_elems:] == sentinel))


def test_graph_replays_match_direct():
    lengths = synth.ragged_lengths(300, hi=40000, seed=4)
    b = ragged.RaggedBatch(lengths)
    x = _inputs(lengths, seed=9)
    xd = torch.from_numpy(x).cuda()
    y1 = torch.zeros(b.out_elems, dtype=torch.complex128, device="cuda")
    b.execute(xd, y1)                       # first call captures the graph
    torch.cuda.synchronize()
    _check_against_oracle(b, x, y1, range(0, 300, 7))
    for r in range(4):                      # replays, with eager work in between
        y2 = torch.zeros_like(y1)
        b.execute(xd, y2)
        z = (xd * 2.0).sum()
        torch.cuda.synchronize()
        assert float(z) == float(z)
        assert torch.equal(y1, y2), r


def test_tune_picks_a_valid_config():
    lengths = synth.ragged_lengths(500, hi=70000, seed=6)
    b = ragged.RaggedBatch(lengths)
    x = _inputs(lengths, seed=10)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros(b.out_elems, dtype=torch.complex128, device="cuda")
    ms = b.tune(xd, yd, reps=1)
    info = b.info()
    print(f"tuned: {info} {ms:.3f} ms")
    assert ms > 0 and (info["threads"], info["lds_points"], info["radix_order"], info["waves_per_simd"]) in CONFIGS
    yd.zero_()
    b.execute(xd, yd)
    torch.cuda.synchronize()
    _check_against_oracle(b, x, yd, range(0, 500, 11))


def test_rejects_bad_arguments():
    with pytest.raises(F._lib.FFTError):
        ragged.RaggedBatch([0, 5])
    b = ragged.RaggedBatch([16, 17])
    with pytest.raises(F._lib.FFTError):
        b.configure(1024, 3000, 0)
    with pytest.raises(F._lib.FFTError):
        b.configure(256, 4096, 0)          # more than 8 points per thread
    xd = torch.zeros(64, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        b.execute(xd, torch.zeros(3, dtype=torch.complex128, device="cuda"))


def test_full_batch_20000():
    print("case ragged 20000 signals", flush=True)
    lengths = synth.ragged_lengths(20000, seed=0)
    b = ragged.RaggedBatch(lengths)
    total = int(lengths.sum())
    xd = synth.uniform_device(total, -1.0, 1.0, seed=0, dtype=torch.float32)
    yd = torch.empty(b.out_elems, dtype=torch.complex128, device="cuda")
    b.execute(xd, yd)
    torch.cuda.synchronize()
    # Parseval per signal: sum x^2 = (|X0|^2 + 2 sum_{0<k<n/2} |Xk|^2 + [n even] |X_{n/2}|^2) / n
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
    assert float(rel.max()) <= 1e-12
    # sampled signals against the oracle and the FP64 per-signal plan path
    rng = np.random.default_rng(1)
    order = np.argsort(lengths)
    idx = np.unique(np.r_[order[-8:], rng.choice(len(lengths), 40, replace=False)])
    x_host = {}
    ycpu = yd.cpu().numpy()
    worst = 0.0
    for i in idx:
        nn_, io, oo = int(lengths[i]), int(b.in_offsets[i]), int(b.out_offsets[i])
        xs = xd[io:io + nn_].double()
        ref = oracle.r2c(xs.cpu().numpy(), [0])
        got = ycpu[oo:oo + nn_ // 2 + 1]
        worst = max(worst, _rel(got, ref))
        plan_out = torch.empty(nn_ // 2 + 1, dtype=torch.complex128, device="cuda")
        p = F.get_plan(F._lib.R2C, [nn_], [8], [16], [0], True)
        p.execute(xs.data_ptr(), plan_out.data_ptr(), 1.0, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert _rel(got, plan_out.cpu().numpy()) <= 1e-12, nn_
    print(f"sampled {len(idx)} signals vs oracle: worst rel-L2 {worst:.2e}")
    assert worst <= 1e-12
