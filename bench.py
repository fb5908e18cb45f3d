"""Throughput of the FP64 drop-in FFT on one MI355X.

One step = one pass of the hot path over the plain FP64 GPU drop-in workload
at scale, inputs resident in HBM:
  * 2-D complex128 c2c over axes [0, 1] of 4096 x 4096, re/im iid N(0, 1)
  * 1-D r2c float64, length 2^22, batch 16, uniform[-1, 1]
  * 1-D r2c float64, length 1,000,003 (prime -> Bluestein), batch 4, uniform[-1, 1]
value = algorithmic GFLOP (5 N log2 N per complex transform, 2.5 N log2 N per
real one, N nominal; 2-D N = all points) / time, summed over the three parts.

The "roofline" object describes the dominant kernel: its launches are timed
with HIP events on the execution stream inside the timed region, and its
algorithmic bytes per launch are the bytes each launch must read plus write
(input read once, output written once).  cpu_baseline times the reference
PocketFFT C++ (built from the reference sources into oracle/_ref) on a bounded
sample of the same workload on this host's cores (the job's CPU share, and one
thread).  `c1` times the reference's CPU-runnable size, r2c [256, 4096] and
[256, 4095] and back, on device buffers, end to end from host arrays through
DataDesc / FFTDesc / apply, and through the C path with a plan per call, beside
the reference PocketFFT on the same arrays.  vendor_fft times torch.fft
(hipFFT / rocFFT on ROCm) on the same device inputs as a black-box comparator;
it is reported beside the result and never enters `value`.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "FFT GFLOP/s (5NlogN) + %HBM roofline; split-FP32 vs FP64 at ≤1e-10 rel"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the [256, 4096] / [256, 4095] r2c leg (device, host end to end, C path)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the forms and widened paths (ragged, MFMA, convolution, DCT, upfirdn, fftpack)")
    ap.add_argument("--only-forms", default="",
                    help="comma list of forms to run alone (ragged,mfma,conv,dct,upfirdn,fftpack), no drop-in step")
    ap.add_argument("--form-reps", type=int, default=10)
    ap.add_argument("--no-f32", action="store_true",
                    help="skip the float32-array run of the drop-in step (profiles of the FP64 launches)")
    ap.add_argument("--no-vendor", action="store_true",
                    help="skip timing torch.fft (hipFFT/rocFFT) on the same inputs as a comparator")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (the reference "
                         "reading for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def copy_calibration():
    import torch
    n = 1 << 27   # 2^27 doubles = 1 GiB
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 3}))


def c_strides(shape, itemsize):
    st, acc = [0] * len(shape), itemsize
    for i in range(len(shape) - 1, -1, -1):
        st[i] = acc
        acc *= shape[i]
    return st


class Part:
    def __init__(self, name, kind, shape, axes, in_dtype, out_shape, out_dtype, flops, comp_bytes, make_input):
        self.name, self.kind, self.shape, self.axes = name, kind, list(shape), list(axes)
        self.in_dtype, self.out_shape, self.out_dtype = in_dtype, list(out_shape), out_dtype
        self.flops, self.comp_bytes, self.make_input = flops, comp_bytes, make_input


def parts():
    from impulse_amd import synth
    n2d = 4096
    nl, bl = 1 << 22, 16
    nb, bb = 1000003, 4
    return [
        Part("c2c_4096x4096_axes01", 0, (n2d, n2d), (0, 1), "c128", (n2d, n2d), "c128",
             5.0 * n2d * n2d * math.log2(n2d * n2d), 2 * n2d * n2d * 16,
             lambda: synth.complex_normal((n2d, n2d), seed=0)),
        Part("r2c_2^22_x16", 1, (bl, nl), (1,), "f64", (bl, nl // 2 + 1), "c128",
             bl * 2.5 * nl * math.log2(nl), bl * (nl * 8 + (nl // 2 + 1) * 16),
             lambda: synth.uniform((bl, nl), -1.0, 1.0, seed=0)),
        Part("r2c_1000003_x4", 1, (bb, nb), (1,), "f64", (bb, nb // 2 + 1), "c128",
             bb * 2.5 * nb * math.log2(nb), bb * (nb * 8 + (nb // 2 + 1) * 16),
             lambda: synth.uniform((bb, nb), -1.0, 1.0, seed=0)),
    ]


def measured_traffic(kernel, threads):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by a plain copy's reading), or None."""
    path = os.path.join(HERE, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    vals = [e["hbm_bytes_per_launch"] for e in rec.get("launches", [])
            if e.get("kernel") == kernel and e.get("grid_threads") in threads]
    return int(sum(vals) / len(vals)) if vals and len(vals) == len(threads) else None


def form_traffic(form):
    """HBM bytes of one call of a form (ragged execute, convolution forward / backward) from the
    committed rocprofv3 --pmc summary, scaled like the drop-in's, or None."""
    try:
        with open(os.path.join(HERE, "profiles", "pmc_traffic.json")) as f:
            return json.load(f).get("forms", {}).get(form)
    except (OSError, ValueError):
        return None


def torch_dtype(torch, s):
    return {"c128": torch.complex128, "f64": torch.float64}[s]


def host_cores():
    """CPUs this process may run on, the machine's count, and the job's CPU share (the box gives one
    GPU's job the share in OMP_NUM_THREADS; elsewhere the affinity mask is the limit)."""
    total = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = total
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    use = min(affinity, share) if share > 0 else affinity
    return max(1, use), total, affinity


def _ref_rate(plist, data, threads, budget_s, max_reps=50):
    from oracle import ref
    total_flops, total_s, reps = 0.0, 0.0, 0
    t_start = time.perf_counter()
    while True:
        for p, x in zip(plist, data):
            t0 = time.perf_counter()
            if p.kind == 0:
                ref.c2c(x, p.axes, True, 1.0, threads)
            else:
                ref.r2c(x, p.axes, True, 1.0, threads)
            total_s += time.perf_counter() - t0
            total_flops += p.flops
        reps += 1
        if time.perf_counter() - t_start > budget_s or reps >= max_reps:
            break
    return total_flops / total_s / 1e9, reps


def cpu_baseline(plist, budget_s):
    """Reference PocketFFT C++ (oracle/_ref) on this host's cores, same workload, bounded time: at the
    job's full CPU share (`value`) and at nthreads = 1."""
    from oracle import ref
    if not ref.available():
        return None
    threads, total, affinity = host_cores()
    data = [p.make_input() for p in plist]
    v_all, reps_all = _ref_rate(plist, data, threads, budget_s)
    v_one, reps_one = _ref_rate(plist, data, 1, budget_s / 2, max_reps=3)
    return {"value": v_all, "unit": "GFLOP/s", "cores": threads, "kind": "reference",
            "host_cpus": total, "affinity_cpus": affinity,
            "single_thread": {"value": v_one, "unit": "GFLOP/s", "cores": 1, "reps": reps_one},
            "sample": f"the full 3-part workload repeated {reps_all}x on the host (reference PocketFFT C++ "
                      f"c2c/r2c, nthreads={threads} = the job's CPU share of a {total}-CPU host, "
                      f"built -O3 -march=x86-64-v3 from the reference sources); single_thread: "
                      f"{reps_one}x at nthreads=1"}


def c1_leg(torch, reps, with_cpu, host_reps=10):
    """Plain FP64 drop-in at the reference's CPU-runnable size: r2c float64 [256, 4096] and
    [256, 4095] (4095 = 3^2*5*7*13) along axis 1 (FFTDesc axes=[1]), uniform[-1, 1] seed 0, then the
    inverse c2r with fct 1/N.  Three ways:
      * device: device-resident buffers, HIP events;
      * host: end to end through DataDesc / FFTDesc / apply on host arrays (the drop-in as impulse
        calls it: PCIe copies through the plan's pinned staging included);
      * c_path: impulse's C path per row, a plan built and destroyed per call
        (c_pocketfft/pocketfft.nim:216-288: make_rfft_plan, rfft_forward, destroy_rfft_plan).
    With with_cpu, the reference PocketFFT (oracle/_ref) is timed on the same arrays at nthreads = 1 and
    at the job's CPU share, and its results are the accuracy reference."""
    from impulse_amd import _lib, synth
    from impulse_amd import fft as F

    L = _lib.lib()
    rows = 256
    sp = torch.cuda.current_stream().cuda_stream
    threads, total_cpus, _ = host_cores()
    out = {"workload": "r2c f64 [256, N] axes [1] + c2r back (fct 1/N), N = 4096 and 4095", "cases": []}
    for n in (4096, 4095):
        nb = n // 2 + 1
        x = synth.uniform((rows, n), -1.0, 1.0, seed=0)
        fl = rows * 2.5 * n * math.log2(n)
        comp = rows * (n * 8 + nb * 16)
        xd = torch.from_numpy(x).cuda()
        Xd = torch.empty((rows, nb), dtype=torch.complex128, device="cuda")
        zd = torch.empty_like(xd)
        pf = F.Plan(_lib.R2C, [rows, n], [n * 8, 8], [nb * 16, 16], [1], True)
        pb = F.Plan(_lib.C2R, [rows, n], [nb * 16, 16], [n * 8, 8], [1], False)
        fwd = _timed(torch, lambda: pf.execute(xd.data_ptr(), Xd.data_ptr(), 1.0, sp), reps)
        bwd = _timed(torch, lambda: pb.execute(Xd.data_ptr(), zd.data_ptr(), 1.0 / n, sp), reps)
        X = Xd.cpu().numpy()
        rt = float(np.linalg.norm(zd.cpu().numpy() - x) / np.linalg.norm(x))
        case = {"n": n, "gflop": round(fl / 1e9, 4),
                "device": {"fwd_ms": round(fwd, 4), "bwd_ms": round(bwd, 4),
                           "fwd_gflops": round(fl / (fwd * 1e-3) / 1e9, 1),
                           "fwd_compulsory_GBs": round(comp / (fwd * 1e-3) / 1e9, 1),
                           "fwd_hbm_roofline_frac": round(comp / (fwd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                           "launches": pf.num_launches},
                "round_trip_rel_l2": rt}
        # end to end on host arrays through the drop-in API
        Xh = np.empty((rows, nb), np.complex128)
        zh = np.empty_like(x)
        fdesc, bdesc = F.FFTDesc([1], True), F.FFTDesc([1], False, 1.0 / n)
        dX, dx, dz = F.DataDesc(Xh, Xh.shape), F.DataDesc(x, x.shape), F.DataDesc(zh, zh.shape)
        F.apply(fdesc, dX, dx)
        F.apply(bdesc, dz, dX)
        tf, tb = [], []
        for _ in range(host_reps):
            t0 = time.perf_counter()
            F.apply(fdesc, dX, dx)
            t1 = time.perf_counter()
            F.apply(bdesc, dz, dX)
            tb.append(time.perf_counter() - t1)
            tf.append(t1 - t0)
        hf, hb = 1e3 * float(np.median(tf)), 1e3 * float(np.median(tb))
        case["host"] = {"fwd_ms": round(hf, 4), "bwd_ms": round(hb, 4), "fwd_gflops": round(fl / (hf * 1e-3) / 1e9, 2),
                        "note": "PCIe-inclusive: host arrays staged through pinned buffers; not `value`"}
        # impulse's C path, one plan per call per row
        buf = np.empty(n)
        t0 = time.perf_counter()
        for r in range(rows):
            buf[:] = x[r]
            h = L.impulse_make_rfft_plan(n)
            if L.impulse_rfft_forward(h, buf.ctypes.data, 1.0) != 0:
                raise RuntimeError("Forward FFT calculation failed.")
            L.impulse_destroy_rfft_plan(h)
        cp = (time.perf_counter() - t0) * 1e3
        case["c_path"] = {"ms": round(cp, 3), "per_call_us": round(cp * 1e3 / rows, 1),
                          "gflops": round(fl / (cp * 1e-3) / 1e9, 3)}
        if with_cpu:
            case["cpu_reference"] = cpu_reference_c1(x, X, n, rows, threads, total_cpus, host_reps)
        out["cases"].append(case)
        del pf, pb, xd, Xd, zd
    return out


def cpu_reference_c1(x, X_gpu, n, rows, threads, total_cpus, reps):
    """CPU baseline leg of the C1 workload: the reference PocketFFT C++ r2c / c2r with the byte strides
    DataDesc.init makes (cpp_pocketfft/pocketfft.nim:178-199) at nthreads = 1 and at the job's CPU share,
    and its C path as impulse calls it (plan per call) and with the plan reused.  Also the GPU result's
    rel-L2 against the reference's."""
    from oracle import ref
    if not ref.available():
        return None
    fl = rows * 2.5 * n * math.log2(n)
    res = {"host_cpus": total_cpus}
    want = ref.r2c(x, [1], True)
    res["gpu_vs_reference_rel_l2"] = float(np.linalg.norm(X_gpu - want) / np.linalg.norm(want))
    for th in (1, threads):
        tf, tb = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            Y = ref.r2c(x, [1], True, 1.0, th)
            t1 = time.perf_counter()
            ref.c2r(Y, n, [1], False, 1.0 / n, th)
            tb.append(time.perf_counter() - t1)
            tf.append(t1 - t0)
        f = 1e3 * float(np.median(tf))
        res[f"cpp_nthreads_{th}"] = {"fwd_ms": round(f, 4), "bwd_ms": round(1e3 * float(np.median(tb)), 4),
                                     "fwd_gflops": round(fl / (f * 1e-3) / 1e9, 2)}
    L = ref.lib()
    buf = np.empty(n)
    t0 = time.perf_counter()
    for r in range(rows):
        buf[:] = x[r]
        h = L.make_rfft_plan(n)
        L.rfft_forward(h, buf.ctypes.data, 1.0)
        L.destroy_rfft_plan(h)
    per_call = (time.perf_counter() - t0) * 1e3
    h = L.make_rfft_plan(n)
    t0 = time.perf_counter()
    for r in range(rows):
        buf[:] = x[r]
        L.rfft_forward(h, buf.ctypes.data, 1.0)
    reused = (time.perf_counter() - t0) * 1e3
    L.destroy_rfft_plan(h)
    res["c_path_plan_per_call"] = {"ms": round(per_call, 3), "gflops": round(fl / (per_call * 1e-3) / 1e9, 2)}
    res["c_path_plan_reused"] = {"ms": round(reused, 3), "gflops": round(fl / (reused * 1e-3) / 1e9, 2)}
    return res


def ragged_form(torch, reps, loop_sample=600):
    """Ragged batch of 20,000 real float32 signals (lengths log-uniform in
    [64, 262144], 70 % 2^a3^b5^c7^d, 30 % with a prime factor > 1000, values
    uniform[-1, 1]) in one hipGraph launch, autotuned, against a loop of FP64
    per-signal r2c plans timed on a random sample of the same signals; then the
    same batch with the radix-16 stages on the matrix cores (exact bf16 slices)."""
    from impulse_amd import _lib, ragged, synth
    from impulse_amd import fft as F

    lengths = synth.ragged_lengths(20000, seed=0)
    b = ragged.RaggedBatch(lengths)
    total = int(lengths.sum())
    x = synth.uniform_device(total, -1.0, 1.0, seed=0, dtype=torch.float32)
    y = torch.empty(b.out_elems, dtype=torch.complex128, device="cuda")
    tune_ms = b.tune(x, y, reps=2)
    info = b.info()
    for _ in range(2):
        b.execute(x, y)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.execute(x, y)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    flops = info["flops"]
    comp = 4 * total + 16 * int((lengths // 2 + 1).sum())

    # per-signal FP64 plan loop on a sample (plans and float64 copies made before timing)
    rng = np.random.default_rng(2)
    idx = rng.choice(len(lengths), loop_sample, replace=False)
    sp = torch.cuda.current_stream().cuda_stream
    jobs = []
    for i in idx:
        n, io = int(lengths[i]), int(b.in_offsets[i])
        xs = x[io:io + n].double().contiguous()
        ys = torch.empty(n // 2 + 1, dtype=torch.complex128, device="cuda")
        jobs.append((F.get_plan(_lib.R2C, [n], [8], [16], [0], True), xs, ys, i))
    for pl, xs, ys, _ in jobs:
        pl.execute(xs.data_ptr(), ys.data_ptr(), 1.0, sp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for pl, xs, ys, _ in jobs:
        pl.execute(xs.data_ptr(), ys.data_ptr(), 1.0, sp)
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    sample_flops = float(sum(2.5 * int(lengths[i]) * math.log2(int(lengths[i])) for i in idx))
    loop_gflops = sample_flops / loop_s / 1e9
    worst = 0.0
    for pl, xs, ys, i in jobs:
        o, n = int(b.out_offsets[i]), int(lengths[i])
        d = torch.linalg.vector_norm(y[o:o + n // 2 + 1] - ys) / torch.linalg.vector_norm(ys)
        worst = max(worst, float(d))
    gf = flops / (ms * 1e-3) / 1e9

    # the matrix-core precision (radix-16 stages as exact bf16-slice MFMA products), tuned on its
    # own; its bins against the same FP64 per-signal plans
    b.set_precision("bf16x")
    tune_mf = b.tune(x, y, reps=2)
    info_mf = b.info()
    for _ in range(2):
        b.execute(x, y)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.execute(x, y)
    e1.record()
    torch.cuda.synchronize()
    ms_mf = e0.elapsed_time(e1) / reps
    worst_mf = 0.0
    for pl, xs, ys, i in jobs:
        o, n = int(b.out_offsets[i]), int(lengths[i])
        d = torch.linalg.vector_norm(y[o:o + n // 2 + 1] - ys) / torch.linalg.vector_norm(ys)
        worst_mf = max(worst_mf, float(d))
    b.set_precision("fp64")
    mf = {"ms": round(ms_mf, 3), "gflops": round(flops / (ms_mf * 1e-3) / 1e9, 1),
          "vs_fp64_ragged": round(ms / ms_mf, 3), "max_rel_l2_vs_fp64_plan": worst_mf,
          "tuned": {k: info_mf[k] for k in ("threads", "lds_points", "radix_order", "waves_per_simd")},
          "tune_best_ms": round(tune_mf, 3)}
    return {
        "form": "ragged", "signals": len(lengths), "samples": total, "gflop": round(flops / 1e9, 3),
        "ms": round(ms, 3), "gflops": round(gf, 1), "compulsory_GBs": round(comp / (ms * 1e-3) / 1e9, 1),
        "hbm_roofline_frac": round(comp / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "launches": 1,
        "dtype": "f32 in, f64 arithmetic, c128 out",
        "tuned": {k: info[k] for k in ("threads", "lds_points", "radix_order", "waves_per_simd", "grid")},
        "tune_best_ms": round(tune_ms, 3), "scratch_bytes": info["scratch_bytes"],
        "fp64_plan_loop": {"sample": loop_sample, "gflops": round(loop_gflops, 2),
                           "projected_ms": round(flops / (loop_gflops * 1e9) * 1e3, 1),
                           "speedup": round(gf / loop_gflops, 1)},
        "max_rel_l2_vs_fp64_plan": worst,
        **ragged_traffic(comp, ms),
        "mfma_bf16x": mf,
    }


def ragged_traffic(comp, ms):
    """The profiled execute's HBM traffic against the compulsory bytes (counters committed in
    profiles/pmc_traffic.json), or nothing if no summary is there."""
    tr = form_traffic("ragged")
    if not tr or not tr.get("hbm_bytes_per_execute"):
        return {}
    t = tr["hbm_bytes_per_execute"]
    return {"traffic": {"hbm_bytes": t, "compulsory_bytes": comp, "over_compulsory": round(t / comp, 2),
                        "hbm_GBs": round(t / (ms * 1e-3) / 1e9, 1), "profiled_config": tr.get("config")}}


def mfma_form(torch, reps):
    """Matrix-core form: 8192 complex64 transforms of 65536 (re/im N(0,1)),
    radix-16 stages on MFMA in two precisions, against the FP64 drop-in c2c
    plan on the same values (complex128 in and out).  Errors are relative L2
    per transform against the FP64 plan over 64 sampled transforms."""
    from impulse_amd import _lib, synth
    from impulse_amd import fft as F
    from impulse_amd.mfma import MfmaFFT

    B, L = 8192, 65536
    flops = 5.0 * L * math.log2(L) * B
    x = synth.complex_normal_device(B * L, seed=0, dtype=torch.complex64)
    rows = np.random.default_rng(4).choice(B, 64, replace=False)
    sp = torch.cuda.current_stream().cuda_stream

    def timed(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    # FP64 drop-in on the same values
    x64 = x.to(torch.complex128)
    ref = torch.empty_like(x64)
    p = F.Plan(_lib.C2C, [B, L], [L * 16, 16], [L * 16, 16], [1], True)
    ms64 = timed(lambda: p.execute(x64.data_ptr(), ref.data_ptr(), 1.0, sp))
    refs = ref.view(B, L)[rows].clone()
    del x64, ref, p
    out = {"form": "mfma", "workload": "c2c complex64 65536 x 8192, re/im N(0,1)", "gflop": round(flops / 1e9, 2),
           "fp64_dropin": {"ms": round(ms64, 3), "gflops": round(flops / (ms64 * 1e-3) / 1e9, 1),
                           "io": "complex128 in and out"}}
    for prec, io_bytes in (("f32", 8 + 8), ("bf16x", 8 + 16)):
        plan = MfmaFFT(L, B, prec)
        y = torch.empty(B * L, dtype=plan.out_dtype, device="cuda")
        ms = timed(lambda: plan(x, y))
        err = (torch.linalg.vector_norm(y.view(B, L)[rows].to(torch.complex128) - refs, dim=1) /
               torch.linalg.vector_norm(refs, dim=1)).cpu().numpy()
        comp = io_bytes * B * L
        out[prec] = {"ms": round(ms, 3), "gflops": round(flops / (ms * 1e-3) / 1e9, 1),
                     "compulsory_GBs": round(comp / (ms * 1e-3) / 1e9, 1),
                     "hbm_roofline_frac": round(comp / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "launches": 2,
                     "rel_l2_vs_fp64": {"max": float(err.max()), "median": float(np.median(err))},
                     "speedup_vs_fp64_dropin": round(ms64 / ms, 2)}
        del plan, y
        torch.cuda.empty_cache()
    # torch.fft (hipFFT / rocFFT) on the same complex64 values, as a black-box comparator
    xv = x.view(B, L)
    msv = timed(lambda: torch.fft.fft(xv, dim=1))
    out["vendor_fft_c64"] = {"library": "torch.fft.fft (hipFFT/rocFFT), black-box comparator", "ms": round(msv, 3),
                             "f32_form_speedup": round(msv / out["f32"]["ms"], 2)}
    return out


def conv_form(torch, reps):
    """Differentiation form: 'same' FFT convolution of 16 float32 images 2048x2048
    (U[0,1]) with a 63x63 kernel (N(0, 1/63^2)), forward and the adjoint
    backward (image and kernel gradients) of loss = sum((y - t)^2)."""
    from impulse_amd import autograd as AG
    from impulse_amd import synth

    B, H, K = 16, 2048, 63
    x = synth.uniform_device(B * H * H, 0.0, 1.0, seed=11, dtype=torch.float32).view(B, H, H)
    k = torch.from_numpy(synth.normal((K, K), seed=12, sigma=1.0 / K).astype(np.float32)).cuda()
    t = synth.uniform_device(B * H * H, -1.0, 1.0, seed=13, dtype=torch.float32).view(B, H, H)
    c = AG._conv(B, H, H, K, K, torch.float32)
    P, Q = c.grid
    y = torch.empty_like(x)
    gx, gk = torch.empty_like(x), torch.empty_like(k)

    def timed(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    fwd = timed(lambda: c.forward(x, k, y))
    gy = 2.0 * (y - t)
    bwd = timed(lambda: c.backward(x, k, gy, gx, gk))
    # transforms per call: forward r2c of the images + kernel, c2r of the product;
    # backward (recomputing spectra) r2c-adjoint, c2r-adjoints for both gradients
    n = P * Q
    fl_r = 2.5 * n * math.log2(n)
    # compulsory bytes: forward reads x and k, writes y; backward reads x, k and dy, writes dx and dk
    img, ker = B * H * H * 4, K * K * 4
    comp_f, comp_b = 2 * img + ker, 3 * img + 2 * ker
    out = {"form": "conv2d", "workload": "16 x 2048^2 f32 images * 63x63 kernel, 'same', FP64 arithmetic",
           "grid": [P, Q], "forward_ms": round(fwd, 3), "backward_ms": round(bwd, 3),
           "forward_gflops": round((B + 1 + B) * fl_r / (fwd * 1e-3) / 1e9, 1),
           "forward_compulsory_GBs": round(comp_f / (fwd * 1e-3) / 1e9, 1),
           "forward_hbm_roofline_frac": round(comp_f / (fwd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "backward_compulsory_GBs": round(comp_b / (bwd * 1e-3) / 1e9, 1),
           "backward_hbm_roofline_frac": round(comp_b / (bwd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "note": "backward recomputes the image and kernel spectra instead of storing them"}
    tr = form_traffic("conv2d")
    if tr:
        # counters of the profiled build (profiles/pmc_traffic.json): HBM bytes per call and the
        # rate they imply at this run's time
        for kind, ms, comp in (("forward", fwd, comp_f), ("backward", bwd, comp_b)):
            t = tr.get(kind, {}).get("hbm_bytes_per_call")
            if t:
                out[f"{kind}_traffic"] = {"hbm_bytes": t, "over_compulsory": round(t / comp, 2),
                                          "hbm_GBs": round(t / (ms * 1e-3) / 1e9, 1)}
    return out


def _timed(torch, fn, reps):
    for _ in range(2):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def dct_form(torch, reps):
    """DCT-II (pocketfft::dct type 2) of a float64 4096 x 4096 array over axes [0, 1] and over its rows,
    uniform[-1, 1].  Compulsory bytes: the array read once and written once per call.  Accuracy:
    DCT-III of the result divided by (2N)^d returns the input."""
    from impulse_amd import _lib, synth
    from impulse_amd import fft as F

    n = 4096
    x = synth.uniform_device(n * n, -1.0, 1.0, seed=21, dtype=torch.float64).view(n, n)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    sp = torch.cuda.current_stream().cuda_stream
    st = [n * 8, 8]
    out = {"form": "dct", "workload": "DCT-II f64 4096 x 4096", "cases": []}
    for axes in ([0, 1], [1]):
        p2 = F.Plan(_lib.DCT, [n, n], st, st, axes, True, r2r_type=2)
        p3 = F.Plan(_lib.DCT, [n, n], st, st, axes, True, r2r_type=3)
        ms = _timed(torch, lambda: p2.execute(x.data_ptr(), y.data_ptr(), 1.0, sp), reps)
        p3.execute(y.data_ptr(), z.data_ptr(), 1.0 / (2.0 * n) ** len(axes), sp)
        torch.cuda.synchronize()
        err = float(torch.linalg.vector_norm(z - x) / torch.linalg.vector_norm(x))
        comp = 2 * n * n * 8
        out["cases"].append({"axes": axes, "ms": round(ms, 3), "launches": p2.num_launches,
                             "compulsory_GBs": round(comp / (ms * 1e-3) / 1e9, 1),
                             "hbm_roofline_frac": round(comp / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                             "round_trip_rel_l2": err})
        del p2, p3
    return out


def fftpack_form(torch, reps):
    """r2r_fftpack (forward, halfcomplex out) over the rows of a float64 4096 x 4096 array and the genuine
    2-D Hartley transform of it (pocketfft_hdronly.h:3392-3445), uniform[-1, 1].  Compulsory bytes: the
    array read once and written once per call."""
    from impulse_amd import _lib, synth
    from impulse_amd import fft as F

    n = 4096
    x = synth.uniform_device(n * n, -1.0, 1.0, seed=41, dtype=torch.float64).view(n, n)
    y = torch.empty_like(x)
    sp = torch.cuda.current_stream().cuda_stream
    st = [n * 8, 8]
    out = {"form": "fftpack", "workload": "f64 4096 x 4096", "cases": []}
    for name, kind, axes in (("r2r_fftpack rows", _lib.FFTPACK, [1]),
                             ("genuine Hartley axes [0, 1]", _lib.HARTLEY_GENUINE, [0, 1])):
        p = F.FftpackPlan(kind, [n, n], st, st, axes, True, True)
        ms = _timed(torch, lambda: p.execute(x.data_ptr(), y.data_ptr(), 1.0, sp), reps)
        comp = 2 * n * n * 8
        out["cases"].append({"case": name, "ms": round(ms, 3), "launches": p.num_launches,
                             "compulsory_GBs": round(comp / (ms * 1e-3) / 1e9, 1),
                             "hbm_roofline_frac": round(comp / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)})
        del p
    return out


def upfirdn_form(torch, reps):
    """upfirdn / resample (impulse/signal.nim:545-809) on float64 signals, uniform[-1, 1]:
    resample by 3/2 of 2^24 samples with the generated 61-tap filter (direct algorithm), and a
    4001-tap FIR over 2^22 samples (overlap-save FFT algorithm, against the direct one).
    Compulsory bytes: the signal read once, the output written once."""
    from impulse_amd import _lib, synth
    from impulse_amd import signal as S

    out = {"form": "upfirdn", "cases": []}
    rng = np.random.default_rng(31)
    cases = [("resample 3/2, 2^24 samples, 61-tap generated filter", 1 << 24, None, 3, 2),
             ("FIR 4001 taps, 2^22 samples", 1 << 22, rng.standard_normal(4001) / 64.0, 1, 1)]
    for name, nt, h, up, down in cases:
        if h is None:
            h0 = S.generate_resampling_filter(up, down)
            h, _ = S.adjust_filter_position(h0, nt, int(math.ceil(nt * up / down)), up, down)
        t = synth.uniform_device(nt, -1.0, 1.0, seed=32, dtype=torch.float64)
        res = {"case": name, "taps": int(h.size), "up": up, "down": down}
        ys = {}
        for algo, label in ((_lib.UPFIRDN_DIRECT, "direct"), (_lib.UPFIRDN_FFT, "fft")):
            p = S.UpfirdnPlan(nt, h, up, down, algorithm=algo)
            y = torch.empty((1, p.nout), dtype=torch.float64, device="cuda")
            ms = _timed(torch, lambda: p.execute(t, y), reps)
            comp = (nt + p.nout) * 8
            res[label] = {"ms": round(ms, 3), "compulsory_GBs": round(comp / (ms * 1e-3) / 1e9, 1),
                          "hbm_roofline_frac": round(comp / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "fft_length": p.fft_length, "blocks": p.blocks}
            ys[label] = y
            del p
        res["auto"] = "fft" if S.UpfirdnPlan(nt, h, up, down).algorithm == _lib.UPFIRDN_FFT else "direct"
        res["direct_vs_fft_rel_l2"] = float(torch.linalg.vector_norm(ys["fft"] - ys["direct"]) /
                                            torch.linalg.vector_norm(ys["direct"]))
        out["cases"].append(res)
    return out


def vendor_fft(torch, plist, bufs, ms_per_step, reps):
    """torch.fft (hipFFT / rocFFT on this ROCm build) on the same device inputs
    and output shapes, timed with events: a black-box comparator, never part of
    `value`."""
    from impulse_amd import _lib
    rows, total = [], 0.0
    for p, (x, _) in zip(plist, bufs):
        if p.kind == _lib.C2C:
            fn = lambda x=x, p=p: torch.fft.fftn(x, dim=tuple(p.axes))
        else:
            fn = lambda x=x, p=p: torch.fft.rfft(x, dim=p.axes[0])
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        total += ms
        rows.append({"part": p.name, "ms": round(ms, 4)})
    return {"library": "torch.fft (hipFFT/rocFFT), black-box comparator on the same device inputs",
            "parts": rows, "step_ms": round(total, 4), "this_engine_speedup": round(total / ms_per_step, 2)}


def f32_arrays(torch, plist, bufs, ms64, steps, warmup):
    """The same three-part workload with float32 / complex64 arrays (the reference's
    FFTDesc[float32], pocketfft::c2c/r2c<float>): FP64 arithmetic inside, the arrays half the
    bytes.  Inputs are the FP64 inputs rounded to float; the error is the rel-L2 of each part's
    output against its FP64 output."""
    from impulse_amd import _lib
    from impulse_amd import fft as F
    f32of = {"c128": torch.complex64, "f64": torch.float32}
    b32, pl32 = [], []
    for p, (x, y) in zip(plist, bufs):
        x32 = x.to(f32of[p.in_dtype])
        y32 = torch.empty(p.out_shape, dtype=f32of[p.out_dtype], device="cuda")
        ib = 8 if p.in_dtype == "c128" else 4
        pl32.append(F.Plan(p.kind, p.shape, c_strides(p.shape, ib), c_strides(p.out_shape, 8), p.axes, True,
                           dtype=_lib.F32))
        b32.append((x32, y32))
    sp = torch.cuda.current_stream().cuda_stream

    def step():
        for pl, (x, y) in zip(pl32, b32):
            pl.execute(x.data_ptr(), y.data_ptr(), 1.0, sp)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    errs = []
    for (x, y), (_, y32) in zip(bufs, b32):
        errs.append(float(torch.linalg.vector_norm(y32.to(y.dtype) - y) / torch.linalg.vector_norm(y)))
    return {"workload": "the same three parts with float32 / complex64 arrays, FP64 arithmetic",
            "ms_per_step": round(ms, 4), "gflops": round(sum(p.flops for p in plist) / (ms * 1e-3) / 1e9, 1),
            "vs_fp64_arrays": round(ms64 / ms, 3), "max_rel_l2_vs_fp64": max(errs)}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the engine has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import torch

    from impulse_amd import _lib
    from impulse_amd import fft as F

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    L = _lib.lib()
    _lib.check(L.impulse_fft_device_check())
    form_fns = {"ragged": ragged_form, "mfma": mfma_form, "conv": conv_form, "dct": dct_form,
                "upfirdn": upfirdn_form, "fftpack": fftpack_form}
    if args.only_forms:
        print(json.dumps({"forms": [form_fns[f](torch, args.form_reps) for f in args.only_forms.split(",")]}))
        return
    plist = parts()

    # device-resident inputs and outputs, plans
    bufs, plans = [], []
    for p in plist:
        x = torch.from_numpy(p.make_input()).cuda()
        y = torch.empty(p.out_shape, dtype=torch_dtype(torch, p.out_dtype), device="cuda")
        ib = 16 if p.in_dtype == "c128" else 8
        pl = F.Plan(p.kind, p.shape, c_strides(p.shape, ib), c_strides(p.out_shape, 16), p.axes, True)
        bufs.append((x, y))
        plans.append(pl)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def step():
        for pl, (x, y) in zip(plans, bufs):
            pl.execute(x.data_ptr(), y.data_ptr(), 1.0, sp)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # timed region: every launch is bracketed by HIP events on `stream` by the library
    part_ms = [0.0] * len(plist)
    for pl in plans:
        _lib.check(L.impulse_fft_plan_profile(pl.handle, args.steps))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    launches = []
    for pi_, (p, pl) in enumerate(zip(plist, plans)):
        n = pl.num_launches
        ms = (ctypes.c_double * n)()
        cnt = ctypes.c_int()
        _lib.check(L.impulse_fft_plan_profile_read(pl.handle, ms, n, ctypes.byref(cnt)))
        for k in range(n):
            name = ctypes.create_string_buffer(64)
            bi, bo, gr, bl_, lds = (ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_uint(), ctypes.c_uint(),
                                    ctypes.c_size_t())
            _lib.check(L.impulse_fft_plan_pass_info(pl.handle, k, name, 64, ctypes.byref(bi), ctypes.byref(bo),
                                                    ctypes.byref(gr), ctypes.byref(bl_), ctypes.byref(lds)))
            avg = ms[k] / max(1, cnt.value)
            launches.append({"part": p.name, "pass": k, "kernel": name.value.decode(), "ms": avg,
                             "bytes": bi.value + bo.value, "grid": gr.value, "block": bl_.value,
                             "lds": lds.value})
            part_ms[pi_] += avg
        _lib.check(L.impulse_fft_plan_profile(pl.handle, 0))

    ms_per_step = elapsed * 1e3 / args.steps
    total_flops = sum(p.flops for p in plist)
    value = total_flops / (ms_per_step * 1e-3) / 1e9

    # dominant kernel: the kernel instantiation (as rocprof names it) with the
    # largest summed launch time per step
    by_kernel = {}
    for l_ in launches:
        k = by_kernel.setdefault(l_["kernel"], {"ms": 0.0, "bytes": 0, "n": 0, "threads": set()})
        k["ms"] += l_["ms"]
        k["bytes"] += l_["bytes"]
        k["n"] += 1
        k["threads"].add(l_["grid"] * l_["block"])
    dom = max(by_kernel, key=lambda k: by_kernel[k]["ms"])
    d = by_kernel[dom]
    avg_ms = d["ms"] / d["n"]
    avg_bytes = d["bytes"] / d["n"]
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, d["threads"]),
                "kernel": dom, "launches_per_step": d["n"], "avg_launch_ms": round(avg_ms, 4),
                "alg_bytes_per_launch": int(avg_bytes)}

    part_rows = []
    for p, pms in zip(plist, part_ms):
        npass = sum(1 for l_ in launches if l_["part"] == p.name)
        part_rows.append({"part": p.name, "ms": round(pms, 4), "gflops": round(p.flops / (pms * 1e-3) / 1e9, 1),
                          "compulsory_GBs": round(p.comp_bytes / (pms * 1e-3) / 1e9, 1),
                          "hbm_roofline_frac": round(p.comp_bytes / (pms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "launches": npass})

    c1 = None if args.no_c1 else c1_leg(torch, max(args.steps, 20), not args.no_cpu)
    forms = [] if args.no_forms else [ragged_form(torch, args.form_reps), mfma_form(torch, args.form_reps),
                                       conv_form(torch, args.form_reps), dct_form(torch, args.form_reps),
                                       upfirdn_form(torch, args.form_reps), fftpack_form(torch, args.form_reps)]
    for f in forms:
        if f.get("form") == "ragged":
            # ragged-batch throughput (FP64 and matrix-core stages) against the FP64 drop-in's
            f["vs_fp64_dropin_gflops"] = round(f["gflops"] / value, 3)
            f["mfma_bf16x"]["vs_fp64_dropin_gflops"] = round(f["mfma_bf16x"]["gflops"] / value, 3)
    vendor = None if args.no_vendor else vendor_fft(torch, plist, bufs, ms_per_step, args.steps)
    f32 = None if args.no_f32 else f32_arrays(torch, plist, bufs, ms_per_step, args.steps, args.warmup)
    cpu = None if args.no_cpu else cpu_baseline(plist, args.cpu_seconds)
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (SplitMix64 seed 0)",
        "config": {"workload": "FP64 drop-in at scale: c2c c128 4096x4096 axes [0,1] + r2c f64 2^22 x16 + "
                               "r2c f64 1000003 x4", "parts": [p.name for p in plist]},
        "roofline": roofline, "cpu_baseline": cpu, "parts": part_rows, "launches": launches, "c1": c1,
        "forms": forms,
        "vendor_fft": vendor, "f32_arrays": f32,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
