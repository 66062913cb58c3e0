#!/usr/bin/env python3
"""Benchmark: batched complex64 forward FFT, N = 2^20, batch 128, on one MI355X.

One step = one forward transform of the whole batch (1 GiB in, 1 GiB out),
device-resident, through the C ABI (gfft_execute) on torch's current stream.
value = GFLOP/s in the 5*N*log2(N) convention.

The JSON line also carries
  roofline      -- the dominant kernel: algorithmic bytes per launch (one read
                   and one write of the elements it transforms) over its mean
                   duration, measured live with one HIP event pair on the
                   execution stream around 20 back-to-back runs of each launch
                   (gfft_plan_profile_launch);
                   `traffic` is the HBM bytes per launch measured by rocprofv3
                   PMC counters, read from profiles/pmc_traffic.json when present;
  e2e           -- the whole transform against the HBM roofline;
  copy_gbs      -- (in roofline) a plain 1 GiB device copy measured in the same run,
                   the practical ceiling of the box, with frac_of_copy beside frac;
  pattern_copy  -- (in roofline, complex64 four-step plans) copies with no
                   arithmetic in the HBM access pattern of each tile pass and the
                   same pipelined loop (genfft_amd/libgfft_pattern_copy.so), in
                   both tile orders: the ceiling a pass with that pattern can
                   reach, and the dominant kernel's fraction of the faster one;
  cpu_baseline  -- genFFT itself (oracle/_ref, built from the reference sources)
                   timed on the host cores over a bounded sample, or the
                   oracle's C restatement when that library is absent;
  fused_filter  -- y = irfft(H * rfft(x)) on x [256, 16384] f32, H [8193]:
                   fused forward + adjoint kernels against the same operator
                   through unfused launches (rfft / irfft plans + elementwise);
  uneven_batch  -- 20,000 complex64 transforms of sizes 2^6..2^16, mixed
                   directions: the batch plan (one persistent launch for the
                   short lengths, per-class launches above) replayed as a HIP
                   graph and run eagerly, against one plan launch per transform;
  precision     -- N = 2^20 batch 64 and one 8192 x 8192 image with log-uniform
                   magnitudes 1e-3..1e3: FP16 (+ per-transform scale), BF16 and
                   FP32 modes, time and rel-L2 against the GPU complex128 plan
                   and numpy's float64 FFT;
  reference_cases -- genFFT's own benchmark cases (test/fft_bench.cpp): FFT_1D
                   and RealFFT_1D latency at batch 1, n = 2..2^21, and FFT_Vert
                   nfft x ncols, on device and host pointers.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-extras]
                      [--leg fused_filter|uneven_batch|precision|reference_cases]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    """Progress on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)
WORKLOADS = {
    # name: (precision, n, batch)
    "c64_n2^20_b128": ("c64", 1 << 20, 128),
    "c64_n2^10_b131072": ("c64", 1 << 10, 131072),
    "c64_n2^14_b8192": ("c64", 1 << 14, 8192),
    "c64_n2^18_b512": ("c64", 1 << 18, 512),
    "c64_n2^22_b32": ("c64", 1 << 22, 32),
    "c128_n2^10_b65536": ("c128", 1 << 10, 65536),
    "c128_n2^14_b4096": ("c128", 1 << 14, 4096),
    "c128_n2^18_b256": ("c128", 1 << 18, 256),
    "c128_n2^20_b64": ("c128", 1 << 20, 64),
    "c128_n2^22_b16": ("c128", 1 << 22, 16),
    # 16-bit storage (FP32 arithmetic), 4 bytes per complex element
    "f16_n2^20_b64": ("f16", 1 << 20, 64),
    "bf16_n2^20_b64": ("bf16", 1 << 20, 64),
    "f16_n2^20_b256": ("f16", 1 << 20, 256),
}
CSIZE = {"c64": 8, "c128": 16, "f16": 4, "bf16": 4}
# real-input and 2D workloads of the same configuration set: name -> (kind, n, n2, batch), complex64 / float32
OTHER = {
    "r2c_f32_n2^20_b256": ("r2c", 1 << 20, 0, 256),
    "c2d_c64_4096x4096": ("c2c_2d", 4096, 4096, 1),
    "c2d_c64_8192x8192": ("c2c_2d", 8192, 8192, 1),
    "c2r_f32_n2^20_b256": ("c2r", 1 << 20, 0, 256),
    "r2c2d_f32_4096x4096": ("r2c_2d", 4096, 4096, 1),
    "c2r2d_f32_4096x4096": ("c2r_2d", 4096, 4096, 1),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="c64_n2^20_b128", choices=sorted(WORKLOADS) + sorted(OTHER))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the fused-filter, uneven-batch, precision and reference-case legs")
    ap.add_argument("--leg", choices=["fused_filter", "uneven_batch", "precision", "reference_cases", "tile_widths"],
                    help="run only this leg and print its JSON object (for profiling)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain 1 GiB device copy (scale reference for PMC byte counters)")
    return ap.parse_args()


def tile_launches(text):
    """The launch lines of a plan description, each with its launch index
    (every line after the `plan` header is one launch, in execution order)."""
    out = []
    idx = -1
    for ln in text.splitlines():
        if not ln.strip() or ln.startswith("plan "):
            continue
        idx += 1
        if ln.startswith("dual"):
            # two tile passes in one launch: "dual ... | <role A fields> | <role B fields>"
            parts = [p.strip() for p in ln.split("|")]
            f = dict(kv.split("=", 1) for kv in parts[0].split()[1:] if "=" in kv)
            f["kind"] = "dual"
            f["roles"] = [dict(kv.split("=", 1) for kv in p.split() if "=" in kv) for p in parts[1:]]
            f["index"] = idx
            out.append(f)
            continue
        if not ln.startswith(("tile", "r2c_rows_unpack", "c2r_pack_cols", "real_rows", "unpack", "pack", "hfill",
                              "copy", "repack")):
            continue
        f = dict(kv.split("=", 1) for kv in ln.split()[1:] if "=" in kv)
        f["kind"] = ln.split()[0]
        f["index"] = idx
        out.append(f)
    return out


def _tile_key(f):
    return f"tile L={f['L']} M={f['M']} {'row' if f['inp'] == '0' else 'col'}->{'row' if f['outp'] == '0' else 'col'}"


def _tile_elems(f):
    return int(f["n_inner"]) * int(f["tiles"]) // int(f["tiles_inner"]) * int(f["L"])


def host_cpus():
    """(usable, host): the CPUs this process may run on -- its affinity mask,
    capped by a cgroup CPU quota when one is set -- and the host's count."""
    host = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = host
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            usable = min(usable, max(1, math.ceil(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    return max(1, usable), host


def cpu_baseline(x_cpu, n, batch, prec, gpu_out_cpu):
    """genFFT (or the oracle's restatement) on every CPU this process may use,
    bounded sample."""
    import ctypes

    import numpy as np
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libgenfft_ref.so")
    orc_so = os.path.join(ROOT, "oracle", "liboracle.so")
    cores, host = host_cpus()
    sample = min(batch, 128)
    xs = np.ascontiguousarray(x_cpu[:sample])
    ys = np.empty_like(xs)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    dbl = int(prec == "c128")
    if os.path.exists(ref_so):
        lib = ctypes.CDLL(ref_so)
        fn = lambda: lib.ref_c2c_batch_mt(P(ys), P(xs), ctypes.c_int(n), ctypes.c_int64(sample), 0, dbl, cores)
        kind = "reference"
    elif os.path.exists(orc_so):
        lib = ctypes.CDLL(orc_so)
        fn = lambda: lib.orc_c2c_batch_mt(P(ys), P(xs), ctypes.c_int64(n), ctypes.c_int64(sample), 0, dbl, cores)
        kind = "port"
    else:
        return None
    best = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t0)
    flops = 5.0 * n * math.log2(n) * sample
    g = gpu_out_cpu[:sample].astype(np.complex128)
    agree = float(np.max(np.linalg.norm(g - ys, axis=1) / np.linalg.norm(ys, axis=1)))
    batch1 = None
    if kind == "reference":
        # genFFT's own benchmark shape (test/fft_bench.cpp:41-100): one complex64
        # transform per call, forward + inverse, one plan, on one thread; the
        # reference_cases leg times the same calls on the GPU
        lib.ref_c2c_pair_seconds.restype = ctypes.c_double
        lib.ref_c2c_pair_seconds.argtypes = [ctypes.c_int] * 3
        batch1 = []
        for k in range(1, 22):
            sec = lib.ref_c2c_pair_seconds(1 << k, max(1, int(3e6 / ((1 << k) * k))), 3)
            batch1.append({"n": 1 << k, "us": round(sec * 1e6, 3)})
    return {
        "value": round(flops / best / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": cores,
        "host_cores": host,
        "kind": kind,
        "sample": f"{sample} forward transforms of the same N={n} batch (first {sample} rows), "
                  f"best of 3, batch strided over {cores} threads sharing one plan (one thread per CPU "
                  f"this process may use: affinity mask and cgroup quota; the host has {host})",
        "seconds": round(best, 4),
        "gpu_vs_cpu_rel_l2_max": agree,
        "batch1_fwd_inv_us": batch1,
        "batch1_note": "genFFT, one complex64 transform per call, forward + unscaled inverse, one plan, one "
                       "thread, best of 3 loops (the FFT_1D rows of reference_cases on the GPU)",
    }


def accuracy(kind, x, y, n, n2, prec, inv=False):
    """Relative L2 error of spot transforms (first, middle, last of the batch;
    the whole image for 2D) against numpy's float64 FFT of the same input --
    an independent reference (the long-double oracle runs in the tests)."""
    import numpy as np
    import torch

    def host(t):
        t = t.detach()
        if t.dtype in (torch.float16, torch.bfloat16):  # [.., 2] pairs
            t = torch.view_as_complex(t.double().contiguous())
        return t.cpu().numpy().astype(np.complex128 if t.is_complex() else np.float64)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    if kind in ("c2c_2d", "r2c_2d", "c2r_2d"):
        xi, yi = host(x).reshape(n2, -1), host(y).reshape(n2, -1)
        if kind == "c2c_2d":
            ref = np.fft.fft2(xi)
        elif kind == "r2c_2d":
            ref = np.fft.fft2(xi)
        else:
            ref = np.fft.irfft2(xi, s=(n2, n)) * (n * n2)
        return {"rel_l2_max": rel(yi, ref), "spot": "whole image", "reference": "numpy.fft float64"}
    b = x.shape[0]
    rows = sorted({0, b // 2, b - 1})
    errs = []
    for r in rows:
        xr, yr = host(x[r]), host(y[r])
        if kind == "r2c":
            ref = np.fft.rfft(xr.real)
        elif kind == "c2r":
            ref = np.fft.irfft(xr, n) * n
        else:
            ref = np.fft.ifft(xr) * n if inv else np.fft.fft(xr)
        errs.append(rel(yr.reshape(ref.shape), ref))
    return {"rel_l2_max": max(errs), "spot": f"transforms {rows}", "reference": "numpy.fft float64"}


def _warm(fn, calls, min_ms=40.0):
    """Untimed calls: at least `calls`, and at least min_ms of wall time with
    the queue kept full.  An MI355X that was idle runs its first ~15-20 ms of
    sustained work slower while its clocks come up (c64 2^14 x 8192: 0.60 ms
    per call falling to 0.49 ms over the first 30 calls)."""
    import torch
    k, t0 = max(1, calls), time.perf_counter()
    while True:
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        if (time.perf_counter() - t0) * 1e3 >= min_ms:
            return
        k *= 2


def _time_ms(fn, reps, warm=2):
    import torch
    _warm(fn, warm)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def pattern_copy_gbs(src, dst, n1, n2, m, mode, order):
    """Algorithmic GB/s of a tile-pattern copy (libgfft_pattern_copy.so, mode
    0: m-column tiles of n1 x n2 matrices read and written; mode 1: m whole
    rows read, written transposed; order 0 interleaved tiles, 1 one run per
    workgroup) over the float32 buffers src / dst, best of 3 x 5 runs on
    torch's current stream."""
    import ctypes
    import torch
    lib = ctypes.CDLL(os.path.join(ROOT, "genfft_amd", "libgfft_pattern_copy.so"))
    fn = lib.gfft_pattern_copy
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    mats = src.numel() * 4 // (n1 * n2 * 8)
    stream = torch.cuda.current_stream().cuda_stream

    def run():
        rc = fn(src.data_ptr(), dst.data_ptr(), mats, n1, n2, m, mode, order, stream)
        if rc != 0:
            raise RuntimeError(f"gfft_pattern_copy({n1}, {n2}, {m}, mode {mode}) returned {rc}")

    ms = min(_time_ms(run, 5, 2) for _ in range(3))
    return 2.0 * mats * n1 * n2 * 8 / (ms * 1e-3) / 1e9


def _device_ms(fn, reps, warm=2):
    """GPU time per call with host overhead hidden: copies of a 256 MiB buffer
    (~0.1 ms each) hold the stream while all reps are enqueued.  Memory-bound
    work keeps the clocks where a sustained load puts them; a one-wave spin
    kernel lets them fall, and the calls after it measured 2-10% slower.  The
    start event must still be pending when enqueueing ends (else the hold is
    lengthened and the run repeated); None if the host never got ahead."""
    import torch
    _warm(fn, warm)
    src = torch.empty(1 << 26, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    copies = 200
    try:
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            for _ in range(copies):
                dst.copy_(src)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            primed = not e0.query()
            torch.cuda.synchronize()
            if primed:
                return e0.elapsed_time(e1) / reps
            copies *= 4
        return None
    finally:
        del src, dst


def bench_filter(dev):
    """Fused spectral filter forward + adjoint vs the unfused launches."""
    import torch

    import genfft_amd as G
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.randn(256, 16384, device=dev, generator=g).requires_grad_(True)
    H = (1 + 0.01 * torch.randn(8193, device=dev, generator=g)).to(torch.complex64).requires_grad_(True)
    gy = torch.randn(256, 16384, device=dev, generator=g)
    t = torch.roll(x.detach(), 3, -1) + 0.1 * torch.randn(256, 16384, device=dev, generator=g)
    def torch_fft_filter(xv, Hv):
        # the same operator composed from PyTorch's own FFTs (rocFFT), native autograd
        return torch.fft.irfft(torch.fft.rfft(xv) * Hv, xv.shape[-1])

    res = {}
    for name, fn in (("fused", G.spectral_filter), ("unfused", G.spectral_filter_plain),
                     ("unfused_torch_fft", torch_fft_filter)):
        def fwd_bwd():
            x.grad = None
            H.grad = None
            fn(x, H).backward(gy)

        def loss_step():
            x.grad = None
            H.grad = None
            ((fn(x, H) - t) ** 2).mean().backward()
        dms = _device_ms(fwd_bwd, 20)
        res[name] = {"fwd_ms": round(_time_ms(lambda: fn(x.detach(), H.detach()), 20), 4),
                     "fwd_adj_ms": round(_time_ms(fwd_bwd, 20), 4),
                     "fwd_adj_device_ms": None if dms is None else round(dms, 4),
                     "loss_step_ms": round(_time_ms(loss_step, 20), 4)}
    return {"config": "x [256,16384] f32 N(0,1), H [8193] complex64 = 1+0.01 N(0,1), "
                      "loss mean|y-t|^2, t = shift(x,3)+0.1 N(0,1)",
            "fused": res["fused"], "unfused": res["unfused"],
            "speedup_fwd_adj": round(res["unfused"]["fwd_adj_ms"] / res["fused"]["fwd_adj_ms"], 2),
            "speedup_fwd_adj_device": None if None in (res["unfused"]["fwd_adj_device_ms"],
                                                       res["fused"]["fwd_adj_device_ms"]) else
            round(res["unfused"]["fwd_adj_device_ms"] / res["fused"]["fwd_adj_device_ms"], 2),
            "timing": "*_ms: events around 20 calls, host launch and autograd-engine time included; "
                      "*_device_ms: the same calls enqueued behind copies that hold the stream (device time only)",
            "speedup_fwd": round(res["unfused"]["fwd_ms"] / res["fused"]["fwd_ms"], 2),
            "speedup_fwd_adj_vs_torch_fft": round(res["unfused_torch_fft"]["fwd_adj_ms"] / res["fused"]["fwd_adj_ms"], 2),
            "baselines": "unfused: this library's rfft / irfft launches plus PyTorch elementwise ops; "
                         "unfused_torch_fft: torch.fft.rfft / irfft (rocFFT) with PyTorch's own autograd",
            "speedup_loss_step": round(res["unfused"]["loss_step_ms"] / res["fused"]["loss_step_ms"], 2),
            # forward: x in, y out; adjoint: x and gy in, gx and gH out (the gH partial rows are the
            # fused path's own workspace, not algorithmic bytes)
            "roofline_fwd_adj_device": None if res["fused"]["fwd_adj_device_ms"] is None else {
                "alg_bytes": _FILTER_ALG_BYTES,
                "achieved_gbs": round(_FILTER_ALG_BYTES / (res["fused"]["fwd_adj_device_ms"] * 1e-3) / 1e9, 1),
                "frac": round(_FILTER_ALG_BYTES / (res["fused"]["fwd_adj_device_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}


_FILTER_ALG_BYTES = 5 * 256 * 16384 * 4 + 3 * 8193 * 8  # x, y, x, gy, gx (float32); H twice, gH


def bench_uneven(dev):
    """20,000 complex64 transforms, sizes 2^k with k uniform in 6..16, mixed directions."""
    import numpy as np
    import torch

    import genfft_amd as G
    rng = np.random.default_rng(0)
    count = 20000
    sizes = (1 << rng.integers(6, 17, count)).astype(np.int64)
    inv = rng.integers(0, 2, count).astype(bool)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    total = int(sizes.sum())
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.complex(torch.rand(total, device=dev, generator=g) * 2 - 1,
                      torch.rand(total, device=dev, generator=g) * 2 - 1)
    y = torch.empty_like(x)
    b = G.UnevenBatch(sizes, offs, offs, inv, torch.complex64)
    ms = {m: _time_ms(lambda: b.execute(x, y, mode=m), r, 1) for m, r in
          (("graph", 20), ("eager", 20), ("per_job", 3))}
    # 500 of the 20,000 jobs against numpy's float64 FFT (y holds the last run's output)
    b.execute(x, y, mode="graph")
    torch.cuda.synchronize()
    xs, ys = x.cpu().numpy().astype(np.complex128), y.cpu().numpy().astype(np.complex128)
    worst = 0.0
    for j in np.random.default_rng(1).choice(count, 500, replace=False):
        o, m = int(offs[j]), int(sizes[j])
        ref = np.fft.ifft(xs[o:o + m]) * m if inv[j] else np.fft.fft(xs[o:o + m])
        worst = max(worst, float(np.linalg.norm(ys[o:o + m] - ref) / np.linalg.norm(ref)))
    return {"config": f"{count} complex64 transforms, sizes 2^6..2^16 (seed 0), 50/50 forward/inverse, "
                      f"{total} elements", "graph_ms": round(ms["graph"], 4), "eager_ms": round(ms["eager"], 4),
            "per_job_ms": round(ms["per_job"], 4),
            "speedup_vs_per_job": round(ms["per_job"] / ms["graph"], 2),
            "rel_l2_max_500_jobs": worst, "accuracy_reference": "numpy.fft float64, 500 random jobs, bound "
                                                                 "2e-7*log2(n) <= 3.2e-6",
            "alg_gbs_graph": round(total * 16 / (ms["graph"] * 1e-3) / 1e9, 1)}


def _vert_args():
    """(nfft, ncols) of test/fft_bench.cpp:155-165 (ncols += 1 up to 16, then
    ncols + 3; nfft * ncols < 2^22), every third ncols of each nfft kept."""
    out = []
    nfft = 1
    while nfft <= (1 << 20):
        cols, ncols = [], 1
        while ncols * nfft < (1 << 22):
            cols.append(ncols)
            ncols += 1 if ncols <= 16 else ncols + 3
        out += [(nfft, c) for c in cols[::3]]
        nfft *= 4
    return out


def bench_reference_cases(dev):
    """The reference's own benchmark cases (test/fft_bench.cpp:41-176), one
    transform per call as genFFT runs them:
      FFT_1D      forward + unscaled inverse of one complex64 transform, n = 2 .. 2^21
      RealFFT_1D  one real float32 transform to n/2 + 1 bins, n = 2 .. 2^21
      FFT_Vert    forward + inverse column FFT, nfft x ncols (subsampled sweep)
    each on host memory (the drop-in path genFFT callers take: staged through
    the device, synchronous) and on device memory (HIP events around the
    calls).  err: the reference's own check, mean |inverse(forward(x))/n - x|."""
    import ctypes

    import numpy as np
    import torch

    import genfft_amd as G
    from genfft_amd import _capi as C
    lib = C.lib()
    rng = np.random.default_rng(0)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)

    def host_us(fn, reps):
        fn()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            best = min(best, time.perf_counter() - t0)
        return best * 1e6

    def case(kind, n, n2, x_np, out_np, inv_plan=None):
        """(host us, device us, err) for one forward (+ inverse) pair."""
        plan = G.Plan(kind, n, n2=n2, half=kind == C.GFFT_R2C)
        back_np = np.empty_like(x_np) if inv_plan else None

        def host():
            C.check(lib.gfft_execute_host(plan._h, P(x_np), P(out_np), 0))
            if inv_plan:
                C.check(lib.gfft_execute_host(plan._h, P(out_np), P(back_np), 1))
        reps = 20 if x_np.nbytes < (1 << 22) else 5
        h = host_us(host, reps)
        err = float(np.mean(np.abs(back_np / n - x_np))) if inv_plan else None
        xd = torch.from_numpy(x_np).to(dev)
        yd = torch.empty(out_np.shape, dtype=torch.from_numpy(out_np).dtype, device=dev)
        zd = torch.empty_like(xd) if inv_plan else None

        def device():
            plan.execute(xd, yd)
            if inv_plan:
                plan.execute(yd, zd, inverse=True)
        d = _time_ms(device, 50, 3) * 1e3
        g = _device_ms(device, 50)
        del plan
        return round(h, 2), round(d, 2), (None if g is None else round(g * 1e3, 2)), err

    out = {"timing": "host_us: best of 5-20 synchronous calls on host buffers (H2D + transform + D2H, the "
                     "genFFT drop-in's memory model); device_us: mean of 50 back-to-back calls on device buffers "
                     "from Python, HIP events (host launch cost included); gpu_us: the same 50 calls enqueued "
                     "behind copies that hold the stream, i.e. the GPU's time alone",
           "FFT_1D": [], "RealFFT_1D": [], "FFT_Vert": []}
    for k in range(1, 22):
        n = 1 << k
        if k % 5 == 1:
            log(f"reference cases: FFT_1D / RealFFT_1D n = 2^{k}")
        x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
        h, d, g, err = case(C.GFFT_C2C, n, 0, x, np.empty_like(x), inv_plan=True)
        out["FFT_1D"].append({"n": n, "host_us": h, "device_us": d, "gpu_us": g, "err": err})
        r = rng.uniform(-1, 1, n).astype(np.float32)
        h, d, g, _ = case(C.GFFT_R2C, n, 0, r, np.empty(n // 2 + 1, np.complex64))
        out["RealFFT_1D"].append({"n": n, "host_us": h, "device_us": d, "gpu_us": g})
    for i, (nfft, ncols) in enumerate(_vert_args()):
        if i % 20 == 0:
            log(f"reference cases: FFT_Vert {nfft} x {ncols}")
        x = (rng.uniform(-1, 1, nfft * ncols) + 1j * rng.uniform(-1, 1, nfft * ncols)).astype(np.complex64)
        h, d, g, err = case(C.GFFT_COLS, nfft, ncols, x, np.empty_like(x), inv_plan=True)
        out["FFT_Vert"].append({"nfft": nfft, "ncols": ncols, "host_us": h, "device_us": d, "gpu_us": g, "err": err})
    return out


def bench_tile_widths(dev):
    """The single-pass kernels the autotuner can pick for complex64 rows of
    1024 points (the LDS tiles of 2-8 rows and 128-512 threads, and the
    one-wavefront kernel with no LDS), each timed by gfft_autotune at batches
    from one row to 1 GiB; and which one the planner takes by default."""
    import ctypes

    import torch

    import genfft_amd as G
    from genfft_amd import _capi as C
    out = []
    for b in (1, 16, 256, 2048, 8192, 32768, 131072):
        C.check(C.lib().gfft_tuning_clear())
        G.clear_plan_cache()
        plan = G.Plan(C.GFFT_C2C, 1024, batch=b)
        default = plan.describe().splitlines()[1]
        x = torch.zeros(b, 1024, dtype=torch.complex64, device=dev)
        y = torch.empty_like(x)
        _warm(lambda: plan.execute(x, y), 2)  # the clocks up before the candidates are timed
        del plan, x, y
        rep = G.autotune(1024, b, torch.complex64, reps=50 if b < 8192 else 20)
        rows = []
        for c in sorted(rep["candidates"], key=lambda c: c["ms"]):
            d = C.GfftDesc(C.GFFT_C2C, 0, 1024, 0, b, 0, 0, 0, 0, 0, 0)
            # the candidate's kernel, as the planner describes it under that tuning entry
            path = os.path.join("/tmp", f"gfft_tile_widths_{os.getpid()}.json")
            with open(path, "w") as f:
                f.write('{"version": 1, "entries": [\n {"precision": 0, "n": 1024, "n1_log2": 0, "col_m": 0, '
                        f'"row_m": 0, "pipe_mask": 3, "row_variant": {c["row_variant"]}}}\n]}}\n')
            C.check(C.lib().gfft_tuning_clear())
            G.load_tuning(path)
            os.remove(path)
            buf = ctypes.create_string_buffer(1 << 14)
            C.check(C.lib().gfft_plan_explain(ctypes.byref(d), buf, len(buf)))
            line = buf.value.decode().splitlines()[1]
            shape = " ".join(w for w in line.split() if w.split("=")[0] in ("M", "NT", "lds", "rad"))
            rows.append({"us": round(c["ms"] * 1e3, 2), "kernel": shape})
        C.check(C.lib().gfft_tuning_clear())
        out.append({"batch": b, "default": " ".join(w for w in default.split()
                                                     if w.split("=")[0] in ("M", "NT", "lds", "rad")),
                    "candidates": rows})
    return {"config": "complex64, n = 1024, forward; each candidate timed by gfft_autotune (events around "
                      "repeated calls after two untimed ones)", "batches": out}


def bench_precision(dev):
    """Log-uniform |x| in 1e-3..1e3, uniform phase, in FP16 + scale, BF16 and FP32:
    N = 2^20 batch 64 (seed 20) and one 8192 x 8192 image (seed 21)."""
    import math as m_

    import numpy as np
    import torch

    import genfft_amd as G
    from genfft_amd import _capi as C

    def workload(shape, seed, kind, n, n2, batch, ref_fn, flops, points):
        g = torch.Generator(device=dev).manual_seed(seed)
        mag = torch.pow(10.0, torch.rand(shape, device=dev, generator=g, dtype=torch.float64) * 6 - 3)
        z = torch.polar(mag, torch.rand(shape, device=dev, generator=g, dtype=torch.float64) * (2 * m_.pi))
        del mag
        res = {}
        for mode in ("f16", "bf16", "f32"):
            if mode == "f32":
                x = z.to(torch.complex64)
                xd = x.to(torch.complex128)
                y = torch.empty_like(x)
                plan = G.Plan(kind, n, n2=n2, precision=C.GFFT_F32, batch=batch)
                run = lambda: plan.execute(x, y)
            else:
                x = torch.view_as_real(z).to(torch.float16 if mode == "f16" else torch.bfloat16).contiguous()
                xd = torch.view_as_complex(x.double().contiguous())
                y = torch.empty_like(x)
                sc = torch.empty(batch, dtype=torch.float32, device=dev)
                plan = G.Plan(kind, n, n2=n2, precision=C.GFFT_F16 if mode == "f16" else C.GFFT_BF16,
                              batch=batch)
                run = lambda: plan.execute_scaled(x, y, sc, True)
            ms = _time_ms(run, 20)
            run()
            ref = ref_fn(xd)
            yd = y.to(torch.complex128) if mode == "f32" else \
                torch.view_as_complex(y.double().contiguous()) / sc.double().view((-1,) + (1,) * (ref.dim() - 1))
            err = yd - ref
            res[mode] = {"ms": round(ms, 4), "gflops": round(flops / (ms * 1e-3) / 1e9, 1),
                         "alg_gbs": round(2.0 * points * (8 if mode == "f32" else 4) / (ms * 1e-3) / 1e9, 1),
                         "rel_l2_vs_c128": float(torch.linalg.vector_norm(err) / torch.linalg.vector_norm(ref)),
                         "max_abs_vs_c128": float(err.abs().max())}
            # an independent reference: numpy's float64 FFT of the same stored input
            # (transforms 0, batch/2 and batch-1 of the 1D batch; the whole 2D image)
            if kind == C.GFFT_C2C:
                rows = [0, batch // 2, batch - 1]
                xs, ys = xd[rows].cpu().numpy(), yd[rows].cpu().numpy()
                rn = np.fft.fft(xs, axis=-1)
            else:
                xs, ys = xd.cpu().numpy(), yd.cpu().numpy()
                rn = np.fft.fft2(xs, axes=(-2, -1))
            res[mode]["rel_l2_vs_numpy_f64"] = float(np.linalg.norm(ys - rn) / np.linalg.norm(rn))
            del x, xd, y, ref, yd, err, plan, xs, ys, rn
        return res

    n, batch = 1 << 20, 64
    out = workload((batch, n), 20, C.GFFT_C2C, n, 0, batch, G.fft, 5.0 * n * m_.log2(n) * batch, n * batch)
    side = 8192
    out["2d_8192x8192"] = workload((1, side, side), 21, C.GFFT_C2C_2D, side, side, 1, G.fft2,
                                   5.0 * side * side * m_.log2(side * side), side * side)
    out["note"] = ("FP16/BF16 include the automatic per-transform scale pass (one extra read); "
                   "rel_l2_vs_c128: against this library's complex128 plan on the whole output; "
                   "rel_l2_vs_numpy_f64: against numpy's float64 FFT of the same stored input (3 transforms "
                   "of the 1D batch, the whole 2D image); against the long-double oracle: "
                   "tests/test_gpu_precision.py")
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are out of scope "
                                   "(each transform batch fits one MI355X); use --gpus 1"}))
        sys.exit(2)
    import numpy as np
    import torch

    import genfft_amd as G
    from genfft_amd import _capi as C

    if a.leg:
        leg = {"fused_filter": bench_filter, "uneven_batch": bench_uneven, "precision": bench_precision,
               "reference_cases": bench_reference_cases, "tile_widths": bench_tile_widths}[a.leg]
        print(json.dumps({a.leg: leg(torch.device("cuda", 0))}), flush=True)
        return
    kind, n2 = "c2c", 0
    if a.workload in OTHER:
        kind, n, n2, batch = OTHER[a.workload]
        prec = "c64"
    else:
        prec, n, batch = WORKLOADS[a.workload]
    half = prec in ("f16", "bf16")
    cdt = torch.complex128 if prec == "c128" else torch.complex64
    csize = CSIZE[prec]
    dev = torch.device("cuda", 0)
    gen = torch.Generator(device=dev).manual_seed(0)
    if kind == "r2c":
        # float32 i.i.d. N(0,1) rows -> n/2 + 1 bins each
        x = torch.randn(batch, n, device=dev, generator=gen)
        y = torch.empty(batch, n // 2 + 1, dtype=torch.complex64, device=dev)
        plan = G.Plan(C.GFFT_R2C, n, precision=0, batch=batch, half=True)
    elif kind == "c2c_2d":
        x = torch.complex(torch.randn(n2, n, device=dev, generator=gen), torch.randn(n2, n, device=dev, generator=gen))
        y = torch.empty_like(x)
        plan = G.Plan(C.GFFT_C2C_2D, n, precision=0, n2=n2, batch=batch)
    elif kind == "c2r":
        # half spectra of float32 N(0,1) rows (bins 0 and n/2 real) -> n reals each, unscaled
        x = torch.fft.rfft(torch.randn(batch, n, device=dev, generator=gen)).contiguous()
        y = torch.empty(batch, n, device=dev)
        plan = G.Plan(C.GFFT_C2R, n, precision=0, batch=batch)
    elif kind == "r2c_2d":
        # real n2 x n image -> the full n2 x n complex spectrum (RealFFT2D::forward)
        x = torch.randn(n2, n, device=dev, generator=gen)
        y = torch.empty(n2, n, dtype=torch.complex64, device=dev)
        plan = G.Plan(C.GFFT_R2C_2D, n, precision=0, n2=n2, batch=batch)
    elif kind == "c2r_2d":
        # the n/2 + 1 non-redundant columns of a real image's spectrum -> the image (RealFFT2D::inverse)
        x = torch.fft.rfft2(torch.randn(n2, n, device=dev, generator=gen)).contiguous()
        y = torch.empty(n2, n, device=dev)
        plan = G.Plan(C.GFFT_C2R_2D, n, precision=0, n2=n2, batch=batch, in_stride=n // 2 + 1, out_stride=n)
    else:
        # re/im i.i.d. N(0,1) from a fixed seed
        x = torch.complex(torch.randn(batch, n, device=dev, generator=gen, dtype=cdt.to_real()),
                          torch.randn(batch, n, device=dev, generator=gen, dtype=cdt.to_real()))
        if half:
            x = torch.view_as_real(x).to(torch.float16 if prec == "f16" else torch.bfloat16).contiguous()
        y = torch.empty_like(x)
        plan = G.Plan(C.GFFT_C2C, n, precision={"c64": 0, "c128": 1, "f16": 2, "bf16": 3}[prec], batch=batch)

    if a.calib_copy:
        # a plain copy of known size: reads and writes exactly 1 GiB
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    inv = kind in ("c2r", "c2r_2d")
    for _ in range(a.warmup):
        plan.execute(x, y, inverse=inv)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(a.steps):
        plan.execute(x, y, inverse=inv)
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_time(e1)
    ms_step = ms_total / a.steps

    if kind in ("r2c", "c2r"):
        # real-input convention: half the complex flops; n reals and n/2 + 1 bins
        flops = 2.5 * n * math.log2(n) * batch
        alg_bytes = (4.0 * n + 8.0 * (n // 2 + 1)) * batch
    elif kind == "r2c_2d":
        N = n * n2
        flops = 2.5 * N * math.log2(N) * batch
        alg_bytes = (4.0 * N + 8.0 * N) * batch
    elif kind == "c2r_2d":
        N = n * n2
        flops = 2.5 * N * math.log2(N) * batch
        alg_bytes = (8.0 * (n // 2 + 1) * n2 + 4.0 * N) * batch
    elif kind == "c2c_2d":
        N = n * n2
        flops = 5.0 * N * math.log2(N) * batch
        alg_bytes = 2.0 * N * csize * batch
    else:
        flops = 5.0 * n * math.log2(n) * batch
        alg_bytes = 2.0 * n * csize * batch
    gflops = flops / (ms_step * 1e-3) / 1e9

    # per-launch kernel time: each launch repeated back to back between one
    # HIP event pair on the execution stream (an event pair around every
    # single launch adds its own cost to the kernel's)
    launches = tile_launches(plan.describe())
    acc = [plan.profile_launch(x, y, l["index"], reps=20, inverse=inv) for l in launches]
    groups = {}
    for l, t in zip(launches, acc):
        if l["kind"] == "tile":
            key = _tile_key(l)
            elems = _tile_elems(l)
        elif l["kind"] == "dual":
            # column pass of one chunk beside the row pass of the previous one:
            # each role reads and writes its own elements
            key = "dual " + " + ".join(_tile_key(r) for r in l["roles"])
            elems = sum(_tile_elems(r) for r in l["roles"])
        elif l["kind"] == "real_rows":
            # rows of n = 2L reals <-> bins in one pass (unpack / pack fused)
            key = f"real rows {l['mode']} L={l['L']} M={l['M']}"
            elems = int(l["rows"]) * int(l["L"])
        elif l["kind"] == "c2r_pack_cols":
            # reads n/2 + 1 bins, writes the n/2-point column-pass output per transform
            key = f"c2r pack+cols L={l['L']} M={l['M']}"
            elems = int(l["tiles"]) // (int(l["n2"]) // int(l["M"])) * int(l["L"]) * int(l["n2"])
        elif l["kind"] == "r2c_rows_unpack":
            # reads the n/2-point column-pass output, writes n/2 + 1 bins per transform
            key = f"r2c rows+unpack R={l['R']} M={l['M']}"
            elems = int(l["tiles"]) // (int(l["n1"]) // int(l["M"])) * int(l["n1"]) * int(l["R"])
        else:
            key, elems = l["kind"], 0
        g = groups.setdefault(key, {"launches": 0, "ms": 0.0, "bytes": 0.0})
        g["launches"] += 1
        g["ms"] += t
        g["bytes"] += 2.0 * elems * csize
    dom_key = max(groups, key=lambda k: groups[k]["ms"])
    dom = groups[dom_key]
    per_launch_ms = dom["ms"] / dom["launches"]
    per_launch_bytes = dom["bytes"] / dom["launches"]
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        t = tj.get(a.workload, {}).get(dom_key)
        if t:
            traffic = t.get("bytes_per_launch")

    # the box's own ceiling: a plain device copy of 1 GiB (torch), best of 5
    src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    copy_ms = min(_time_ms(lambda: dst.copy_(src), 5, 2) for _ in range(3))
    copy_gbs = 2.0 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    # the same copy in the access pattern of each complex64 four-step tile pass
    # with 16384-element tiles (tile keys "tile L=<L> M=<M> col->col" / "row->col")
    pattern = None
    if prec == "c64" and kind == "c2c":
        import re
        pats = {}
        for k in groups:
            mt = re.fullmatch(r"tile L=(\d+) M=(\d+) (col->col|row->col)", k)
            if mt and int(mt.group(1)) * int(mt.group(2)) == 16384 and n % int(mt.group(1)) == 0:
                L_, M_ = int(mt.group(1)), int(mt.group(2))
                shape = (L_, n // L_, M_, 0) if mt.group(3) == "col->col" else (n // L_, L_, M_, 1)
                pats[k] = {o: pattern_copy_gbs(src, dst, *shape, i) for i, o in enumerate(("interleaved", "runs"))}
        if pats:
            pattern = {k: {o: round(g, 1) for o, g in v.items()} for k, v in pats.items()}
            if dom_key in pats:
                # against the faster of the two tile orders
                pattern["frac_of_pattern"] = round(achieved / max(pats[dom_key].values()), 4)
    del src, dst

    # the launches were repeated one by one above: one whole call restores y
    plan.execute(x, y, inverse=inv)
    torch.cuda.synchronize()
    acc_line = accuracy(kind, x, y, n, n2, prec, inv) if not half else None
    if half:  # 16-bit storage: against the float64 FFT of the stored (dequantized) input
        acc_line = accuracy("c2c", x, y, n, n2, prec)
    cpu = None
    if not a.no_cpu and not half and kind == "c2c":
        log("cpu baseline")
        cpu = cpu_baseline(x.cpu().numpy(), n, batch, prec, y.cpu().numpy())
    del x, y, plan
    torch.cuda.empty_cache()
    extras = {}
    if not a.no_extras and a.workload == "c64_n2^20_b128":
        for name, leg in (("fused_filter", bench_filter), ("uneven_batch", bench_uneven),
                          ("precision", bench_precision), ("reference_cases", bench_reference_cases)):
            log(f"leg {name}")
            extras[name] = leg(dev)

    line = {
        "metric": "batched c2c fp32 FFT GFLOP/s & %HBM peak at N=2^20; fused fwd+adjoint speedup",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"c64": "f32", "c128": "f64", "f16": "f16 storage, f32 math", "bf16": "bf16 storage, f32 math"}[prec],
        "data": "synthetic: re/im i.i.d. N(0,1), torch generator seed 0, resident in HBM",
        "config": {"workload": (f"batched 1D c2c forward, "
                                f"{ {'c64': 'complex64', 'c128': 'complex128', 'f16': 'FP16 pairs', 'bf16': 'BF16 pairs'}[prec]}, "
                                f"N=2^{int(math.log2(n))}, batch {batch}") if kind == "c2c" else {
                       "r2c": f"batched real float32 -> n/2+1 bins, N=2^{int(math.log2(n))}, batch {batch}",
                       "c2r": f"batched n/2+1 bins -> real float32 (unscaled inverse), N=2^{int(math.log2(n))}, "
                              f"batch {batch}",
                       "c2c_2d": f"2D complex64 forward {n}x{n2}",
                       "r2c_2d": f"2D real float32 {n}x{n2} -> full complex64 spectrum",
                       "c2r_2d": f"2D n/2+1 spectrum columns -> real float32 {n}x{n2} (unscaled inverse)"}[kind],
                   "n": n, "batch": batch, "precision": prec, "plan": [k for k in groups]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom_key,
                     "copy_gbs": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4),
                     "kernel_ms": round(per_launch_ms, 5), "alg_bytes_per_launch": per_launch_bytes,
                     "pattern_copy_gbs": pattern},
        "e2e": {"alg_bytes": alg_bytes, "achieved_gbs": round(alg_bytes / (ms_step * 1e-3) / 1e9, 1),
                "hbm_frac": round(alg_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "frac_of_copy": round(alg_bytes / (ms_step * 1e-3) / 1e9 / copy_gbs, 4),
                "kernels": {k: {"launches": v["launches"], "ms_per_step": round(v["ms"], 5)}
                            for k, v in groups.items()},
                "wall_s": round(wall, 3)},
        "cpu_baseline": cpu,
        "accuracy": acc_line,
    }
    line.update(extras)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
