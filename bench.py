#!/usr/bin/env python3
"""Benchmark: batched complex64 forward FFT, N = 2^20, batch 128, on one MI355X.

One step = one forward transform of the whole batch (1 GiB in, 1 GiB out),
device-resident, through the C ABI (gfft_execute) on torch's current stream.
value = GFLOP/s in the 5*N*log2(N) convention.

The JSON line also carries
  roofline      -- the dominant kernel: algorithmic bytes per launch (one read
                   and one write of the elements it transforms) over its mean
                   duration, measured live with HIP events recorded on the
                   execution stream around every launch (gfft_plan_profile);
                   `traffic` is the HBM bytes per launch measured by rocprofv3
                   PMC counters, read from profiles/pmc_traffic.json when present;
  e2e           -- the whole transform against the HBM roofline;
  cpu_baseline  -- genFFT itself (oracle/_ref, built from the reference sources)
                   timed on the host cores over a bounded sample, or the
                   oracle's C restatement when that library is absent.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
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
WORKLOADS = {
    # name: (precision, n, batch)
    "c64_n2^20_b128": ("c64", 1 << 20, 128),
    "c64_n2^10_b131072": ("c64", 1 << 10, 131072),
    "c64_n2^14_b8192": ("c64", 1 << 14, 8192),
    "c64_n2^18_b512": ("c64", 1 << 18, 512),
    "c64_n2^22_b32": ("c64", 1 << 22, 32),
    "c128_n2^20_b64": ("c128", 1 << 20, 64),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c64_n2^20_b128", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain 1 GiB device copy (scale reference for PMC byte counters)")
    return ap.parse_args()


def tile_launches(text):
    out = []
    for ln in text.splitlines():
        if not ln.startswith(("tile", "unpack", "pack", "hfill", "copy")):
            continue
        f = dict(kv.split("=", 1) for kv in ln.split()[1:] if "=" in kv)
        f["kind"] = ln.split()[0]
        out.append(f)
    return out


def cpu_baseline(x_cpu, n, batch, prec, gpu_out_cpu):
    """genFFT (or the oracle's restatement) on the host cores, bounded sample."""
    import ctypes

    import numpy as np
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libgenfft_ref.so")
    orc_so = os.path.join(ROOT, "oracle", "liboracle.so")
    cores = max(1, min(16, os.cpu_count() or 1))
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
    return {
        "value": round(flops / best / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": cores,
        "kind": kind,
        "sample": f"{sample} forward transforms of the same N={n} batch (first {sample} rows), "
                  f"best of 3, batch strided over {cores} threads sharing one plan",
        "seconds": round(best, 4),
        "gpu_vs_cpu_rel_l2_max": agree,
    }


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

    prec, n, batch = WORKLOADS[a.workload]
    cdt = torch.complex64 if prec == "c64" else torch.complex128
    csize = 8 if prec == "c64" else 16
    dev = torch.device("cuda", 0)
    gen = torch.Generator(device=dev).manual_seed(0)
    # re/im i.i.d. N(0,1) from a fixed seed
    x = torch.complex(torch.randn(batch, n, device=dev, generator=gen, dtype=cdt.to_real()),
                      torch.randn(batch, n, device=dev, generator=gen, dtype=cdt.to_real()))
    y = torch.empty_like(x)
    plan = G.Plan(C.GFFT_C2C, n, precision=0 if prec == "c64" else 1, batch=batch)

    if a.calib_copy:
        # a plain copy of known size: reads and writes exactly 1 GiB
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    for _ in range(a.warmup):
        plan.execute(x, y)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(a.steps):
        plan.execute(x, y)
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_time(e1)
    ms_step = ms_total / a.steps

    flops = 5.0 * n * math.log2(n) * batch
    alg_bytes = 2.0 * n * csize * batch
    gflops = flops / (ms_step * 1e-3) / 1e9

    # per-launch profile (HIP events around each kernel on the execution stream)
    launches = tile_launches(plan.describe())
    reps = 10
    acc = [0.0] * len(launches)
    for _ in range(reps):
        ms = plan.profile(x, y)
        for i, v in enumerate(ms[:len(launches)]):
            acc[i] += v / reps
    groups = {}
    for l, t in zip(launches, acc):
        if l["kind"] == "tile":
            key = f"tile L={l['L']} M={l['M']} {'row' if l['inp'] == '0' else 'col'}->" \
                  f"{'row' if l['outp'] == '0' else 'col'}"
            elems = int(l["n_inner"]) * int(l["tiles"]) // int(l["tiles_inner"]) * int(l["L"])
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

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(x.cpu().numpy(), n, batch, prec, y.cpu().numpy())

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
        "dtype": "f32" if prec == "c64" else "f64",
        "data": "synthetic: re/im i.i.d. N(0,1), torch generator seed 0, resident in HBM",
        "config": {"workload": f"batched 1D c2c forward, {'complex64' if prec == 'c64' else 'complex128'}, "
                               f"N=2^{int(math.log2(n))}, batch {batch}",
                   "n": n, "batch": batch, "precision": prec, "plan": [k for k in groups]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom_key,
                     "kernel_ms": round(per_launch_ms, 5), "alg_bytes_per_launch": per_launch_bytes},
        "e2e": {"alg_bytes": alg_bytes, "achieved_gbs": round(alg_bytes / (ms_step * 1e-3) / 1e9, 1),
                "hbm_frac": round(alg_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "kernels": {k: {"launches": v["launches"], "ms_per_step": round(v["ms"], 5)}
                            for k, v in groups.items()},
                "wall_s": round(wall, 3)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
