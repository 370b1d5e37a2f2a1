#!/usr/bin/env python3
"""Benchmark: the ten open-earth-benchmarks stencils, plain FP64, on one MI355X.

A step launches each of the ten stencils once, in a fixed order, on its own
256x256x80 fields (halo 3, k-halo 1), all resident in HBM.  value = interior
grid points processed per second over the whole step (10 x 256*256*80 points
per step), in Gpts/s.  Each kernel's launches are bracketed by HIP events on
the stream they run on.  The per-stencil Gpts/s and fraction of the 8 TB/s
HBM peak come from those events, computed with the algorithmic bytes per point
of each stencil (inputs read once, outputs written once, FP64).  Between two
launches of one stencil the other nine move ~3 GB, so nothing it reads is
still in the 256 MB MALL when it runs again.

cpu_baseline: the reference's own kernels (OpenMP variants, serial DEFAULT for
fvtp2d, which has no OpenMP variant), compiled by oracle/Makefile, timed on
this host over a bounded sample of the same workload.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256] [--nz 80]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "open-earth-benchmarks_amd"))

METRIC = "Gpts/s + % HBM peak per stencil; FP32/batched/adjoint vs plain FP64 at stated err"
HBM_PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"


def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--size", type=int, default=256, help="nx = ny")
    p.add_argument("--nz", type=int, default=80)
    p.add_argument("--kernels", default="all", help="comma-separated subset (default: all ten)")
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--cpu-budget-s", type=float, default=20.0, help="wall-clock budget of the CPU baseline")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="also time a plain device copy of this many MiB (HBM counter calibration)")
    return p.parse_args(argv)


def _traffic(size: int, nz: int, kernel: str):
    """HBM bytes per launch of `kernel` from the committed PMC summary, if one
    exists for this shape (profiles/hbm_traffic_<size>x<size>x<nz>.json)."""
    f = PROFILES / f"hbm_traffic_{size}x{size}x{nz}.json"
    if not f.exists():
        return None
    try:
        data = json.loads(f.read_text())
        return data["kernels"][kernel]["hbm_bytes_per_launch"]
    except (KeyError, ValueError):
        return None


def cpu_baseline(names, hosts, L, budget_s: float):
    """Times the reference's kernels (or the oracle when the reference build is
    absent) on this host over whole steps of the same workload."""
    sys.path.insert(0, str(ROOT))
    from oracle import binding as B

    avail = B.cpu_count_physical()
    threads = max(1, min(16, avail))
    if B.ref_available():
        kind = "reference"
        B.ref_lib().ref_set_threads(threads)
        run = lambda n, h: B.run_ref(n, h, L)  # noqa: E731
        what = "reference kernels (OpenMP variants; fvtp2d serial DEFAULT)"
    else:
        kind = "port"
        threads = 1
        run = lambda n, h: B.run_oracle(n, h, L)  # noqa: E731
        what = "C oracle (serial)"
    pts = L.nx * L.ny * L.nz
    steps, t_total, per = 0, 0.0, {n: [] for n in names}
    t_start = time.perf_counter()
    while True:
        for n in names:
            t0 = time.perf_counter()
            run(n, hosts[n])
            per[n].append(time.perf_counter() - t0)
        steps += 1
        t_total = time.perf_counter() - t_start
        if t_total >= budget_s or steps >= 5:
            break
    value = len(names) * pts * steps / t_total / 1e9
    return {
        "value": round(value, 6),
        "unit": "Gpts/s",
        "cores": threads,
        "kind": kind,
        "sample": f"{steps} step(s) of {len(names)} stencils at {L.nx}x{L.ny}x{L.nz}, {what}, {threads} thread(s)",
        "per_stencil_ms": {n: round(1e3 * statistics.median(v), 3) for n, v in per.items()},
    }


def main(argv=None):
    args = parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (run with --gpus 1)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from oeb import stencils
    from oeb.generators import generate
    from oeb.layout import Layout, Layout1D
    from oeb.spec import KERNELS, ORDER
    from oeb.storage import Storage1D, Storage3D

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    torch.cuda.set_device(0)
    names = ORDER if args.kernels == "all" else [k.strip() for k in args.kernels.split(",")]
    n = args.size
    L = Layout(n, n, args.nz, h=3, hk=1, align=16)
    L1 = Layout1D(n, 3)
    pts = n * n * args.nz

    # Inputs: seeded host arrays, copied once to HBM; every stencil has its own.
    t0 = time.perf_counter()
    hosts, dev = {}, {}
    for name in names:
        h = generate(name, L, "baseline", args.seed)
        hosts[name] = h
        d = {}
        for a in KERNELS[name].args:
            if a.kind == "scratch":
                continue
            d[a.name] = (Storage1D if a.kind == "in1" else Storage3D).to_device(L1 if a.kind == "in1" else L,
                                                                               h[a.name])
        dev[name] = d
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0

    def step(ev=None):
        for name in names:
            if ev is not None:
                ev[name][0].record()
            stencils.run(name, dev[name], domain=(n, n, args.nz))
            if ev is not None:
                ev[name][1].record()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    events = [{nm: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for nm in names}
              for _ in range(args.steps)]
    begin, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    begin.record()
    for s in range(args.steps):
        step(events[s])
    end.record()
    torch.cuda.synchronize()
    total_ms = begin.elapsed_time(end)

    per = {}
    for name in names:
        durs = [events[s][name][0].elapsed_time(events[s][name][1]) for s in range(args.steps)]
        avg_ms = sum(durs) / len(durs)
        bpp = KERNELS[name].bytes_per_point_f64
        gbs = bpp * pts / (avg_ms * 1e-3) / 1e9
        per[name] = {
            "us": round(avg_ms * 1e3, 3),
            "us_median": round(statistics.median(durs) * 1e3, 3),
            "gpts_s": round(pts / (avg_ms * 1e-3) / 1e9, 3),
            "bytes_per_pt": bpp,
            "gbs": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
        }

    copy_line = None
    if args.copy_mib > 0:
        nbytes = args.copy_mib << 20
        src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        copy_line = {"bytes_read": nbytes, "bytes_written": nbytes, "us": round(ms * 1e3, 2),
                     "gbs": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}

    dom = max(names, key=lambda nm: per[nm]["us"])
    bpp = KERNELS[dom].bytes_per_point_f64
    achieved = bpp * pts / (per[dom]["us"] * 1e-6) / 1e9
    step_ms = total_ms / args.steps
    value = len(names) * pts / (step_ms * 1e-3) / 1e9

    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{len(names)} stencils, plain FP64, forward, {n}x{n}x{args.nz} interior, halo 3, k-halo 1",
            "stencils": names,
            "nx": n, "ny": n, "nz": args.nz,
            "points_per_stencil": pts,
            "layout": f"i-fastest, row pitch {L.sj} (interior rows 128 B aligned)",
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": _traffic(n, args.nz, dom),
            "algorithmic_bytes_per_launch": bpp * pts,
        },
        "per_stencil": per,
        "setup_s": round(t_setup, 2),
    }
    if copy_line:
        result["device_copy"] = copy_line
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(names, hosts, L, args.cpu_budget_s)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
