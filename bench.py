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

fp32: the same step in the FP32 form (oeb_<name>_f32: FP32 storage and
arithmetic, FP64 islands in fastwavesuv, p_grad_c and nh_p_grad), on the same
inputs rounded to float32.  Reported per stencil: time, speed-up over FP64,
bytes per point of the FP32 form, and the relative max / RMS error of each
output against the FP64 path's output (which equals the FP64 oracle's bit for
bit), with the fraction of points whose error exceeds 1e-5 of the output's
max-norm.

target_512x512x80: the stated target workload, the same ten-stencil step at
512x512x80 in plain FP64 and in the FP32 form (inputs generated on the
device with the baseline recipes): per stencil, µs, Gpts/s and fraction of
HBM peak (with whether it meets 0.70), the FP32 speed-up and its error
against the FP64 outputs per output field.

hbm_ceiling: a streaming device copy (oeb_stream_copy, 16-byte non-temporal
loads and stores) of 336 MB, 1 GiB and 2 GiB, the practical HBM ceiling.

config1: hdiffsa at the reference's default shape 64x64x60 (k-halo 0), GPU
and the reference's OpenMP variant on the host, µs per call.

chain: the forward chain hadvuv5th -> hdiffsmag -> fastwavesuv -> uvbke with
periodic halo refills (oeb.chain: one launch refills both winds), 10 steps
at the same size, enqueued eagerly and replayed as one recorded HIP graph;
ms per step for each, per form.

ensemble: 48 members, nx and ny drawn from {96,...,512} and nz from
{60, 80, 137}, FP32 with FP64 islands, 10 steps of the same chain: one launch
per member per op (plain), against one persistent launch per op for all
members (oeb_batch), eager and replayed from a HIP graph, first with the
built-in k-chunks, then with the autotuner's choice per (stencil, member
shape).  Gpts/s counts 4 stencils x all member points per step.

adjoint: per stencil (hdiffsa, hdiffsmag, hadvuv5th, fvtp2d, fastwavesuv), the
time of the hand-written adjoint against its forward at the same size; and the
chain hdiffsa -> hadvuv5th -> fvtp2d -> fastwavesuv (oeb.adjoint_chain), 100
steps checkpointed every 10, FP64 and FP32: forward time, forward + backward
time (the backward recomputes each segment), their ratio, and peak memory.

cpu_baseline: the reference's own kernels (OpenMP variants, serial DEFAULT for
fvtp2d, which has no OpenMP variant), compiled by oracle/Makefile, timed on
this host over a bounded sample of the same workload, with one OpenMP thread
per physical core of the process's CPUs (at most 16, the CPU share of one
GPU), OMP_PROC_BIND=close and OMP_PLACES=cores.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256] [--nz 80]
                  [--precision both|f64|f32]
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
    # 100 untimed steps (about 60 ms at 256x256x80): after 3 the FP64 step
    # measured 0.623-0.629 ms, after 100 or 300 0.590-0.591 ms, the same as
    # every later measurement in the process.
    p.add_argument("--warmup", type=int, default=100)
    p.add_argument("--size", type=int, default=256, help="nx = ny")
    p.add_argument("--nz", type=int, default=80)
    p.add_argument("--kernels", default="all", help="comma-separated subset (default: all ten)")
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--cpu-budget-s", type=float, default=20.0, help="wall-clock budget of the CPU baseline")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--precision", choices=("both", "f64", "f32"), default="both",
                   help="forms to time: the FP64 line plus its fp32 block (both), or one form alone")
    p.add_argument("--target-size", type=int, default=512,
                   help="nx = ny of the target block (the stated 512x512x80 workload, FP64 and FP32); 0 skips it")
    p.add_argument("--no-target", action="store_true", help="skip the target block")
    p.add_argument("--no-chain", action="store_true", help="skip the chain block (eager vs HIP-graph steps)")
    p.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble block")
    p.add_argument("--members", type=int, default=48, help="members of the ensemble block")
    p.add_argument("--no-adjoint", action="store_true", help="skip the adjoint block")
    p.add_argument("--adjoint-steps", type=int, default=100,
                   help="steps of the checkpointed adjoint chain (0: per-stencil adjoint timings only)")
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


def host_info() -> dict:
    """The host the CPU baseline ran on: CPU model, sockets and physical cores
    of the machine (from /proc/cpuinfo), logical CPUs, and the CPUs and
    physical cores this process may run on (its affinity mask)."""
    sys.path.insert(0, str(ROOT))
    from oracle import binding as B

    model, sockets, cores = None, set(), set()
    try:
        with open("/proc/cpuinfo") as f:
            phys = None
            for line in f:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                if key == "model name" and model is None:
                    model = val
                elif key == "physical id":
                    phys = val
                    sockets.add(val)
                elif key == "core id":
                    cores.add((phys, val))
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        affinity = os.cpu_count()
    return {"model": model, "sockets": len(sockets) or None, "physical_cores": len(cores) or None,
            "logical_cpus": os.cpu_count(), "affinity_cpus": affinity,
            "affinity_physical_cores": B.cpu_count_physical()}


# The CPU share of one GPU on the benchmark machines: at most 16 threads.
CPU_SHARE = 16


def _omp_env() -> dict:
    """OpenMP placement of the CPU baseline (threads on cores,
    packed).  Set before the reference's library, and with it its OpenMP
    runtime, is loaded."""
    env = {"OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}
    for k, v in env.items():
        os.environ.setdefault(k, v)
    return {k: os.environ[k] for k in env}


def cpu_threads() -> int:
    """Threads of the CPU baseline: every physical core this process may run
    on, up to the CPU share of one GPU (CPU_SHARE)."""
    sys.path.insert(0, str(ROOT))
    from oracle import binding as B

    return max(1, min(CPU_SHARE, B.cpu_count_physical()))


def cpu_baseline(names, hosts, L, budget_s: float):
    """Times the reference's kernels (or the oracle when the reference build is
    absent) on this host over whole steps of the same workload."""
    sys.path.insert(0, str(ROOT))
    from oracle import binding as B

    omp = _omp_env()
    threads = cpu_threads()
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
        "threads_rule": f"physical cores of the affinity mask, at most {CPU_SHARE} (the CPU share of one GPU)",
        "omp": omp,
        "per_stencil_ms": {n: round(1e3 * statistics.median(v), 3) for n, v in per.items()},
        "host": host_info(),
    }


def config1_block(seed: int, budget_s: float, with_cpu: bool = True, reps: int = 200):
    """hdiffsa at the reference's default shape, 64x64x60 (halo 3, k-halo 0):
    the GPU per call (HIP events around `reps` back-to-back calls) and the
    reference's OpenMP variant on this host (median of timed calls)."""
    import torch

    from oeb import stencils
    from oeb.generators import generate
    from oeb.layout import Layout, Layout1D
    from oeb.spec import KERNELS
    from oeb.storage import Storage1D, Storage3D

    L = Layout(64, 64, 60, h=3, hk=0, align=16)
    L1 = Layout1D(64, 3)
    host = generate("hdiffsa", L, "baseline", seed)
    dev = {a.name: (Storage1D.to_device(L1, host[a.name]) if a.kind == "in1" else Storage3D.to_device(L, host[a.name]))
           for a in KERNELS["hdiffsa"].args if a.kind != "scratch"}
    run = lambda: stencils.run("hdiffsa", dev, domain=(64, 64, 60))  # noqa: E731
    for _ in range(10):
        run()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    torch.cuda.synchronize()
    gpu_us = e0.elapsed_time(e1) * 1e3 / reps
    pts = 64 * 64 * 60
    out = {"workload": "hdiffsa, plain FP64, 64x64x60 interior, halo 3, k-halo 0 (the reference's default shape)",
           "gpu_us_per_call": round(gpu_us, 2), "gpu_gpts_s": round(pts / (gpu_us * 1e-6) / 1e9, 3),
           "gpu_timing": f"HIP events around {reps} back-to-back calls (inputs resident in the MALL)"}
    if with_cpu:
        sys.path.insert(0, str(ROOT))
        from oracle import binding as B

        _omp_env()
        threads = cpu_threads()
        if B.ref_available():
            B.ref_lib().ref_set_threads(threads)
            fn, kind = (lambda: B.run_ref("hdiffsa", host, L)), "reference"
        else:
            threads = 1
            fn, kind = (lambda: B.run_oracle("hdiffsa", host, L)), "port"
        ts, t_start = [], time.perf_counter()
        while len(ts) < 5 or (time.perf_counter() - t_start < budget_s and len(ts) < 200):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        cpu_us = statistics.median(ts) * 1e6
        out["cpu"] = {"us_per_call": round(cpu_us, 1), "gpts_s": round(pts / (cpu_us * 1e-6) / 1e9, 4),
                      "kind": kind, "threads": threads, "calls": len(ts)}
        out["gpu_over_cpu"] = round(cpu_us / gpu_us, 1)
    return out


class LaunchEvents:
    """HIP events created with hipEventDisableSystemFence, for timing launches
    one by one on a stream.  A default event ends with a system-scope fence:
    recorded before and after each launch of the ten-stencil step it added
    about 5 us to every launch (laplace 20.7 us against 15.6 in the rocprofv3
    kernel trace), this form about 2 us (17.5; fvtp2d 143.7 against 143.2).
    The HIP runtime is the one the process already loaded (torch's), found in
    /proc/self/maps, so the events and the stream handles come from one
    runtime."""

    _hip = None

    @classmethod
    def _runtime(cls):
        import ctypes as C

        if cls._hip is None:
            path = None
            with open("/proc/self/maps") as f:
                for line in f:
                    if "libamdhip64.so" in line:
                        path = line.split()[-1]
                        break
            if path is None:
                raise RuntimeError("the HIP runtime is not loaded in this process")
            hip = C.CDLL(path)
            hip.hipEventCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
            hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
            hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
            hip.hipEventDestroy.argtypes = [C.c_void_p]
            cls._hip = hip
        return cls._hip

    def __init__(self, n: int):
        import ctypes as C

        hip = self._runtime()
        self.ev = []
        for _ in range(n):
            e = C.c_void_p()
            rc = hip.hipEventCreateWithFlags(C.byref(e), 0x20000000)  # hipEventDisableSystemFence
            if rc != 0:
                raise RuntimeError(f"hipEventCreateWithFlags failed ({rc})")
            self.ev.append(e)

    def record(self, i: int, stream: int) -> None:
        rc = self._hip.hipEventRecord(self.ev[i], stream)
        if rc != 0:
            raise RuntimeError(f"hipEventRecord failed ({rc})")

    def elapsed_ms(self, a: int, b: int) -> float:
        import ctypes as C

        t = C.c_float()
        rc = self._hip.hipEventElapsedTime(C.byref(t), self.ev[a], self.ev[b])
        if rc != 0:
            raise RuntimeError(f"hipEventElapsedTime failed ({rc})")
        return t.value

    def close(self) -> None:
        for e in self.ev:
            self._hip.hipEventDestroy(e)
        self.ev = []


def time_steps(names, dev, run, steps: int, warmup: int):
    """Runs `warmup` untimed steps, then `steps` timed ones with events only
    around the whole timed region, then `steps` more with events around each
    launch (LaunchEvents, on the stream the launches go to).  Returns the mean
    step time (ms) of the first timed pass and each stencil's launch durations
    (ms, from the second pass).  The per-launch events add their own gaps
    between launches, so they stay out of the step time."""
    import torch

    stream = torch.cuda.current_stream().cuda_stream

    def step(ev=None, base=0):
        for i, name in enumerate(names):
            if ev is not None:
                ev.record(base + 2 * i, stream)
            run(name, dev[name])
            if ev is not None:
                ev.record(base + 2 * i + 1, stream)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    begin, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    begin.record()
    for _ in range(steps):
        step()
    end.record()
    torch.cuda.synchronize()
    step_ms = begin.elapsed_time(end) / steps
    ev = LaunchEvents(2 * len(names) * steps)
    torch.cuda.synchronize()
    for s in range(steps):
        step(ev, 2 * len(names) * s)
    torch.cuda.synchronize()
    durs = {nm: [ev.elapsed_ms(2 * len(names) * s + 2 * i, 2 * len(names) * s + 2 * i + 1) for s in range(steps)]
            for i, nm in enumerate(names)}
    ev.close()
    return step_ms, durs


def per_stencil(names, durs, pts, bytes_per_pt):
    out = {}
    for name in names:
        avg_ms = sum(durs[name]) / len(durs[name])
        bpp = bytes_per_pt[name]
        gbs = bpp * pts / (avg_ms * 1e-3) / 1e9
        out[name] = {
            "us": round(avg_ms * 1e3, 3),
            "us_median": round(statistics.median(durs[name]) * 1e3, 3),
            "gpts_s": round(pts / (avg_ms * 1e-3) / 1e9, 3),
            "bytes_per_pt": bpp,
            "gbs": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
        }
    return out


def fp32_errors(names, dev64, dev32, L) -> dict:
    """Per output: relative max and RMS error of the FP32 form's output against
    the FP64 form's (on the same inputs), and the fraction of points off by
    more than 1e-5 of the FP64 output's max-norm."""
    from oeb.spec import KERNELS

    out = {}
    for name in names:
        errs = {}
        for o in KERNELS[name].outputs:
            i0, i1, j0, j1 = KERNELS[name].region(o, L.nx, L.ny)
            x = L.region(dev32[name][o].buf, i0, i1, j0, j1).double()
            y = L.region(dev64[name][o].buf, i0, i1, j0, j1)
            diff = (x - y).abs()
            ymax = y.abs().max()
            errs[o] = {"rel_max": float(f"{float(diff.max() / ymax):.3e}"),
                       "rel_rms": float(f"{float(diff.pow(2).mean().sqrt() / y.pow(2).mean().sqrt()):.3e}"),
                       "frac_above_1e-5": float(f"{float((diff > 1e-5 * ymax).double().mean()):.3e}")}
        out[name] = errs
    return out


def deviation_fields(dev32: dict, dev64: dict, L) -> dict:
    """The FP32 step's fields with the islands of fastwavesuv, p_grad_c and
    nh_p_grad as FP32 deviations from FP64 per-level means
    (oeb_<name>_f32d), split on the device from the FP64 form's island
    arrays."""
    from oeb.islands import split_device
    from oeb.layout import Layout1D
    from oeb.spec import DEVIATION_STENCILS
    from oeb.storage import DeviationIsland, Storage1D, Storage3D

    out = {name: dict(fields) for name, fields in dev32.items()}
    for name, islands in DEVIATION_STENCILS.items():
        if name not in out:
            continue
        for a in islands:
            d, base = split_device(L, dev64[name][a].buf)
            out[name][a] = DeviationIsland(Storage3D(L, d), Storage1D(Layout1D(L.nz, L.hk), base))
    return out


def deviation_summary(names, step64, per64, step32d, per32d, errs, pts) -> dict:
    """The FP32 step with deviation islands: its rate and speed-up over FP64,
    and the three stencils that take the deviation form."""
    from oeb.spec import DEVIATION_STENCILS

    v64 = len(names) * pts / (step64 * 1e-3) / 1e9
    v32d = len(names) * pts / (step32d * 1e-3) / 1e9
    per = {}
    for name in names:
        if name in DEVIATION_STENCILS:
            per[name] = dict(per32d[name], speedup=round(per64[name]["us"] / per32d[name]["us"], 3),
                             error_vs_f64=errs[name])
    return {"value": round(v32d, 3), "unit": "Gpts/s", "ms_per_step": round(step32d, 4),
            "speedup_vs_f64": round(v32d / v64, 3),
            "islands": "fastwavesuv (ppuv, hhl), p_grad_c (gz, pkc) and nh_p_grad (gz, pp, pk3) as FP32 "
                       "deviations from FP64 level means",
            "per_stencil": per}


def target_block(n: int, nz: int, seed: int, steps: int, warmup: int) -> dict:
    """The stated target workload, n x n x nz (512 x 512 x 80): the ten-stencil
    step in plain FP64 and in the FP32 form, timed as the headline step is
    (per-stencil HIP events on the launch stream), with each FP64 stencil's
    fraction of the 8 TB/s HBM peak against the 0.70 target, and each FP32
    stencil's speed-up and error against the FP64 outputs.  Inputs are
    generated on the device (oeb.generators.generate_device: the baseline
    recipes), since host generation of 67 fields of 185 MB dominates."""
    import torch

    from oeb import stencils
    from oeb.generators import generate_device
    from oeb.layout import Layout, Layout1D
    from oeb.spec import KERNELS, ORDER
    from oeb.storage import Storage1D, Storage3D

    L = Layout(n, n, nz, h=3, hk=1, align=16)
    L1 = Layout1D(n, 3)
    pts = n * n * nz
    dev = {"f64": {}, "f32": {}}
    for name in ORDER:
        h = generate_device(name, L, seed)
        for prec in ("f64", "f32"):
            d = {}
            for a in KERNELS[name].args:
                if a.kind == "scratch":
                    continue
                t = h[a.name]
                if prec == "f32" and KERNELS[name].arg_precision(a.name, "f32") == "f32":
                    t = t.float()
                d[a.name] = Storage1D(L1, t) if a.kind == "in1" else Storage3D(L, t)
            dev[prec][name] = d
    torch.cuda.synchronize()
    run = lambda name, fields: stencils.run(name, fields, domain=(n, n, nz))  # noqa: E731
    res = {}
    for prec in ("f64", "f32"):
        step_ms, durs = time_steps(ORDER, dev[prec], run, steps, warmup)
        res[prec] = (step_ms, per_stencil(ORDER, durs, pts, {nm: KERNELS[nm].bytes_per_point(prec) for nm in ORDER}))
    (s64, p64), (s32, p32) = res["f64"], res["f32"]
    errs = fp32_errors(ORDER, dev["f64"], dev["f32"], L)
    # The deviation-island step writes the FP32 step's output arrays, so it
    # runs after their errors are taken.
    dev["f32d"] = deviation_fields(dev["f32"], dev["f64"], L)
    step_ms, durs = time_steps(ORDER, dev["f32d"], run, steps, warmup)
    s32d, p32d = step_ms, per_stencil(ORDER, durs, pts, {nm: KERNELS[nm].bytes_per_point("f32d") for nm in ORDER})
    dsum = deviation_summary(ORDER, s64, p64, s32d, p32d, fp32_errors(ORDER, dev["f64"], dev["f32d"], L), pts)
    for name in ORDER:
        p64[name]["meets_0.70"] = p64[name]["hbm_frac"] >= 0.70
        p32[name]["speedup"] = round(p64[name]["us"] / p32[name]["us"], 3)
        p32[name]["error_vs_f64"] = errs[name]
    del dev
    torch.cuda.empty_cache()
    v64, v32 = len(ORDER) * pts / (s64 * 1e-3) / 1e9, len(ORDER) * pts / (s32 * 1e-3) / 1e9
    return {
        "workload": f"10 stencils, forward, {n}x{n}x{nz} interior, halo 3, k-halo 1; FP64 and FP32 (FP64 islands)",
        "data": "synthetic, generated on the device (baseline recipes)",
        "steps": steps,
        "f64": {"value": round(v64, 3), "unit": "Gpts/s", "ms_per_step": round(s64, 4),
                "stencils_at_0.70": sum(1 for nm in ORDER if p64[nm]["meets_0.70"]), "per_stencil": p64},
        "f32": {"value": round(v32, 3), "unit": "Gpts/s", "ms_per_step": round(s32, 4),
                "speedup_vs_f64": round(v32 / v64, 3),
                "error_reference": "FP64 outputs of the same inputs (equal to the FP64 oracle's bit for bit)",
                "per_stencil": p32},
        "f32_deviation_islands": dsum,
    }


def ceiling_block(sizes=(336 << 20, 1 << 30, 2 << 30), reps: int = 10) -> dict:
    """The practical HBM ceiling: oeb_stream_copy (16-byte non-temporal loads
    and stores, a grid striding through the buffer) of `sizes` bytes, HIP
    events around `reps` back-to-back copies; GB/s counts the bytes read plus
    the bytes written."""
    import ctypes as C

    import torch

    from oeb import abi, stencils

    lib = abi.load()
    out = {}
    for nbytes in sizes:
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)

        def copy():
            abi.check(lib.oeb_stream_copy(C.c_void_p(dst.data_ptr()), C.c_void_p(src.data_ptr()), C.c_size_t(nbytes),
                                          stencils._stream(None)))

        for _ in range(3):
            copy()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            copy()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        assert bool(torch.equal(dst[:: max(1, dst.numel() // 4096)], src[:: max(1, src.numel() // 4096)]))
        gbs = 2 * nbytes / (ms * 1e-3) / 1e9
        out[f"{nbytes / 1e6:.0f}MB"] = {"us": round(ms * 1e3, 1), "gbs": round(gbs, 1),
                                        "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}
        del src, dst
    torch.cuda.empty_cache()
    return out


def chain_block(L, forms, seed, steps: int = 10, reps: int = 5):
    """ms per step of the forward chain, eager and as a replayed graph."""
    import torch

    from oeb.chain import Chain, Halo, forward_host_fields, forward_program, to_device

    prog = forward_program()
    out = {"stencils": ["hadvuv5th", "hdiffsmag", "fastwavesuv", "uvbke"],
           "halo_refills_per_step": sum(1 for op in prog if isinstance(op, Halo)),
           "fields_refilled_per_step": sum(len(op.fields) for op in prog if isinstance(op, Halo)),
           "steps_per_measurement": steps}
    side = torch.cuda.Stream()
    for prec in forms:
        ch = Chain(forward_program(), to_device(forward_host_fields(L, prec, seed), L), (L.nx, L.ny, L.nz))
        s = side.cuda_stream
        for _ in range(2):
            ch.step(s)
        side.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        eager = []
        for _ in range(reps):
            e0.record(side)
            for _ in range(steps):
                ch.step(s)
            e1.record(side)
            side.synchronize()
            eager.append(e0.elapsed_time(e1) / steps)
        g = ch.capture(s, steps)
        g.launch(s)
        side.synchronize()
        graph = []
        for _ in range(reps):
            e0.record(side)
            g.launch(s)
            e1.record(side)
            side.synchronize()
            graph.append(e0.elapsed_time(e1) / steps)
        g.close()
        u = ch.fields["u"].buf
        out[prec] = {"eager_ms_per_step": round(statistics.median(eager), 4),
                     "graph_ms_per_step": round(statistics.median(graph), 4),
                     "state_finite": bool(torch.isfinite(L.interior(u)).all())}
    return out


def ensemble_block(n_members: int, seed: int, steps: int = 10, reps: int = 3):
    """ms per step of the forward chain over an uneven ensemble: per-member
    launches against batched launches (eager and from a graph), and the
    batched graph after autotuning (the tuned knobs kept only if the step
    gets faster with them)."""
    import torch

    from oeb.ensemble import draw_shapes, forward_ensemble
    from oeb.tune import Tuner
    from oeb.tune import clear as tune_clear

    shapes = draw_shapes(n_members, seed)
    t0 = time.perf_counter()
    ens = forward_ensemble(shapes, "f32", seed)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    side = torch.cuda.Stream()
    s = side.cuda_stream
    pts = sum(nx * ny * nz for nx, ny, nz in shapes)

    def timed(fn):
        fn()
        side.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        out = []
        for _ in range(reps):
            e0.record(side)
            fn()
            e1.record(side)
            side.synchronize()
            out.append(e0.elapsed_time(e1) / steps)
        return statistics.median(out)

    def plain():
        for _ in range(steps):
            ens.step_plain(s)

    def batched_ms(batch):
        graph = ens.capture(batch, s, steps)
        ms = timed(lambda: graph.launch(s))
        graph.close()
        return ms

    plain_ms = timed(plain)
    batch = ens.record()
    records, launches = batch.info()

    def batched():
        for _ in range(steps):
            batch.launch(s)

    batch_ms = timed(batched)
    untuned_graph_ms = batched_ms(batch)
    # The persistent form: a grid of 4 workgroups per CU walks each run's block
    # list with stride gridDim.x (the grid knob of kernel "batch").
    from oeb.tune import set_knob

    set_knob("batch", None, None, "grid", 4 * 256)
    persistent_graph_ms = batched_ms(batch)
    set_knob("batch", None, None, "grid", -1)
    batch.close()

    # Autotuned knobs per (stencil, member shape), accepted only if the
    # batched graph step gets faster with them (Ensemble.autotune).
    t1 = time.perf_counter()
    tuner = Tuner(path=Path(os.environ.get("TMPDIR", "/tmp")) / "oeb_bench_tuning.json", reps=5)
    tuner.cache = {}
    tuned = ens.autotune(tuner, "f32", s, steps=steps, reps=reps)
    choices = tuned["choices"]
    tune_s = time.perf_counter() - t1
    batch = ens.record()
    graph_ms = batched_ms(batch)
    batch.close()
    tune_clear()
    finite = all(bool(torch.isfinite(m.fields["u"].layout.interior(m.fields["u"].buf)).all()) for m in ens.members)
    gpts = lambda ms: round(4 * pts / (ms * 1e-3) / 1e9, 3)  # noqa: E731
    return {
        "members": n_members,
        "shapes": [list(x) for x in shapes],
        "member_points": pts,
        "precision": "f32 (FP64 islands)",
        "steps_per_measurement": steps,
        "launches_per_step": {"plain": records, "batched": launches},
        "plain_ms_per_step": round(plain_ms, 3),
        "batched_ms_per_step": round(batch_ms, 3),
        "batched_graph_untuned_ms_per_step": round(untuned_graph_ms, 3),
        "batched_graph_persistent_ms_per_step": round(persistent_graph_ms, 3),
        "persistent_grid": 4 * 256,
        "batched_graph_ms_per_step": round(graph_ms, 3),
        "autotune": {"choices": len(choices), "seconds": round(tune_s, 2),
                     "non_default": sum(1 for v in choices.values() if v["knobs"]),
                     "knob_histogram": _knob_histogram(choices),
                     "step_default_ms": round(tuned["default_ms"], 3), "step_tuned_ms": round(tuned["tuned_ms"], 3),
                     "accepted": tuned["accepted"]},
        "plain_gpts_s": gpts(plain_ms),
        "batched_graph_gpts_s": gpts(graph_ms),
        "speedup_graph_vs_plain": round(plain_ms / graph_ms, 3),
        "setup_s": round(setup, 2),
        "state_finite": finite,
    }


def _knob_histogram(choices: dict) -> dict:
    """How often each knob value was chosen over the default."""
    h: dict = {}
    for v in choices.values():
        for k, x in v["knobs"].items():
            h[f"{k}={x}"] = h.get(f"{k}={x}", 0) + 1
    return dict(sorted(h.items()))


def adjoint_block(L, forms, seed, steps: int, segment: int = 10, reps: int = 5):
    import torch

    from oeb import autograd as AG
    from oeb.adjoint_chain import AdjointChain
    from oeb.generators import cast, generate
    from oeb.layout import Layout1D
    from oeb.spec import KERNELS
    from oeb.storage import Storage1D, Storage3D

    L1 = Layout1D(L.ny, L.h)
    dom = (L.nx, L.ny, L.nz)
    diff = {"hdiffsa": ("in",), "hdiffsmag": ("uin", "vin"), "hadvuv5th": ("uin", "vin"), "fvtp2d": ("q",),
            "fastwavesuv": ("uin", "vin", "ppuv")}
    fns = {"hdiffsa": AG.Hdiffsa, "hdiffsmag": AG.Hdiffsmag, "hadvuv5th": AG.Hadvuv5th, "fvtp2d": AG.Fvtp2d,
           "fastwavesuv": AG.Fastwavesuv}
    out = {"per_stencil": {}, "chain": {}}

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return statistics.median(ts)

    def kernel_times(name, prec, host, consts, dnames, stream, reps_graph=10):
        """Forward and adjoint entry points on preallocated buffers, each
        recorded reps_graph times into one HIP graph, timed over replays: the
        kernels alone, without autograd's allocations."""
        from oeb import stencils
        from oeb.chain import Graph
        from oeb import abi

        spec = KERNELS[name]
        fwd = dict(consts)
        for n in dnames:
            fwd[n] = Storage3D.to_device(L, host[n])
        for a in spec.args:
            if a.kind == "out":
                fwd[a.name] = Storage3D(L, torch.zeros(L.size, dtype=fwd[dnames[0]].buf.dtype, device="cuda"))
        adj = dict(fwd)
        names, kinds = abi.ADJOINTS[name]
        for n, k in zip(names, kinds):
            if k == "s" or n in adj:
                continue
            dt = torch.float64 if (k == "D" or prec == "f64") else torch.float32
            if n.endswith("_bar") and n[:-4] in spec.outputs:
                adj[n] = Storage3D(L, torch.ones(L.size, dtype=dt, device="cuda"))
            else:
                adj[n] = Storage3D(L, torch.zeros(L.size, dtype=dt, device="cuda"))
        lib = abi.load()
        sh = stencils._stream(stream.cuda_stream)

        def graph(fn):
            abi.check(lib.oeb_graph_begin(sh))
            for _ in range(reps_graph):
                fn()
            g = C.c_void_p()
            abi.check(lib.oeb_graph_end(sh, C.byref(g)))
            return Graph(g.value)

        with torch.cuda.stream(stream):
            gf = graph(lambda: stencils.run(name, fwd, None, domain=dom, stream=stream.cuda_stream))
            ga = graph(lambda: AG.adjoint(name, adj, None, dom, prec, stream=stream.cuda_stream))
            # Untimed launches of both graphs first: timed right after the
            # buffers and graphs were made, the first graph of a series ran
            # 4-7% slower than the same work measured a few milliseconds
            # later (hadvuv5th FP64 adjoint at 256x256x80: 158 against 148 us).
            for _ in range(3):
                gf.launch(stream.cuda_stream)
                ga.launch(stream.cuda_stream)
            stream.synchronize()
            res = []
            for gr in (gf, ga):
                gr.launch(stream.cuda_stream)
                ts = []
                for _ in range(reps):
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    gr.launch(stream.cuda_stream)
                    e1.record(stream)
                    stream.synchronize()
                    ts.append(e0.elapsed_time(e1) / reps_graph)
                res.append(statistics.median(ts))
            gf.close()
            ga.close()
        return res

    import ctypes as C

    kstream = torch.cuda.Stream()
    for prec in forms:
        per = {}
        for name, dnames in diff.items():
            host = cast(name, generate(name, L, "baseline", seed), prec)
            consts = {a.name: (Storage1D.to_device(L1, host[a.name]) if a.kind == "in1"
                               else Storage3D.to_device(L, host[a.name]))
                      for a in KERNELS[name].args if a.kind in ("in", "in1") and a.name not in dnames}
            xs = [torch.from_numpy(host[n]).to("cuda").requires_grad_(True) for n in dnames]
            ys = fns[name].apply(*xs, consts, L, dom)
            ys = ys if isinstance(ys, tuple) else (ys,)
            gs = [torch.ones_like(y) for y in ys]
            with torch.no_grad():
                f_ms = timed(lambda: fns[name].apply(*xs, consts, L, dom))
            b_ms = timed(lambda: torch.autograd.grad(ys, xs, gs, retain_graph=True))
            kf_ms, ka_ms = kernel_times(name, prec, host, consts, dnames, kstream)
            per[name] = {"kernel_forward_us": round(kf_ms * 1e3, 2), "kernel_adjoint_us": round(ka_ms * 1e3, 2),
                         "kernel_adjoint_over_forward": round(ka_ms / kf_ms, 3),
                         "autograd_forward_us": round(f_ms * 1e3, 2), "autograd_backward_us": round(b_ms * 1e3, 2),
                         "autograd_backward_over_forward": round(b_ms / f_ms, 3)}
            del xs, ys, gs, consts
        out["per_stencil"][prec] = per

        if steps <= 0:
            continue
        chain = AdjointChain(L, prec, seed)
        with torch.no_grad():
            fwd_ms = timed(lambda: chain.run(steps))

        def fwd_bwd():
            x = [t.clone().requires_grad_(True) for t in chain.state0]
            q, u, v = chain.run(steps, x, segment=segment)
            loss = (q * chain.interior_mask).sum() + (u * chain.interior_mask).sum() + (v * chain.interior_mask).sum()
            return torch.autograd.grad(loss, x)

        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        fb_ms = timed(fwd_bwd)
        peak = torch.cuda.max_memory_allocated() - base
        with torch.no_grad():
            q, u, v = chain.run(steps)
        out["chain"][prec] = {"steps": steps, "checkpoint_every": segment, "forward_ms": round(fwd_ms, 2),
                              "forward_backward_ms": round(fb_ms, 2),
                              "gradient_over_forward": round(fb_ms / fwd_ms, 3),
                              "peak_extra_memory_gb": round(peak / 1e9, 2),
                              "state_finite": bool(torch.isfinite(q).all() and torch.isfinite(u).all())}
        del chain
    return out


def _progress(msg: str) -> None:
    """One line on stderr as each block of the run ends (the JSON line is the
    only stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    args = parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (run with --gpus 1)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from oeb import stencils
    from oeb.generators import cast, generate
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
    forms = ["f64", "f32"] if args.precision == "both" else [args.precision]

    # Inputs: seeded host arrays, copied once to HBM; every stencil has its own
    # arrays in each form.
    t0 = time.perf_counter()
    hosts = {}
    dev = {p: {} for p in forms}
    for name in names:
        h = generate(name, L, "baseline", args.seed)
        hosts[name] = h
        for prec in forms:
            hp = cast(name, h, prec)
            d = {}
            for a in KERNELS[name].args:
                if a.kind == "scratch":
                    continue
                cls, lay = (Storage1D, L1) if a.kind == "in1" else (Storage3D, L)
                d[a.name] = cls.to_device(lay, hp[a.name])
            dev[prec][name] = d
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    _progress(f"inputs ready ({t_setup:.1f} s)")

    run = lambda name, fields: stencils.run(name, fields, domain=(n, n, args.nz))  # noqa: E731
    timed = {}
    for prec in forms:
        step_ms, durs = time_steps(names, dev[prec], run, args.steps, args.warmup)
        bpp = {nm: KERNELS[nm].bytes_per_point(prec) for nm in names}
        timed[prec] = (step_ms, per_stencil(names, durs, pts, bpp))
        _progress(f"{prec} step timed: {step_ms:.4f} ms")

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

    main_prec = forms[0]
    step_ms, per = timed[main_prec]
    dom = max(names, key=lambda nm: per[nm]["us"])
    bpp = KERNELS[dom].bytes_per_point(main_prec)
    achieved = bpp * pts / (per[dom]["us"] * 1e-6) / 1e9
    value = len(names) * pts / (step_ms * 1e-3) / 1e9
    form_text = "plain FP64" if main_prec == "f64" else "FP32 with FP64 islands"

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
        "dtype": main_prec,
        "data": "synthetic",
        "config": {
            "workload": f"{len(names)} stencils, {form_text}, forward, {n}x{n}x{args.nz} interior, halo 3, k-halo 1",
            "stencils": names,
            "nx": n, "ny": n, "nz": args.nz,
            "points_per_stencil": pts,
            "layout": f"i-fastest, row pitch {L.sj} (interior rows 64 B aligned in FP32, 128 B in FP64)",
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": _traffic(n, args.nz, dom) if main_prec == "f64" else None,
            "algorithmic_bytes_per_launch": bpp * pts,
        },
        "per_stencil": per,
        "setup_s": round(t_setup, 2),
    }

    if main_prec == "f64" and "f32" in forms:
        # FP32 outputs of the last timed step against the FP64 outputs of the same inputs.
        step32, per32 = timed["f32"]
        errs = fp32_errors(names, dev["f64"], dev["f32"], L)
        for name in names:
            per32[name]["speedup"] = round(per[name]["us"] / per32[name]["us"], 3)
            per32[name]["error_vs_f64"] = errs[name]
        value32 = len(names) * pts / (step32 * 1e-3) / 1e9
        # After the FP32 errors: this step writes the FP32 step's output arrays.
        dev32d = deviation_fields(dev["f32"], dev["f64"], L)
        step32d, durs32d = time_steps(names, dev32d, run, args.steps, args.warmup)
        per32d = per_stencil(names, durs32d, pts, {nm: KERNELS[nm].bytes_per_point("f32d") for nm in names})
        result["fp32"] = {
            "value": round(value32, 3),
            "unit": "Gpts/s",
            "ms_per_step": round(step32, 4),
            "speedup_vs_f64": round(value32 / value, 3),
            "error_reference": "FP64 outputs of the same step (equal to the FP64 oracle's bit for bit)",
            "per_stencil": per32,
            "deviation_islands": deviation_summary(names, step_ms, per, step32d, per32d,
                                                   fp32_errors(names, dev["f64"], dev32d, L), pts),
        }
        del dev32d
    if not args.no_target and args.target_size > 0:
        del dev
        torch.cuda.empty_cache()
        result[f"target_{args.target_size}x{args.target_size}x{args.nz}"] = target_block(
            args.target_size, args.nz, args.seed, args.steps, args.warmup)
        _progress("target block done")
    result["config1"] = config1_block(args.seed, 5.0, with_cpu=not args.no_cpu_baseline)
    result["hbm_ceiling"] = ceiling_block()
    _progress("config1 and copy-ceiling blocks done")
    if not args.no_chain:
        result["chain"] = chain_block(L, forms, args.seed)
        _progress("chain block done")
    if not args.no_ensemble and args.members > 0:
        result["ensemble"] = ensemble_block(args.members, args.seed)
        _progress("ensemble block done")
    if not args.no_adjoint:
        result["adjoint"] = adjoint_block(L, forms, args.seed, args.adjoint_steps)
        _progress("adjoint block done")
    if copy_line:
        result["device_copy"] = copy_line
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(names, hosts, L, args.cpu_budget_s)
        _progress("cpu baseline done")
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
