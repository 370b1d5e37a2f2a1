"""Autotuner: picks each stencil's k-chunk (levels per workgroup) and vector
path per (kernel, precision, domain) by timing the candidates on the GPU,
and caches the choices in a JSON file.

The library looks the choice up on every call (oeb_tune_set / the tuning
table, after the OEB_KCHUNK_<K> / OEB_PAIRS_<K> environment overrides).
Tile shapes are compiled into the kernels; what varies at run time is the
k-chunk, which trades parallelism (more workgroups) against reuse of the
vertical neighbours carried in registers, and the 16-byte vector path.

Timing: before every timed launch a buffer larger than the 256 MB MALL is
overwritten, so each candidate reads its inputs from HBM, as it does inside
a chain.  The median of `reps` launches decides, and a candidate replaces
the library's default only when it is at least MARGIN faster; a choice of
(kchunk 0, vector -1) means "the default".

The reference has no tuner: its drivers fix every loop order at compile
time.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import statistics
from pathlib import Path

from . import abi, stencils
from .spec import KERNELS

KCHUNKS = (1, 2, 4, 8, 16, 32)
# Stencils with a 16-byte vector path that the host may switch.
VECTOR_KERNELS = {"laplace", "uvbke", "fastwavesuv", "p_grad_c", "nh_p_grad"}
FLUSH_BYTES = 512 << 20
# A candidate replaces the library's default only when it is this much faster.
MARGIN = 0.05


def cache_path() -> Path:
    p = os.environ.get("OEB_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "oeb" / "tuning.json"


def _esize(precision: str) -> int:
    return 8 if precision == "f64" else 4


def set_choice(kernel: str, precision: str, domain, kchunk: int = 0, vector: int = -1) -> None:
    """Puts (kernel, precision, domain) -> (kchunk, vector) in the library's
    table; kchunk 0 and vector -1 leave that knob at its default."""
    nx, ny, nz = domain
    abi.check(abi.load().oeb_tune_set(kernel.encode(), _esize(precision), C.byref(abi.OebDomain(nx, ny, nz)),
                                      int(kchunk), int(vector)))


def get_choice(kernel: str, precision: str, domain) -> tuple[int, int]:
    nx, ny, nz = domain
    kc, vec = C.c_int(), C.c_int()
    abi.check(abi.load().oeb_tune_get(kernel.encode(), _esize(precision), C.byref(abi.OebDomain(nx, ny, nz)),
                                      C.byref(kc), C.byref(vec)))
    return kc.value, vec.value


def clear() -> None:
    abi.check(abi.load().oeb_tune_clear())


def _key(kernel: str, precision: str, domain, device: str) -> str:
    nx, ny, nz = domain
    return f"{device}|{kernel}|{precision}|{nx}x{ny}x{nz}"


def device_name() -> str:
    import torch

    p = torch.cuda.get_device_properties(0)
    return getattr(p, "gcnArchName", None) or p.name


class Tuner:
    """Times candidates and keeps the choices, in memory and in the cache."""

    def __init__(self, path: Path | None = None, reps: int = 3):
        self.path = Path(path) if path is not None else cache_path()
        self.reps = reps
        self.cache: dict = {}
        if self.path.exists():
            try:
                self.cache = json.loads(self.path.read_text())
            except ValueError:
                self.cache = {}
        self._flush = None

    def apply_cached(self) -> int:
        """Loads every cached choice for this device into the library."""
        dev = device_name()
        n = 0
        for key, v in self.cache.items():
            d, kernel, precision, shape = key.split("|")
            if d != dev:
                continue
            set_choice(kernel, precision, tuple(int(x) for x in shape.split("x")), v["kchunk"], v["vector"])
            n += 1
        return n

    def save(self) -> None:
        self.path.parent.mkdir(parents=True, exist_ok=True)
        self.path.write_text(json.dumps(self.cache, indent=1, sort_keys=True))

    def _time(self, run, stream) -> float:
        import torch

        if self._flush is None:
            self._flush = torch.empty(FLUSH_BYTES // 4, dtype=torch.float32, device="cuda")
        s = torch.cuda.ExternalStream(stream) if stream else torch.cuda.current_stream()
        out = []
        for _ in range(self.reps):
            with torch.cuda.stream(s):
                self._flush.zero_()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(s)
                run()
                e1.record(s)
            s.synchronize()
            out.append(e0.elapsed_time(e1))
        return statistics.median(out)

    def tune(self, kernel: str, fields: dict, domain, precision: str, stream=None, scalars: dict | None = None,
             force: bool = False) -> dict:
        """Chooses (kchunk, vector) for `kernel` on these device fields (its
        outputs are overwritten).  A cached choice is applied without timing
        unless `force`."""
        dev = device_name()
        key = _key(kernel, precision, domain, dev)
        if key in self.cache and not force:
            v = self.cache[key]
            set_choice(kernel, precision, domain, v["kchunk"], v["vector"])
            return v
        nz = domain[2]
        vectors = (1, 0) if kernel in VECTOR_KERNELS else (-1,)
        cands = [(kc, vec) for kc in KCHUNKS if kc <= max(1, nz) for vec in vectors]
        run = lambda: stencils.run(kernel, fields, scalars, domain=domain, stream=stream)  # noqa: E731
        timings = {}
        # The library's own choice first (kchunk 0, vector -1: no table entry).
        for kc, vec in [(0, -1)] + cands:
            set_choice(kernel, precision, domain, kc, vec)
            run()
            timings[(kc, vec)] = self._time(run, stream)
        (kc, vec), ms = min(timings.items(), key=lambda kv: kv[1])
        if ms > (1.0 - MARGIN) * timings[(0, -1)]:
            # Not clearly faster than the default: launches of 10-100 us time
            # within a few percent of each other from run to run, and a
            # choice made on that noise does not carry over to other contexts
            # (a batched launch over many members).
            (kc, vec), ms = (0, -1), timings[(0, -1)]
        set_choice(kernel, precision, domain, kc, vec)
        v = {"kchunk": kc, "vector": vec, "ms": round(ms, 5),
             "candidates": {f"{a}/{b}": round(t, 5) for (a, b), t in timings.items()}}
        self.cache[key] = v
        return v
