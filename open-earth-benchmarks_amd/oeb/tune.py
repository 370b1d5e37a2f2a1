"""Autotuner: picks each stencil's launch parameters ("knobs") per (kernel,
precision, domain) by timing candidates on the GPU, and caches the choices
in a JSON file.

The library looks its knobs up on every call in a process-wide table
(oeb_tune_knob_set); an unset knob takes the kernel's built-in default.
Knobs (include/oeb.h): form (row march or tiles), tile (tile shape), order
(the tile -> XCD mapping), kchunk (levels per workgroup), rows and depth
(row chunk and prefetch of a row march), lanes (workgroup size) and vector
(16-byte bodies).

Search.  A kernel's knob space is a list of branches: a branch fixes the
knobs that select one form of the kernel and lists the knobs of that form
with their candidate values.  Within a branch the tuner descends one knob at
a time (each value with the other knobs at their best so far), in the order
listed.  The best of all branches replaces the built-in default only when it
is at least MARGIN faster.

Timing: before every timed launch a buffer larger than the 256 MB MALL is
overwritten, so each candidate reads its inputs from HBM, as it does inside
a chain.  The median of `reps` launches decides.

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

KNOBS = ("kchunk", "vector", "form", "rows", "depth", "order", "tile", "lanes", "grid")
KCHUNKS = (1, 2, 4, 8, 16)
ORDERS = (0, 1, 2, 3)
ROWS = (8, 12, 16, 24, 32)
FLUSH_BYTES = 512 << 20
# A candidate replaces the library's default only when it is this much faster.
MARGIN = 0.05


def _tiled(extra=()):
    return [({}, [("kchunk", KCHUNKS), ("order", ORDERS)] + list(extra))]


# Branches of each kernel's knob space: (fixed knobs, [(knob, candidates)]).
SPACES = {
    "laplace": [({"form": 0}, [("vector", (1, 0)), ("kchunk", KCHUNKS), ("order", ORDERS)]),
                ({"form": 1}, [("vector", (1, 0)), ("rows", ROWS), ("depth", (1, 2, 4))])],
    "hdiffsa": [({"form": 1}, [("rows", ROWS), ("depth", (2, 4))]),
                ({"form": 0}, [("kchunk", KCHUNKS)])],
    "hdiffsmag": [({"form": 1}, [("vector", (1, 0)), ("rows", ROWS), ("depth", (1, 2))]),
                  ({"form": 0}, [("kchunk", KCHUNKS)])],
    "hadvuv": _tiled([("lanes", (1024, 512, 256))]),
    "hadvuv5th": _tiled([("lanes", (1024, 512, 256))]),
    "fastwavesuv": _tiled([("vector", (1, 0)), ("tile", (1, 0))]),
    "uvbke": _tiled([("vector", (1, 0))]),
    "p_grad_c": _tiled([("vector", (1, 0))]),
    "nh_p_grad": _tiled([("vector", (1, 0))]),
    "fvtp2d": [({"form": 0}, [("tile", (14, 13, 16, 15, 17, 8, 3)), ("order", (2, 3, 0))]),
               ({"form": 1}, [("lanes", (512, 256, 1024)), ("rows", ROWS), ("depth", (1, 2))]),
               ({"form": 2}, [("tile", (0, 1, 2)), ("order", (2, 3, 0)), ("kchunk", (1, 2))])],
}
# Kernels with a 16-byte vector path (kept for oeb_tune_set callers).
VECTOR_KERNELS = {"laplace", "uvbke", "fastwavesuv", "p_grad_c", "nh_p_grad", "hdiffsmag"}


def cache_path() -> Path:
    p = os.environ.get("OEB_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "oeb" / "tuning.json"


def _esize(precision: str | None) -> int:
    return {None: 0, "f64": 8, "f32": 4}[precision]


def _dom(domain):
    return None if domain is None else C.byref(abi.OebDomain(*domain))


def set_knob(kernel: str, precision: str | None, domain, knob: str, value: int) -> None:
    """Puts knob = value for (kernel, precision, domain) in the library's
    table; value -1 removes it.  precision None stands for both, domain None
    for every domain."""
    abi.check(abi.load().oeb_tune_knob_set(kernel.encode(), _esize(precision), _dom(domain), knob.encode(),
                                           int(value)))


def get_knob(kernel: str, precision: str | None, domain, knob: str) -> int:
    v = C.c_int()
    abi.check(abi.load().oeb_tune_knob_get(kernel.encode(), _esize(precision), _dom(domain), knob.encode(),
                                           C.byref(v)))
    return v.value


def set_knobs(kernel: str, precision: str | None, domain, knobs: dict) -> None:
    """Sets every knob of `knobs` and unsets the others for this key."""
    for k in KNOBS:
        set_knob(kernel, precision, domain, k, knobs.get(k, -1))


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


def _label(knobs: dict) -> str:
    return ",".join(f"{k}={knobs[k]}" for k in sorted(knobs)) or "default"


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
            set_knobs(kernel, precision, tuple(int(x) for x in shape.split("x")), v["knobs"])
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

    def search(self, kernel: str, run, domain, precision: str, stream=None, space=None) -> dict:
        """Times the default and the candidates of `space` (SPACES[kernel] by
        default) for `run`, a callable that launches the kernel once, and
        leaves the winner in the library's table.  Returns the choice."""
        space = SPACES.get(kernel, []) if space is None else space
        timings: dict[str, float] = {}

        def measure(knobs: dict) -> float:
            lab = _label(knobs)
            if lab not in timings:
                set_knobs(kernel, precision, domain, knobs)
                run()
                timings[lab] = self._time(run, stream)
            return timings[lab]

        base = measure({})
        best, best_ms = {}, base
        for fixed, knobs in space:
            cur = dict(fixed)
            cur_ms = measure(cur)
            for knob, values in knobs:
                for v in values:
                    cand = dict(cur, **{knob: v})
                    ms = measure(cand)
                    if ms < cur_ms:
                        cur, cur_ms = cand, ms
            if cur_ms < best_ms:
                best, best_ms = cur, cur_ms
        if best_ms > (1.0 - MARGIN) * base:
            # Not clearly faster than the default: launches of 10-100 us time
            # within a few percent of each other from run to run, and a
            # choice made on that noise does not carry over to other contexts
            # (a batched launch over many members).
            best, best_ms = {}, base
        set_knobs(kernel, precision, domain, best)
        return {"knobs": best, "ms": round(best_ms, 5), "default_ms": round(base, 5),
                "candidates": {k: round(t, 5) for k, t in timings.items()}}

    def tune(self, kernel: str, fields: dict, domain, precision: str, stream=None, scalars: dict | None = None,
             force: bool = False) -> dict:
        """Chooses the knobs of `kernel` on these device fields (its outputs
        are overwritten).  A cached choice is applied without timing unless
        `force`."""
        dev = device_name()
        key = _key(kernel, precision, domain, dev)
        if key in self.cache and not force:
            v = self.cache[key]
            set_knobs(kernel, precision, domain, v["knobs"])
            return v
        run = lambda: stencils.run(kernel, fields, scalars, domain=domain, stream=stream)  # noqa: E731
        v = self.search(kernel, run, domain, precision, stream)
        self.cache[key] = v
        return v
