"""Ensembles: members of uneven (nx, ny, nz) advanced through one chain
program together.

A plain step launches every op once per member.  A batched step records the
same calls into an oeb_batch, member after member within each op, and
replays each op for all members as one persistent launch over their joint
tile list (oeb_batch_*).  A recorded batch can itself be captured in an
oeb_graph with several steps.

The reference has no ensembles: its kernels take one square domain
(domain_size for both i and j, src/<name>.cpp:4-5).  Members here may be
rectangular.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import abi, stencils
from .chain import FORWARD_WINDS, Call, Chain, Graph, forward_program
from .layout import Layout, Layout1D
from .spec import KERNELS
from .storage import Storage1D, Storage3D

# Stencils with a batched form (oeb_batch_record); the others return
# OEB_ERR_ARG while a batch records.  Halo refills are batched too.
BATCHABLE = frozenset({"hadvuv", "hadvuv5th", "hdiffsmag", "fastwavesuv", "uvbke"})

# Member extents of the ensemble workload.
SIZES_IJ = (96, 128, 192, 256, 384, 512)
SIZES_K = (60, 80, 137)


def draw_shapes(n: int, seed: int = 0) -> list[tuple[int, int, int]]:
    """n member shapes (nx, ny, nz), nx and ny drawn independently from
    SIZES_IJ and nz from SIZES_K."""
    rng = np.random.default_rng(np.random.SeedSequence([seed, 4242]))
    return [(int(rng.choice(SIZES_IJ)), int(rng.choice(SIZES_IJ)), int(rng.choice(SIZES_K))) for _ in range(n)]


class Batch:
    """An oeb_batch: launches recorded for many members, replayed as one
    persistent launch per run of one kernel."""

    def __init__(self):
        h = C.c_void_p()
        abi.check(abi.load().oeb_batch_create(C.byref(h)))
        self._h = h

    def __enter__(self):
        abi.check(abi.load().oeb_batch_record(self._h))
        return self

    def __exit__(self, exc_type, exc, tb):
        abi.check(abi.load().oeb_batch_stop(self._h))
        return False

    def launch(self, stream=None) -> None:
        abi.check(abi.load().oeb_batch_launch(self._h, stencils._stream(stream)))

    def info(self) -> tuple[int, int]:
        r, n = C.c_int(), C.c_int()
        abi.check(abi.load().oeb_batch_info(self._h, C.byref(r), C.byref(n)))
        return r.value, n.value

    def close(self) -> None:
        if self._h:
            abi.check(abi.load().oeb_batch_destroy(self._h))
            self._h = C.c_void_p(0)

    def __del__(self):  # pragma: no cover - best effort at interpreter exit
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Members (Chains with one program) stepped together."""

    def __init__(self, members: list[Chain]):
        if not members:
            raise ValueError("an ensemble needs at least one member")
        prog = members[0].program
        if any(m.program != prog for m in members):
            raise ValueError("all members must run the same program")
        self.members = members
        self.program = prog

    def step_plain(self, stream=None) -> None:
        """One step with one launch per member per op."""
        for op in self.program:
            for m in self.members:
                m.run_op(op, stream)

    def record(self, steps: int = 1) -> Batch:
        """Records `steps` steps; the returned batch runs them with one
        launch per op (per step)."""
        bad = sorted({op.stencil for op in self.program if isinstance(op, Call) and op.stencil not in BATCHABLE})
        if bad:
            raise ValueError(f"stencils without a batched form cannot be recorded: {', '.join(bad)} "
                             f"(batched: {', '.join(sorted(BATCHABLE))})")
        b = Batch()
        with b:
            for _ in range(steps):
                self.step_plain()
        return b

    def capture(self, batch: Batch, stream, steps: int = 1) -> Graph:
        """Records `steps` launches of `batch` on `stream` as one graph."""
        lib = abi.load()
        s = stencils._stream(stream)
        abi.check(lib.oeb_graph_begin(s))
        try:
            for _ in range(steps):
                batch.launch(stream)
        finally:
            g = C.c_void_p()
            rc = lib.oeb_graph_end(s, C.byref(g))
        abi.check(rc)
        return Graph(g.value)


    def graph_step_ms(self, stream, steps: int = 10, reps: int = 3) -> float:
        """Median ms per step of the batched step replayed from a graph of
        `steps` launches (the members' state advances)."""
        import statistics

        import torch

        batch = self.record()
        try:
            graph = self.capture(batch, stream, steps)
            try:
                s = torch.cuda.ExternalStream(stream)
                graph.launch(stream)
                s.synchronize()
                out = []
                for _ in range(reps):
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(s)
                    graph.launch(stream)
                    e1.record(s)
                    s.synchronize()
                    out.append(e0.elapsed_time(e1) / steps)
                return statistics.median(out)
            finally:
                graph.close()
        finally:
            batch.close()

    def autotune(self, tuner, precision: str, stream, steps: int = 10, reps: int = 3, margin: float = 0.02) -> dict:
        """tune(), then a check in context: the batched step, replayed from a
        graph, is timed with the built-in defaults and with the tuned table,
        and the tuned choices stay only if the step is at least `margin`
        faster; otherwise they are removed from the library's table.  Each
        candidate is timed alone (with the MALL flushed), while the step runs
        every member's blocks in one launch per op, where a choice that wins
        alone can lose."""
        from .tune import set_knobs

        default_ms = self.graph_step_ms(stream, steps, reps)
        choices = self.tune(tuner, precision, stream)
        tuned_ms = self.graph_step_ms(stream, steps, reps)
        accepted = tuned_ms < (1.0 - margin) * default_ms
        if not accepted:
            for stencil, domain in choices:
                set_knobs(stencil, precision, domain, {})
        return {"choices": choices, "default_ms": default_ms, "tuned_ms": tuned_ms, "accepted": accepted}

    def tune(self, tuner, precision: str, stream=None) -> dict:
        """Runs the autotuner (oeb.tune.Tuner) for every stencil of the
        program on every member shape (each (stencil, shape) once).  Outputs
        go to temporary buffers, so the members' state is untouched."""
        import torch

        out = {}
        for m in self.members:
            for op in self.program:
                if not isinstance(op, Call) or (op.stencil, m.domain) in out:
                    continue
                fields = {}
                for a in KERNELS[op.stencil].args:
                    if a.name not in op.bind:
                        continue
                    f = m.fields[op.bind[a.name]]
                    if a.kind == "out":
                        f = Storage3D(f.layout, torch.empty_like(f.buf))
                    fields[a.name] = f
                out[(op.stencil, m.domain)] = tuner.tune(op.stencil, fields, m.domain, precision, stream, op.scalars)
        return out


# ----------------------------------------- device-side fields of the forward chain
def _torch_fields(L: Layout, precision: str, seed: int, device="cuda") -> dict:
    """Fields of the forward chain generated on the device with torch, with
    the distributions of oeb.generators' baseline family (winds N(0, 10 m/s)
    box-filtered, coefficients U(0.01, 0.1), pressure 1e5 exp(-z/8000) with
    1% column noise, smooth-terrain hhl, cos(latitude) metric terms).  Used
    for large ensembles, where the host generators would dominate set-up."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    fdt = torch.float32 if precision == "f32" else torch.float64
    shape = (L.levels, L.rows, L.sj)

    def uni(lo, hi, dtype=fdt):
        return (torch.rand(shape, generator=g, device=device, dtype=torch.float64) * (hi - lo) + lo).to(dtype)

    def wind():
        n = torch.randn(shape, generator=g, device=device, dtype=torch.float64) * 10.0
        w = (n + n.roll(1, 2) + n.roll(-1, 2) + n.roll(1, 1) + n.roll(-1, 1)) / 5.0
        return w.to(fdt)

    k = torch.arange(-L.hk, L.nz + L.hk, device=device, dtype=torch.float64)
    z = 250.0 * k

    def pressure():
        eps = torch.randn((L.rows, L.sj), generator=g, device=device, dtype=torch.float64) * 0.01
        return (1e5 * torch.exp(-z / 8000.0))[:, None, None] * (1.0 + eps)[None]

    def hhl():
        jj = torch.arange(L.rows, device=device, dtype=torch.float64)[:, None] - L.h
        ii = torch.arange(L.sj, device=device, dtype=torch.float64)[None, :] - L.lead
        p1, p2 = (torch.rand(2, generator=g, device=device, dtype=torch.float64) * 2 * math.pi).tolist()
        t = 0.5 + 0.25 * (torch.sin(2 * math.pi * ii / L.nx + p1) + torch.sin(2 * math.pi * jj / L.ny + p2))
        return 250.0 * k[:, None, None] * (1.0 + 0.1 * t)[None] + 500.0 * t[None]

    L1 = Layout1D(L.ny, L.h)
    j1 = torch.arange(-L.h, L.ny + L.h, device=device, dtype=torch.float64)

    def coslat():
        step = 160.0 / (L.ny - 1) if L.ny > 1 else 0.0
        return torch.cos(torch.deg2rad(-80.0 + step * j1)).to(fdt)

    def coef1():
        return (torch.rand(L1.size, generator=g, device=device, dtype=torch.float64) * 0.09 + 0.01).to(fdt)

    recipes = {
        "hadvuv5th": {"acrlat0": coef1, "acrlat1": coef1, "tgrlatda0": coef1, "tgrlatda1": coef1},
        "hdiffsmag": {"mask": lambda: uni(0.01, 0.1), "crlavo": coslat, "crlavu": coslat, "crlato": coslat,
                      "crlatu": coslat, "acrlat0": coef1},
        "fastwavesuv": {"utens": lambda: uni(0.01, 0.1), "vtens": lambda: uni(0.01, 0.1),
                        "wgtfac": lambda: uni(0.01, 0.1), "rho": lambda: uni(0.01, 0.1), "fx": coef1,
                        "ppuv": pressure, "hhl": hhl},
        "uvbke": {"cosa": lambda: uni(0.01, 0.1), "rsina": lambda: uni(0.01, 0.1)},
    }
    out = {}
    for st, rec in recipes.items():
        for name, make in rec.items():
            a = make()
            if precision == "f32" and name in KERNELS[st].islands:
                a = a.to(torch.float64)
            out[f"{st}.{name}"] = a.reshape(-1).contiguous()
    out["u"], out["v"] = wind().reshape(-1).contiguous(), wind().reshape(-1).contiguous()
    for outs in (o for _, o in FORWARD_WINDS.values()):
        for name in outs.values():
            if name not in out:
                out[name] = torch.full((L.size,), float("nan"), device=device, dtype=fdt)
    return {n: (Storage1D(L1, a) if a.numel() == L1.size else Storage3D(L, a)) for n, a in out.items()}


def forward_ensemble(shapes, precision: str = "f32", seed: int = 0) -> Ensemble:
    """Members of the forward chain (oeb.chain.forward_program) with the given
    (nx, ny, nz) shapes, fields generated on the device."""
    members = []
    for m, (nx, ny, nz) in enumerate(shapes):
        L = Layout(nx, ny, nz, h=3, hk=1)
        members.append(Chain(forward_program(), _torch_fields(L, precision, seed * 1000 + m), (nx, ny, nz)))
    return Ensemble(members)
