"""Chains of stencils: a time step written as a program of halo fills and
stencil calls over named fields, run eagerly on a stream or recorded once as
a HIP graph and replayed.

The reference couples nothing: each kernel writes only its output region
(kernels/<name>.h), and only the drivers' fillMath (include/util.h:155-163)
writes halos.  A chain maps each stencil's outputs onto the next one's inputs
by name, and refills, before each call, the halo that call reads
(oeb_halo_fill_n_*: periodic or zero-gradient in i/j, zero-gradient in k;
one launch for all fields of a refill).

The program is plain data (Halo and Call records), so the same step runs on
the GPU here and, in the tests, on the host with the oracle.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

from . import abi, stencils
from .generators import cast, generate
from .layout import Layout, Layout1D
from .spec import KERNELS
from .storage import Storage1D, Storage3D, require_device


@dataclass(frozen=True)
class Halo:
    """Refill the halo of `fields`: width h in i/j, hk in k."""

    fields: tuple[str, ...]
    h: int
    hk: int = 0
    mode: int = abi.OEB_HALO_PERIODIC


@dataclass(frozen=True)
class Call:
    """Run `stencil` with each reference argument bound to a named field.
    Scratch arguments are left unbound (NULL)."""

    stencil: str
    bind: dict
    scalars: dict = field(default_factory=dict)


def halo_fill(f: Storage3D, domain, h: int, hk: int = 0, mode: int = abi.OEB_HALO_PERIODIC, stream=None,
              check_device: bool = True) -> None:
    """Refills the halo of one field (oeb_halo_fill_*).  Host buffers raise
    TypeError (check_device=False hands them to the library, which rejects
    them: the validation tests use that)."""
    if check_device:
        require_device("halo_fill", {"field": f})
    lib = abi.load()
    nx, ny, nz = domain
    dom = abi.OebDomain(nx, ny, nz)
    fn = getattr(lib, f"oeb_halo_fill_{f.precision}")
    abi.check(fn(C.byref(dom), f.ref(), int(h), int(hk), int(mode), stencils._stream(stream)))


def halo_fill_n(fs: list, domain, h: int, hk: int = 0, mode: int = abi.OEB_HALO_PERIODIC, stream=None,
                check_device: bool = True) -> None:
    """Refills the halos of 1..8 fields of one layout and precision in one
    launch (oeb_halo_fill_n_*); the same result as halo_fill on each."""
    if check_device:
        require_device("halo_fill_n", {f"fields[{i}]": f for i, f in enumerate(fs)})
    lib = abi.load()
    nx, ny, nz = domain
    dom = abi.OebDomain(nx, ny, nz)
    precs = {f.precision for f in fs}
    if len(precs) != 1:
        raise ValueError(f"halo_fill_n: fields of one precision expected, got {sorted(precs)}")
    fn = getattr(lib, f"oeb_halo_fill_n_{precs.pop()}")
    ptrs = (abi.P3 * len(fs))(*[C.pointer(f.desc) for f in fs])
    abi.check(fn(C.byref(dom), ptrs, len(fs), int(h), int(hk), int(mode), stencils._stream(stream)))


class Graph:
    """A recorded step (oeb_graph_*).  launch() replays it on a stream."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    def launch(self, stream) -> None:
        abi.check(abi.load().oeb_graph_launch(self._h, stencils._stream(stream)))

    def close(self) -> None:
        if self._h:
            abi.check(abi.load().oeb_graph_destroy(self._h))
            self._h = C.c_void_p(0)

    def __del__(self):  # pragma: no cover - best effort at interpreter exit
        try:
            self.close()
        except Exception:
            pass


class Chain:
    """One time step (`program`) over device fields (name -> Storage3D or
    Storage1D), all of one Layout, in one precision."""

    def __init__(self, program: list, fields: dict, domain: tuple[int, int, int]):
        self.program = list(program)
        self.fields = fields
        self.domain = tuple(domain)
        for op in self.program:
            names = op.fields if isinstance(op, Halo) else tuple(op.bind.values())
            missing = [n for n in names if n not in fields]
            if missing:
                raise KeyError(f"chain: fields {missing} are not defined")

    def run_op(self, op, stream=None) -> None:
        """Enqueues one op of the program on `stream`."""
        if isinstance(op, Halo):
            halo_fill_n([self.fields[n] for n in op.fields], self.domain, op.h, op.hk, op.mode, stream)
        else:
            spec = KERNELS[op.stencil]
            arrays = [self.fields[op.bind[a.name]] if a.name in op.bind else None for a in spec.args]
            sc = dict(spec.scalars)
            sc.update(op.scalars)
            stencils.call(op.stencil, arrays, [sc[n] for n, _ in spec.scalars], self.domain, stream)

    def step(self, stream=None) -> None:
        """Enqueues one step on `stream` (default: torch's current stream)."""
        for op in self.program:
            self.run_op(op, stream)

    def capture(self, stream, steps: int = 1) -> Graph:
        """Records `steps` steps enqueued on `stream` (a created stream, not
        the null stream) as one graph.  Nothing runs until Graph.launch."""
        lib = abi.load()
        s = stencils._stream(stream)
        abi.check(lib.oeb_graph_begin(s))
        try:
            for _ in range(steps):
                self.step(stream)
        except Exception:
            g = C.c_void_p()
            lib.oeb_graph_end(s, C.byref(g))
            if g:
                lib.oeb_graph_destroy(g)
            raise
        g = C.c_void_p()
        abi.check(lib.oeb_graph_end(s, C.byref(g)))
        return Graph(g.value)


# ------------------------------------------------- the forward workload chain
# hadvuv5th -> hdiffsmag -> fastwavesuv -> uvbke, on the prognostic winds
# (u, v).  Each stencil's wind outputs are the next one's wind inputs, and
# uvbke writes the winds of the next step.  Before each call the halo its
# wind inputs are read with is refilled (periodic): reach 3 for hadvuv5th,
# 1 for hdiffsmag and uvbke; fastwavesuv reads them at the centre only.
# Every other input is a static field of that stencil's generator.
FORWARD_WINDS = {
    "hadvuv5th": ({"uin": "u", "vin": "v"}, {"uout": "u1", "vout": "v1"}),
    "hdiffsmag": ({"uin": "u1", "vin": "v1"}, {"uout": "u2", "vout": "v2"}),
    "fastwavesuv": ({"uin": "u2", "vin": "v2"}, {"uout": "u3", "vout": "v3"}),
    "uvbke": ({"uc": "u3", "vc": "v3"}, {"ub": "u", "vb": "v"}),
}
FORWARD_REACH = {"hadvuv5th": 3, "hdiffsmag": 1, "fastwavesuv": 0, "uvbke": 1}
# Scalars of the chain that differ from the drivers' values.  With the drivers'
# dt = 10 (fastwavesuv) and dt5 = 112.5 (uvbke) each step multiplies the winds
# by ~1e4 and hadvuv5th squares them, so the state overflows within a few
# steps.  With these the winds settle at O(10) m/s.
FORWARD_SCALARS = {"fastwavesuv": {"dt": 1e-3}, "uvbke": {"dt5": 9.0}}


def forward_program() -> list:
    prog = []
    for st, (ins, outs) in FORWARD_WINDS.items():
        if FORWARD_REACH[st]:
            prog.append(Halo(tuple(ins.values()), FORWARD_REACH[st]))
        bind = {a.name: f"{st}.{a.name}" for a in KERNELS[st].args if a.kind in ("in", "in1")}
        bind.update(ins)
        bind.update(outs)
        prog.append(Call(st, bind, dict(FORWARD_SCALARS.get(st, {}))))
    return prog


def forward_host_fields(L: Layout, precision: str = "f64", seed: int = 0) -> dict:
    """Host arrays of the forward chain: the winds u, v (from hadvuv5th's
    generator), the intermediates u1..v3 (NaN), and every static input as
    '<stencil>.<arg>'."""
    out = {}
    for st, (ins, outs) in FORWARD_WINDS.items():
        h = cast(st, generate(st, L, "baseline", seed), precision)
        for a in KERNELS[st].args:
            if a.kind in ("in", "in1") and a.name not in ins:
                out[f"{st}.{a.name}"] = h[a.name]
        if st == "hadvuv5th":
            out["u"], out["v"] = h["uin"], h["vin"]
        for name in outs.values():
            if name not in out:
                out[name] = h[next(k for k, v in outs.items() if v == name)]  # NaN-filled output buffer
    return out


def to_device(host: dict, L: Layout) -> dict:
    """Device copies of chain fields: arrays of L.size elements are 3-D, the
    rest 1-D over j."""
    L1 = Layout1D(L.ny, L.h)
    out = {}
    for n, a in host.items():
        if a.size == L.size:
            out[n] = Storage3D.to_device(L, a)
        elif a.size == L1.size:
            out[n] = Storage1D.to_device(L1, a)
        else:
            raise ValueError(f"chain field {n}: {a.size} elements fit neither the 3-D nor the 1-D layout")
    return out
