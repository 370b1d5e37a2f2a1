"""Registry of the ten stencils: argument lists in the reference's order, driver
scalars, halo needs, output regions and algorithmic bytes per point.

Argument order and names follow the reference signatures (kernels/<name>.h,
cited per entry).  Kinds:
  out      3-D output, written on its region
  in       3-D input
  in1      1-D input over j
  scratch  a reference temporary, accepted and never touched (passed as NULL)
Scalar defaults are the values the reference drivers use (src/<name>.cpp).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

EDDLAT = math.ldexp(1.0, -11)  # src/hdiffsmag.cpp:48, src/hadvuv.cpp:53
EDDLON = math.ldexp(1.5, -11)  # src/hdiffsmag.cpp:49, src/hadvuv.cpp:54


@dataclass(frozen=True)
class Arg:
    name: str
    kind: str  # 'out' | 'in' | 'in1' | 'scratch'


@dataclass(frozen=True)
class Kernel:
    name: str
    args: tuple[Arg, ...]
    scalars: tuple[tuple[str, float], ...]
    hk: int  # k-halo the stencil reads (0 or 1)
    h: int  # i/j halo the stencil reads
    # output name -> (i0, i1, j0, j1) offsets: region is [i0, nx+i1) x [j0, ny+j1)
    regions: dict = field(default_factory=dict)
    ref: str = ""

    @property
    def outputs(self) -> list[str]:
        return [a.name for a in self.args if a.kind == "out"]

    @property
    def inputs3(self) -> list[str]:
        return [a.name for a in self.args if a.kind == "in"]

    @property
    def inputs1(self) -> list[str]:
        return [a.name for a in self.args if a.kind == "in1"]

    @property
    def bytes_per_point_f64(self) -> int:
        """Algorithmic HBM bytes per interior point: each 3-D input read once and
        each 3-D output written once, 8 B each (1-D arrays and halos excluded)."""
        return 8 * (len(self.inputs3) + len(self.outputs))

    @property
    def islands(self) -> frozenset[str]:
        """3-D inputs that stay FP64 in the FP32 form (oeb_<name>_f32)."""
        return ISLANDS.get(self.name, frozenset())

    def bytes_per_point(self, precision: str = "f64") -> int:
        """Algorithmic bytes per interior point in the given form: FP32 arrays
        count 4 B, FP64 (and FP64 island) arrays 8 B; in "f32d" (the
        deviation-island form, oeb_<name>_f32d) the islands count 4 B (their
        per-level base profiles are negligible)."""
        if precision == "f64":
            return self.bytes_per_point_f64
        n3 = self.inputs3 + self.outputs
        wide = self.islands if precision == "f32" else self.islands - set(DEVIATION_STENCILS.get(self.name, ()))
        return sum(8 if a in wide else 4 for a in n3)

    def arg_precision(self, arg: str, precision: str) -> str:
        """Element type of array `arg` in the given form."""
        return "f64" if precision == "f64" or arg in self.islands else "f32"

    def region(self, out: str, nx: int, ny: int) -> tuple[int, int, int, int]:
        i0, i1, j0, j1 = self.regions.get(out, (0, 0, 0, 0))
        return (i0, nx + i1, j0, ny + j1)


# FP64 islands of the FP32 form: pressure and geopotential (and, in
# fastwavesuv, the half-level heights) whose small differences cancel.
ISLANDS = {
    "fastwavesuv": frozenset({"ppuv", "hhl"}),
    "p_grad_c": frozenset({"gz", "pkc"}),
    "nh_p_grad": frozenset({"gz", "pp", "pk3"}),
}
# Stencils with an _f32d entry point, and the islands each takes as FP32
# deviations from an FP64 per-level base profile (oeb.islands).
DEVIATION_STENCILS = {"fastwavesuv": ("ppuv", "hhl"), "p_grad_c": ("gz", "pkc"), "nh_p_grad": ("gz", "pp", "pk3")}


def _a(spec: str) -> tuple[Arg, ...]:
    """'out:ub in:uc ...' -> Args."""
    out = []
    for tok in spec.split():
        kind, name = tok.split(":")
        out.append(Arg(name, {"o": "out", "i": "in", "1": "in1", "s": "scratch"}[kind]))
    return tuple(out)


KERNELS: dict[str, Kernel] = {}


def _reg(k: Kernel) -> None:
    KERNELS[k.name] = k


_reg(Kernel("laplace", _a("i:in o:out"), (), hk=0, h=1, ref="kernels/laplace.h:4"))
_reg(Kernel("hdiffsa", _a("o:out i:in i:mask 1:crlato 1:crlatu s:lap s:flx s:fly"), (), hk=0, h=2,
            ref="kernels/hdiffsa.h:4-5"))
_reg(Kernel("hdiffsmag",
            _a("o:uout o:vout i:uin i:vin i:mask 1:crlavo 1:crlavu 1:crlato 1:crlatu 1:acrlat0 s:T_sqr_s s:S_sqr_uv"),
            (("eddlat", EDDLAT), ("eddlon", EDDLON), ("tau_smag", 0.025), ("weight_smag", 0.01)),
            hk=0, h=1, ref="kernels/hdiffsmag.h:4-7"))
_HADV = ("o:uout o:vout i:uin i:vin 1:acrlat0 1:acrlat1 1:tgrlatda0 1:tgrlatda1 s:uatupos s:vatupos s:uatvpos "
         "s:vatvpos s:uavg s:vavg s:ures s:vres")
_reg(Kernel("hadvuv", _a(_HADV), (("eddlat", EDDLAT), ("eddlon", EDDLON)), hk=0, h=2, ref="kernels/hadvuv.h:29-32"))
_reg(Kernel("hadvuv5th", _a(_HADV), (("eddlat", EDDLAT), ("eddlon", EDDLON)), hk=0, h=3,
            ref="kernels/hadvuv5th.h:41-44"))
_reg(Kernel("fastwavesuv",
            _a("o:uout o:vout i:uin i:vin i:utens i:vtens i:wgtfac i:ppuv i:hhl i:rho 1:fx s:ppgk s:ppgc s:ppgu "
               "s:ppgv"),
            (("edadlat", math.ldexp(1.0, -11)), ("dt", 10.0)), hk=1, h=1, ref="kernels/fastwavesuv.h:4-7"))
_reg(Kernel("uvbke", _a("o:ub o:vb i:uc i:vc i:cosa i:rsina"), (("dt5", 0.5 * 225.0),), hk=0, h=1,
            ref="kernels/uvbke.h:4-5"))
_reg(Kernel("p_grad_c", _a("o:uout o:vout i:uin i:vin i:rdxc i:rdyc i:delpc i:gz i:pkc s:wk"), (("dt2", 0.1),),
            hk=1, h=1, ref="kernels/p_grad_c.h:4-6"))
_reg(Kernel("nh_p_grad", _a("o:uout o:vout i:uin i:vin i:rdx i:rdy i:gz i:pp i:pk3 i:wk1 s:wk s:du s:dv"),
            (("dt", 0.1),), hk=1, h=1, ref="kernels/nh_p_grad.h:4-6"))
_reg(Kernel("fvtp2d",
            _a("o:q_i o:q_j o:fx1 o:fx2 o:fy1 o:fy2 i:q i:crx i:cry i:ra_x i:ra_y i:xfx i:yfx i:area s:fxx s:fyy "
               "s:al s:almq s:br s:b0 s:smt5"),
            (("p1", 0.583333), ("p2", 0.083333)), hk=0, h=3,
            regions={"q_i": (-3, 2, 0, 0), "q_j": (0, 0, -3, 2), "fx2": (0, 1, -3, 2), "fy2": (-3, 2, 0, 1)},
            ref="kernels/fvtp2d.h:4-7"))

ORDER = ["laplace", "hdiffsa", "hdiffsmag", "hadvuv", "hadvuv5th", "fastwavesuv", "uvbke", "p_grad_c", "nh_p_grad",
         "fvtp2d"]
assert sorted(ORDER) == sorted(KERNELS)
