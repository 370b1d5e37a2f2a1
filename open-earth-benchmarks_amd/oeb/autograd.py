"""torch.autograd.Function wrappers: the HIP forward stencils with the
hand-written HIP adjoints (oeb_<name>_adj_*) as their backward.

Each function takes the differentiated fields as flat device tensors in a
Layout (the allocation, halo and padding included) and the constant fields as
a dict of Storage objects.  Outputs are flat tensors of the same Layout,
zero outside each output's region.  Gradients come back on the whole
allocation: the adjoint writes the box the forward reads, the rest is zero.
Limiter, upwind, clamp and PPM branches are frozen at the forward choice.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import abi, stencils
from .layout import Layout
from .spec import KERNELS
from .storage import Storage3D, require_device


def _adjoint(name: str, fields: dict, scalars: dict, domain, precision: str, stream=None) -> None:
    """Calls oeb_<name>_adj_<precision> with arrays looked up by argument name
    (abi.ADJOINTS) on `stream` (default: torch's current stream)."""
    names, kinds = abi.ADJOINTS[name]
    require_device(f"{name} adjoint", {n: fields[n] for n, ch in zip(names, kinds) if ch != "s"})
    lib = abi.load()
    nx, ny, nz = domain
    dom = abi.OebDomain(nx, ny, nz)
    args = []
    for n, ch in zip(names, kinds):
        if ch == "s":
            args.append((C.c_double if precision == "f64" else C.c_float)(float(scalars[n])))
        else:
            args.append(fields[n].ref())
    abi.check(getattr(lib, f"oeb_{name}_adj_{precision}")(C.byref(dom), *args, stencils._stream(stream)))


def adjoint(name: str, fields: dict, scalars: dict | None = None, domain=None, precision: str = "f64",
            stream=None) -> None:
    """The adjoint entry point of stencil `name` on caller-provided buffers:
    the forward inputs, the output cotangents *_bar, the input cotangents
    (overwritten on the box the forward reads) and the scratch arrays, all
    looked up by argument name."""
    _adjoint(name, fields, _scalars(name, scalars), domain, precision, stream)


def _scalars(name: str, scalars: dict | None) -> dict:
    sc = dict(KERNELS[name].scalars)
    sc.update(scalars or {})
    return sc


def _run(name: str, fields: dict, scalars: dict, domain) -> None:
    stencils.run(name, fields, scalars, domain=domain)


class Hdiffsa(torch.autograd.Function):
    """out = hdiffsa(in); differentiated in `in` (kernels/hdiffsa.h:141-169)."""

    @staticmethod
    def forward(ctx, x, consts: dict, L: Layout, domain, scalars=None):
        out = torch.zeros_like(x)
        f = dict(consts, **{"in": Storage3D(L, x), "out": Storage3D(L, out)})
        _run("hdiffsa", f, _scalars("hdiffsa", scalars), domain)
        ctx.save_for_backward(x)
        ctx.meta = (consts, L, domain, scalars)
        return out

    @staticmethod
    def backward(ctx, gout):
        (x,) = ctx.saved_tensors
        consts, L, domain, scalars = ctx.meta
        gin = torch.zeros_like(x)
        f = dict(consts, **{"in": Storage3D(L, x), "out_bar": Storage3D(L, gout.contiguous()),
                            "in_bar": Storage3D(L, gin), "lap": Storage3D(L, torch.empty_like(x))})
        _adjoint("hdiffsa", f, _scalars("hdiffsa", scalars), domain, Storage3D(L, x).precision)
        return gin, None, None, None, None


class Hadvuv5th(torch.autograd.Function):
    """(uout, vout) = hadvuv5th(uin, vin); differentiated in uin and vin
    (kernels/hadvuv5th.h:177-201)."""

    @staticmethod
    def forward(ctx, u, v, consts: dict, L: Layout, domain, scalars=None):
        uo, vo = torch.zeros_like(u), torch.zeros_like(v)
        f = dict(consts, uin=Storage3D(L, u), vin=Storage3D(L, v), uout=Storage3D(L, uo), vout=Storage3D(L, vo))
        _run("hadvuv5th", f, _scalars("hadvuv5th", scalars), domain)
        ctx.save_for_backward(u, v)
        ctx.meta = (consts, L, domain, scalars)
        return uo, vo

    @staticmethod
    def backward(ctx, gu, gv):
        u, v = ctx.saved_tensors
        consts, L, domain, scalars = ctx.meta
        ub, vb = torch.zeros_like(u), torch.zeros_like(v)
        f = dict(consts, uin=Storage3D(L, u), vin=Storage3D(L, v), uout_bar=Storage3D(L, gu.contiguous()),
                 vout_bar=Storage3D(L, gv.contiguous()), uin_bar=Storage3D(L, ub), vin_bar=Storage3D(L, vb))
        for n in ("uatupos", "vatupos", "uatvpos", "vatvpos", "uavg", "vavg", "ures", "vres"):
            f[n] = Storage3D(L, torch.empty_like(u))
        _adjoint("hadvuv5th", f, _scalars("hadvuv5th", scalars), domain, Storage3D(L, u).precision)
        return ub, vb, None, None, None, None


class Fastwavesuv(torch.autograd.Function):
    """(uout, vout) = fastwavesuv(uin, vin, ppuv); differentiated in uin, vin
    and ppuv (kernels/fastwavesuv.h:304-341).  ppuv is FP64 in both forms."""

    @staticmethod
    def forward(ctx, u, v, pp, consts: dict, L: Layout, domain, scalars=None):
        uo, vo = torch.zeros_like(u), torch.zeros_like(v)
        f = dict(consts, uin=Storage3D(L, u), vin=Storage3D(L, v), ppuv=Storage3D(L, pp), uout=Storage3D(L, uo),
                 vout=Storage3D(L, vo))
        _run("fastwavesuv", f, _scalars("fastwavesuv", scalars), domain)
        ctx.save_for_backward(pp)
        ctx.meta = (consts, L, domain, scalars, Storage3D(L, u).precision)
        return uo, vo

    @staticmethod
    def backward(ctx, gu, gv):
        (pp,) = ctx.saved_tensors
        consts, L, domain, scalars, prec = ctx.meta
        ub, vb = torch.zeros_like(gu), torch.zeros_like(gv)
        pb = torch.zeros_like(pp)
        f = dict(consts, ppuv=Storage3D(L, pp), uout_bar=Storage3D(L, gu.contiguous()),
                 vout_bar=Storage3D(L, gv.contiguous()), uin_bar=Storage3D(L, ub), vin_bar=Storage3D(L, vb),
                 ppuv_bar=Storage3D(L, pb))
        _adjoint("fastwavesuv", f, _scalars("fastwavesuv", scalars), domain, prec)
        return ub, vb, pb, None, None, None, None


class Hdiffsmag(torch.autograd.Function):
    """(uout, vout) = hdiffsmag(uin, vin); differentiated in uin and vin
    (kernels/hdiffsmag.h:205-263); the clamps of the Smagorinsky
    coefficient are frozen."""

    @staticmethod
    def forward(ctx, u, v, consts: dict, L: Layout, domain, scalars=None):
        uo, vo = torch.zeros_like(u), torch.zeros_like(v)
        f = dict(consts, uin=Storage3D(L, u), vin=Storage3D(L, v), uout=Storage3D(L, uo), vout=Storage3D(L, vo))
        _run("hdiffsmag", f, _scalars("hdiffsmag", scalars), domain)
        ctx.save_for_backward(u, v)
        ctx.meta = (consts, L, domain, scalars)
        return uo, vo

    @staticmethod
    def backward(ctx, gu, gv):
        u, v = ctx.saved_tensors
        consts, L, domain, scalars = ctx.meta
        ub, vb = torch.zeros_like(u), torch.zeros_like(v)
        f = dict(consts, uin=Storage3D(L, u), vin=Storage3D(L, v), uout_bar=Storage3D(L, gu.contiguous()),
                 vout_bar=Storage3D(L, gv.contiguous()), uin_bar=Storage3D(L, ub), vin_bar=Storage3D(L, vb))
        for n in ("work0", "work1", "work2", "work3"):
            f[n] = Storage3D(L, torch.empty_like(u))
        _adjoint("hdiffsmag", f, _scalars("hdiffsmag", scalars), domain, Storage3D(L, u).precision)
        return ub, vb, None, None, None, None


FVTP2D_OUTPUTS = ("q_i", "q_j", "fx1", "fx2", "fy1", "fy2")


class Fvtp2d(torch.autograd.Function):
    """(q_i, q_j, fx1, fx2, fy1, fy2) = fvtp2d(q); differentiated in q
    (kernels/fvtp2d.h:4-177); the PPM monotonicity switches and upwind sides
    are frozen."""

    @staticmethod
    def forward(ctx, q, consts: dict, L: Layout, domain, scalars=None):
        outs = {n: torch.zeros_like(q) for n in FVTP2D_OUTPUTS}
        f = dict(consts, q=Storage3D(L, q), **{n: Storage3D(L, t) for n, t in outs.items()})
        _run("fvtp2d", f, _scalars("fvtp2d", scalars), domain)
        ctx.save_for_backward(q, outs["q_i"], outs["q_j"])
        ctx.meta = (consts, L, domain, scalars)
        return tuple(outs[n] for n in FVTP2D_OUTPUTS)

    @staticmethod
    def backward(ctx, *grads):
        q, qi, qj = ctx.saved_tensors
        consts, L, domain, scalars = ctx.meta
        qb = torch.zeros_like(q)
        f = dict(consts, q=Storage3D(L, q), q_i=Storage3D(L, qi), q_j=Storage3D(L, qj), q_bar=Storage3D(L, qb))
        for n, gr in zip(FVTP2D_OUTPUTS, grads):
            f[f"{n}_bar"] = Storage3D(L, (gr if gr is not None else torch.zeros_like(q)).contiguous())
        # The adjoint writes every point of a work array that it reads (each
        # on its output's region), so they need no clearing.
        for n in ("work0", "work1", "work2", "work3"):
            f[n] = Storage3D(L, torch.empty_like(q))
        _adjoint("fvtp2d", f, _scalars("fvtp2d", scalars), domain, Storage3D(L, q).precision)
        return qb, None, None, None, None
