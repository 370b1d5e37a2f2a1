"""The reference's operator API, over liboeb.

One function per stencil, with the reference's name and argument order
(kernels/<name>.h).  Arguments are Storage3D / Storage1D objects with DEVICE
buffers.  Scratch arguments may be None: the GPU keeps every temporary on
chip.  Like the reference, the extents come from module globals
(``domain_size`` and ``domain_height``, src/<name>.cpp:4-5), set with
set_domain(); a ``domain=(nx, ny, nz)`` keyword overrides them per call.

The reference functions return void and have no error path.  Here a
rejected call (bad layout, aliasing, launch failure) raises
abi.OebError.  Launches are asynchronous on ``stream``: a hipStream_t
handle, or by default torch's current stream.
"""
from __future__ import annotations

import ctypes as C

from . import abi
from .spec import KERNELS
from .storage import DeviationIsland, require_device

domain_size = 64  # src/<name>.cpp:4
domain_height = 60  # src/<name>.cpp:5


def set_domain(size: int, height: int) -> None:
    global domain_size, domain_height
    domain_size, domain_height = int(size), int(height)


def _stream(stream):
    if stream is not None:
        return C.c_void_p(int(stream))
    try:
        import torch

        if torch.cuda.is_available():
            return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    except ImportError:  # pragma: no cover
        pass
    return C.c_void_p(0)


def _domain(domain) -> abi.OebDomain:
    if domain is None:
        return abi.OebDomain(domain_size, domain_size, domain_height)
    nx, ny, nz = domain
    return abi.OebDomain(nx, ny, nz)


def call(name: str, arrays: list, scalars: list, domain=None, stream=None, precision=None,
         check_device: bool = True) -> None:
    """Calls oeb_<name>_<precision> with arrays in reference order (None ->
    NULL).  The precision is that of the first array (an output) unless
    given.  In the FP32 form the FP64 island arrays of fastwavesuv, p_grad_c
    and nh_p_grad must be float64, every other array float32; the element
    types are checked here, because the C-ABI cannot see them.  Islands
    given as DeviationIsland call oeb_<name>_f32d.  Host buffers
    raise TypeError here (check_device=False hands them to the library,
    whose own check rejects them: the C-ABI validation tests use that)."""
    lib = abi.load()
    if check_device:
        require_device(name, {a.name: x for a, x in zip(KERNELS[name].args, arrays)})
    dom = _domain(domain)
    if precision is None:
        precision = arrays[0].precision if arrays and arrays[0] is not None else "f64"
    sig = abi.signature(name, precision)
    assert len(arrays) + len(scalars) == len(sig), name
    # FP32 islands given as DeviationIsland select oeb_<name>_f32d: all or none.
    devs = [isinstance(a, DeviationIsland) for a, kind in zip(arrays, sig) if kind == "D"]
    entry = f"oeb_{name}_{precision}"
    if any(devs):
        if not all(devs) or name not in abi.SIGNATURES_F32D:
            raise TypeError(f"{name}_{precision}: islands must be all FP64 arrays or all DeviationIsland"
                            + ("" if name in abi.SIGNATURES_F32D else f" ({name} has no deviation form)"))
        entry = f"oeb_{name}_f32d"
    cargs = []
    spec_args = KERNELS[name].args
    for a, kind, spec_arg in zip(arrays, sig, spec_args):
        if isinstance(a, DeviationIsland):
            if kind != "D":
                raise TypeError(f"{name}_{precision}: array '{spec_arg.name}' is not an island")
            cargs.extend(a.refs())
            continue
        if a is not None:
            want = "f64" if (precision == "f64" or kind == "D") else "f32"
            if a.precision != want:
                raise TypeError(f"{name}_{precision}: array '{spec_arg.name}' must be {want}, got {a.precision}")
        cargs.append(None if a is None else a.ref())
    for s_, kind in zip(scalars, sig[len(arrays):]):
        cargs.append(C.c_double(float(s_)) if kind == "d" else C.c_float(float(s_)))
    abi.check(getattr(lib, entry)(C.byref(dom), *cargs, _stream(stream)))


def run(name: str, fields: dict, scalars: dict | None = None, domain=None, stream=None, precision=None,
        check_device: bool = True) -> None:
    """Calls stencil `name` with arrays looked up by reference argument name
    (missing scratch -> NULL) and scalars defaulting to the driver values."""
    spec = KERNELS[name]
    arrays = [fields.get(a.name) if a.kind == "scratch" else fields[a.name] for a in spec.args]
    sc = dict(spec.scalars)
    if scalars:
        sc.update(scalars)
    call(name, arrays, [sc[n] for n, _ in spec.scalars], domain, stream, precision, check_device)


# ------------------------------------------------- reference signatures
def laplace(in_, out, *, domain=None, stream=None):
    """kernels/laplace.h:4 laplace(const Storage3D& in, Storage3D& out)."""
    call("laplace", [in_, out], [], domain, stream)


def hdiffsa(out, in_, mask, crlato, crlatu, lap=None, flx=None, fly=None, *, domain=None, stream=None):
    """kernels/hdiffsa.h:4-5."""
    call("hdiffsa", [out, in_, mask, crlato, crlatu, lap, flx, fly], [], domain, stream)


def hdiffsmag(uout, vout, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, T_sqr_s, S_sqr_uv, eddlat,
              eddlon, tau_smag, weight_smag, *, domain=None, stream=None):
    """kernels/hdiffsmag.h:4-7."""
    call("hdiffsmag", [uout, vout, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, T_sqr_s, S_sqr_uv],
         [eddlat, eddlon, tau_smag, weight_smag], domain, stream)


def hadvuv(uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uatupos, vatupos, uatvpos, vatvpos, uavg,
           vavg, ures, vres, eddlat, eddlon, *, domain=None, stream=None):
    """kernels/hadvuv.h:29-32."""
    call("hadvuv", [uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uatupos, vatupos, uatvpos,
                    vatvpos, uavg, vavg, ures, vres], [eddlat, eddlon], domain, stream)


def hadvuv5th(uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uatupos, vatupos, uatvpos, vatvpos,
              uavg, vavg, ures, vres, eddlat, eddlon, *, domain=None, stream=None):
    """kernels/hadvuv5th.h:41-44."""
    call("hadvuv5th", [uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uatupos, vatupos, uatvpos,
                       vatvpos, uavg, vavg, ures, vres], [eddlat, eddlon], domain, stream)


def fastwavesuv(uout, vout, uin, vin, utens, vtens, wgtfac, ppuv, hhl, rho, fx, ppgk, ppgc, ppgu, ppgv, edadlat, dt,
                *, domain=None, stream=None):
    """kernels/fastwavesuv.h:4-7."""
    call("fastwavesuv", [uout, vout, uin, vin, utens, vtens, wgtfac, ppuv, hhl, rho, fx, ppgk, ppgc, ppgu, ppgv],
         [edadlat, dt], domain, stream)


def uvbke(ub, vb, uc, vc, cosa, rsina, dt5=0.5 * 225.0, *, domain=None, stream=None):
    """kernels/uvbke.h:4-5; dt5 defaults to the driver's 0.5 * 225 (src/uvbke.cpp:11-12)."""
    call("uvbke", [ub, vb, uc, vc, cosa, rsina], [dt5], domain, stream)


def p_grad_c(uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, wk, dt2, *, domain=None, stream=None):
    """kernels/p_grad_c.h:4-6."""
    call("p_grad_c", [uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, wk], [dt2], domain, stream)


def nh_p_grad(uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, wk1, wk, du, dv, dt, *, domain=None, stream=None):
    """kernels/nh_p_grad.h:4-6."""
    call("nh_p_grad", [uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, wk1, wk, du, dv], [dt], domain, stream)


def fvtp2d(q_i, q_j, fx1, fx2, fy1, fy2, q, crx, cry, ra_x, ra_y, xfx, yfx, area, fxx=None, fyy=None, al=None,
           almq=None, br=None, b0=None, smt5=None, p1=0.583333, p2=0.083333, *, domain=None, stream=None):
    """kernels/fvtp2d.h:4-7; p1/p2 default to the driver's constants (src/fvtp2d.cpp:11-12)."""
    call("fvtp2d", [q_i, q_j, fx1, fx2, fy1, fy2, q, crx, cry, ra_x, ra_y, xfx, yfx, area, fxx, fyy, al, almq, br,
                    b0, smt5], [p1, p2], domain, stream)
