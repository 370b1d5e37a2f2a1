"""Adjoints: the HIP adjoints behind oeb.autograd against torch.autograd
through the ATen compositions of tests/aten_ref.py (FP64, 1e-12 relative to
the largest gradient), and against FP64 central differences of the HIP
forward (eps 1e-6) at 1000 seeded points, to 1e-7 relative to the largest
gradient.

Central differences: the points of one pass are spaced more than twice the
stencil's reach apart, so each output is moved by at most one of them; the
difference quotient of point q is sum_p ybar(p) (y+(p) - y-(p)) / (2 eps) over
the outputs p within reach of q.  A perturbation can flip a frozen branch
(limiter, upwind direction); such points are identified by the
second difference and reported, and at most 1% may be excluded.
"""
import numpy as np
import pytest

import aten_ref as A
from common import inputs
from oeb import autograd as AG
from oeb.generators import cast, generate
from oeb.layout import Layout, Layout1D
from oeb.spec import KERNELS
from oeb.storage import Storage1D, Storage3D

pytestmark = pytest.mark.gpu

SHAPE = (40, 36, 5)
FD_POINTS = 1000
FD_EPS = 1e-6
FD_TOL = 1e-7
ATEN_TOL = 1e-12


def _dev(name, L, host, skip=()):
    import torch

    L1 = Layout1D(L.ny, L.h)
    out = {}
    for a in KERNELS[name].args:
        if a.kind in ("in", "in1") and a.name not in skip:
            out[a.name] = (Storage1D if a.kind == "in1" else Storage3D).to_device(L1 if a.kind == "in1" else L,
                                                                                host[a.name])
    return out


def _t(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _relmax(x, y):
    import torch

    return float((x - y).abs().max() / y.abs().max().clamp_min(1e-300))


# ------------------------------------------------------------ the problems
# Each builder takes (seed, precision, shape, rounded) and returns (L, forward
# (*diff) -> tuple of outputs, ATen forward, diff inputs, read boxes of the
# diff inputs as (i0, i1, j0, j1, k0, k1), reach).  precision "f32" runs the
# FP32 form (FP64 islands).  With precision "f32" or rounded=True every input
# is first rounded to what the FP32 form stores, so the FP64 and FP32 problems
# of one seed have the same inputs; the ATen forward always takes those
# inputs in FP64.
def _host(name, L, seed, precision, rounded):
    # hdiffsmag takes the stress inputs (acrlat0 ~ 10^U(0, 4)): with the
    # baseline ones its Smagorinsky coefficient clamps to 0 everywhere, so the
    # adjoint would be the identity.
    host = generate(name, L, "stress" if name == "hdiffsmag" else "baseline", seed)
    if precision == "f32" or rounded:
        host = {k: v.astype(np.float64) for k, v in cast(name, host, "f32").items()}
    return host, cast(name, host, precision)


def _hdiffsa(seed, precision="f64", shape=SHAPE, rounded=False):
    L = Layout(*shape, h=3, hk=1)
    host, hp = _host("hdiffsa", L, seed, precision, rounded)
    c = _dev("hdiffsa", L, hp, skip=("in",))
    dom = (L.nx, L.ny, L.nz)
    fwd = lambda x: (AG.Hdiffsa.apply(x, c, L, dom),)  # noqa: E731
    ref = lambda x: (A.hdiffsa(x, _t(host["mask"]), _t(host["crlato"]), _t(host["crlatu"]), L),)  # noqa: E731
    return L, fwd, ref, [_t(hp["in"])], [(-2, L.nx + 2, -2, L.ny + 2, 0, L.nz)], 2


def _hadvuv5th(seed, precision="f64", shape=SHAPE, rounded=False):
    L = Layout(*shape, h=3, hk=1)
    host, hp = _host("hadvuv5th", L, seed, precision, rounded)
    c = _dev("hadvuv5th", L, hp, skip=("uin", "vin"))
    dom = (L.nx, L.ny, L.nz)
    sc = dict(KERNELS["hadvuv5th"].scalars)
    fwd = lambda u, v: AG.Hadvuv5th.apply(u, v, c, L, dom)  # noqa: E731
    ref = lambda u, v: A.hadvuv5th(u, v, _t(host["acrlat0"]), _t(host["acrlat1"]), _t(host["tgrlatda0"]),  # noqa
                                   _t(host["tgrlatda1"]), L, sc["eddlat"], sc["eddlon"])
    box = (-3, L.nx + 3, -3, L.ny + 3, 0, L.nz)
    return L, fwd, ref, [_t(hp["uin"]), _t(hp["vin"])], [box, box], 3


def _fastwavesuv(seed, precision="f64", shape=SHAPE, rounded=False):
    L = Layout(*shape, h=3, hk=1)
    host, hp = _host("fastwavesuv", L, seed, precision, rounded)
    c = _dev("fastwavesuv", L, hp, skip=("uin", "vin", "ppuv"))
    dom = (L.nx, L.ny, L.nz)
    sc = dict(KERNELS["fastwavesuv"].scalars)
    fwd = lambda u, v, p: AG.Fastwavesuv.apply(u, v, p, c, L, dom)  # noqa: E731
    ref = lambda u, v, p: A.fastwavesuv(u, v, p, _t(host["utens"]), _t(host["vtens"]), _t(host["wgtfac"]),  # noqa
                                        _t(host["hhl"]), _t(host["rho"]), _t(host["fx"]), L, sc["edadlat"],
                                        sc["dt"])
    inner = (0, L.nx, 0, L.ny, 0, L.nz)
    return (L, fwd, ref, [_t(hp["uin"]), _t(hp["vin"]), _t(hp["ppuv"])],
            [inner, inner, (0, L.nx + 1, 0, L.ny + 1, -1, L.nz + 1)], 1)


def _hdiffsmag(seed, precision="f64", shape=SHAPE, rounded=False):
    L = Layout(*shape, h=3, hk=1)
    host, hp = _host("hdiffsmag", L, seed, precision, rounded)
    c = _dev("hdiffsmag", L, hp, skip=("uin", "vin"))
    dom = (L.nx, L.ny, L.nz)
    sc = dict(KERNELS["hdiffsmag"].scalars)
    fwd = lambda u, v: AG.Hdiffsmag.apply(u, v, c, L, dom)  # noqa: E731
    ref = lambda u, v: A.hdiffsmag(u, v, _t(host["mask"]), _t(host["crlavo"]), _t(host["crlavu"]),  # noqa: E731
                                   _t(host["crlato"]), _t(host["crlatu"]), _t(host["acrlat0"]), L, sc["eddlat"],
                                   sc["eddlon"], sc["tau_smag"], sc["weight_smag"])
    box = (-1, L.nx + 1, -1, L.ny + 1, 0, L.nz)
    return L, fwd, ref, [_t(hp["uin"]), _t(hp["vin"])], [box, box], 2


def _fvtp2d(seed, precision="f64", shape=SHAPE, rounded=False):
    L = Layout(*shape, h=3, hk=1)
    host, hp = _host("fvtp2d", L, seed, precision, rounded)
    c = _dev("fvtp2d", L, hp, skip=("q",))
    dom = (L.nx, L.ny, L.nz)
    sc = dict(KERNELS["fvtp2d"].scalars)
    names = ("crx", "cry", "ra_x", "ra_y", "xfx", "yfx", "area")
    fwd = lambda q: AG.Fvtp2d.apply(q, c, L, dom)  # noqa: E731
    ref = lambda q: A.fvtp2d(q, *[_t(host[n]) for n in names], L, sc["p1"], sc["p2"])  # noqa: E731
    return L, fwd, ref, [_t(hp["q"])], [(-3, L.nx + 3, -3, L.ny + 3, 0, L.nz)], 6


PROBLEMS = {"hdiffsa": _hdiffsa, "hadvuv5th": _hadvuv5th, "fastwavesuv": _fastwavesuv, "hdiffsmag": _hdiffsmag,
            "fvtp2d": _fvtp2d}


def _grads(fwd, xs, ybars):
    import torch

    xs = [x.clone().requires_grad_(True) for x in xs]
    ys = fwd(*xs)
    loss = sum((y * yb).sum() for y, yb in zip(ys, ybars))
    return torch.autograd.grad(loss, xs)


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_adjoint_matches_aten_autograd(name):
    import torch

    L, fwd, ref, xs, boxes, _ = PROBLEMS[name](seed=31)
    print(f"[gpu] adjoint vs autograd {name}", flush=True)
    ys = fwd(*xs)
    g = torch.Generator(device="cuda").manual_seed(5)
    ybars = [torch.randn(y.shape, generator=g, device="cuda", dtype=y.dtype) for y in ys]
    ours = _grads(fwd, xs, ybars)
    theirs = _grads(ref, xs, ybars)
    # The forward outputs agree first (both FP64, same expressions).
    for y, yr in zip(ys, ref(*xs)):
        assert _relmax(y, yr) <= 1e-13
    for a, b in zip(ours, theirs):
        err = _relmax(a, b)
        assert err <= ATEN_TOL, f"{name}: adjoint vs autograd rel max {err:.2e}"


def _central_differences(name, L, fwd, xs, boxes, reach, ybars, grads, tol):
    """Checks grads (cotangents of xs for the output cotangents ybars)
    against central differences of fwd at FD_POINTS seeded points, to tol
    relative to the largest gradient; returns (points, branch flips)."""
    import torch

    with torch.no_grad():
        ys = fwd(*xs)
    gmax = max(float(gr.abs().max()) for gr in grads)
    rng = np.random.default_rng(7)
    spacing = 2 * reach + 1
    checked, bad, flips = 0, [], 0
    per_field = FD_POINTS // len(xs) + 1
    for f, (x, box) in enumerate(zip(xs, boxes)):
        i0, i1, j0, j1, k0, k1 = box
        done = 0
        while done < per_field:
            oi, oj = rng.integers(0, spacing, size=2)
            ii = np.arange(i0 + oi, i1, spacing)
            jj = np.arange(j0 + oj, j1, spacing)
            kk = rng.integers(k0, k1)
            I, J = np.meshgrid(ii, jj)
            I, J = I.ravel(), J.ravel()
            sel = rng.permutation(len(I))[: per_field - done]
            I, J = I[sel], J[sel]
            idx = torch.from_numpy(L.index(I, J, kk).astype(np.int64)).to("cuda")
            eps = FD_EPS * torch.clamp(x[idx].abs(), min=1.0)

            def run(sign):
                xp = [xx.clone() for xx in xs]
                xp[f][idx] += sign * eps
                with torch.no_grad():
                    return fwd(*xp)

            yp, ym, y0 = run(1.0), run(-1.0), ys
            for n, q in enumerate(idx.tolist()):
                fd = 0.0
                curv = 0.0
                for y_p, y_m, y_0, yb in zip(yp, ym, y0, ybars):
                    v = lambda a: a.view(L.levels, L.rows, L.sj)  # noqa: E731
                    k_, j_, i_ = np.unravel_index(q, (L.levels, L.rows, L.sj))
                    sl = (slice(max(k_ - 1, 0), k_ + 2), slice(max(j_ - reach, 0), j_ + reach + 1),
                          slice(max(i_ - reach, 0), i_ + reach + 1))
                    fd += float((v(yb)[sl] * (v(y_p)[sl] - v(y_m)[sl])).sum())
                    curv += float((v(yb)[sl] * (v(y_p)[sl] + v(y_m)[sl] - 2.0 * v(y_0)[sl])).abs().sum())
                fd /= 2.0 * float(eps[n])
                adj = float(grads[f][q])
                checked += 1
                if abs(fd - adj) > tol * gmax:
                    # A frozen branch flipped under the perturbation: the
                    # outputs jump, and the second difference shows it.
                    if curv > 1e-3 * abs(fd) * float(eps[n]):
                        flips += 1
                    else:
                        bad.append((f, q, fd, adj))
            done += len(idx)
    print(f"  {name}: {checked} points, {flips} branch flips excluded, {len(bad)} failures", flush=True)
    assert checked >= FD_POINTS
    assert flips <= 0.01 * checked
    assert not bad, f"{name}: {len(bad)} points off by more than {tol:.0e} x max|grad|: {bad[:3]}"
    return checked, flips


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_adjoint_central_differences(name):
    import torch

    L, fwd, _, xs, boxes, reach = PROBLEMS[name](seed=32)
    print(f"[gpu] adjoint vs central differences {name}", flush=True)
    with torch.no_grad():
        ys = fwd(*xs)
    g = torch.Generator(device="cuda").manual_seed(6)
    ybars = [torch.randn(y.shape, generator=g, device="cuda", dtype=y.dtype) for y in ys]
    grads = _grads(fwd, xs, ybars)
    _central_differences(name, L, fwd, xs, boxes, reach, ybars, grads, FD_TOL)


# ----------------------------------------------------------- FP32 adjoints
# The FP32 adjoints (FP32 storage and arithmetic, FP64 islands) differentiate
# the FP32 forward, whose branches (limiters, upwind directions, clamps, PPM
# monotonicity) they freeze.  Independent checks, on inputs rounded to what the
# FP32 form stores:
#   - against torch.autograd through the ATen compositions in FP64 on those
#     inputs.  The values differ by FP32 rounding (bound F32_ATEN_TOL of the
#     largest gradient) everywhere except where the FP32 and FP64 forwards
#     take different branches (products within rounding of zero): there a
#     point's gradient jumps, and the fraction of such points is bounded
#     (F32_FLIP_FRAC; hdiffsa's flux limiter flips at ~1e-4 of the points of
#     256x256x8, the others at none).
#   - against central differences of the FP64 HIP forward (eps 1e-6) at 1000
#     seeded points, to F32_FD_TOL of the largest gradient, branch flips
#     excluded as in FP64.
F32_ATEN_TOL = 1e-5
F32_FLIP_FRAC = 1e-3
F32_FD_TOL = 1e-5


def _aten_vs_f32(name, seed, shape):
    import torch

    L, fwd, ref, xs, boxes, _ = PROBLEMS[name](seed=seed, precision="f32", shape=shape)
    ys = fwd(*xs)
    g = torch.Generator(device="cuda").manual_seed(5)
    ybars = [torch.randn(y.shape, generator=g, device="cuda", dtype=y.dtype) for y in ys]
    ours = _grads(fwd, xs, ybars)
    xs64 = [x.double() for x in xs]
    theirs = _grads(ref, xs64, [yb.double() for yb in ybars])
    for a, b in zip(ours, theirs):
        a = a.double()
        gmax = float(b.abs().max())
        diff = (a - b).abs()
        frac = float((diff > F32_ATEN_TOL * gmax).double().mean())
        rms = float(diff.pow(2).mean().sqrt() / b.pow(2).mean().sqrt())
        print(f"  {name} {shape}: FP32 adjoint vs FP64 autograd rel max {float(diff.max()) / gmax:.2e} "
              f"rms {rms:.2e} frac>{F32_ATEN_TOL:.0e} {frac:.1e}", flush=True)
        # Points where the two forwards take different branches are off by
        # O(1); they set the max and the RMS, so only their fraction is bounded.
        assert frac <= F32_FLIP_FRAC, f"{name}: {frac:.1e} of the points off by more than {F32_ATEN_TOL:.0e}"


@pytest.mark.parametrize("shape", [SHAPE, (256, 256, 8)], ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("name", list(PROBLEMS))
def test_f32_adjoint_matches_aten_autograd(name, shape):
    print(f"[gpu] FP32 adjoint vs autograd {name} {shape}", flush=True)
    _aten_vs_f32(name, 33, shape)


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_f32_adjoint_central_differences(name):
    import torch

    L, fwd32, _, xs32, boxes, reach = PROBLEMS[name](seed=34, precision="f32")
    _, fwd64, _, xs64, _, _ = PROBLEMS[name](seed=34, rounded=True)
    for a, b in zip(xs32, xs64):
        assert torch.equal(a.double(), b)
    print(f"[gpu] FP32 adjoint vs central differences {name}", flush=True)
    with torch.no_grad():
        ys = fwd32(*xs32)
    g = torch.Generator(device="cuda").manual_seed(6)
    ybars = [torch.randn(y.shape, generator=g, device="cuda", dtype=y.dtype) for y in ys]
    grads = [gr.double() for gr in _grads(fwd32, xs32, ybars)]
    _central_differences(name, L, fwd64, xs64, boxes, reach, [yb.double() for yb in ybars], grads, F32_FD_TOL)


@pytest.mark.parametrize("rows", [1, 3, 64])
@pytest.mark.parametrize("shape", [(40, 36, 5), (67, 9, 3)], ids=lambda s: "x".join(map(str, s)))
def test_hdiffsa_adjoint_strip_matches_two_pass(shape, rows, knobs):
    """The fused row-march hdiffsa adjoint and the two-pass one (lap_bar
    through memory), forced through the form knob, give the same
    doubles, at row chunks of one row, of a number that does not divide
    ny + 4, and of more than ny + 4."""
    import torch

    L = Layout(*shape, h=3, hk=1)
    host = generate("hdiffsa", L, "baseline", 41)
    c = _dev("hdiffsa", L, host, skip=("in",))
    dom = (L.nx, L.ny, L.nz)
    x = _t(host["in"])
    g = torch.Generator(device="cuda").manual_seed(3)
    ybar = torch.randn(x.shape, generator=g, device="cuda", dtype=x.dtype)
    print(f"[gpu] hdiffsa adjoint strip vs two-pass {shape} rows={rows}", flush=True)
    out = []
    for strip in (0, 1):
        knobs.set("hdiffsa_adj", "form", strip)
        knobs.set("hdiffsa_adj", "rows", rows)
        out.append(_grads(lambda v: (AG.Hdiffsa.apply(v, c, L, dom),), [x], [ybar])[0])
    assert torch.equal(out[0], out[1])


@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("rows", [1, 4, 64])
@pytest.mark.parametrize("shape", [(40, 36, 5), (67, 9, 3)], ids=lambda s: "x".join(map(str, s)))
def test_hadvuv5th_adjoint_strip_matches_two_pass(shape, rows, depth, knobs):
    """The fused row-march hadvuv5th adjoint (x contributions gathered from
    LDS, y contributions pushed down each column), with rows loaded one and
    two ahead, and the two-pass one (eight scratch arrays), forced through
    the form knob, agree to 1e-13 of the largest cotangent: the same
    terms, summed in another order."""
    import torch

    L = Layout(*shape, h=3, hk=1)
    host = generate("hadvuv5th", L, "baseline", 43)
    c = _dev("hadvuv5th", L, host, skip=("uin", "vin"))
    dom = (L.nx, L.ny, L.nz)
    xs = [_t(host["uin"]), _t(host["vin"])]
    g = torch.Generator(device="cuda").manual_seed(4)
    ybars = [torch.randn(x.shape, generator=g, device="cuda", dtype=x.dtype) for x in xs]
    print(f"[gpu] hadvuv5th adjoint strip vs two-pass {shape} rows={rows} depth={depth}", flush=True)
    out = []
    knobs.set("hadvuv5th_adj", "tile", 0)
    knobs.set("hadvuv5th_adj", "depth", depth)
    for strip in (0, 1):
        knobs.set("hadvuv5th_adj", "form", strip)
        knobs.set("hadvuv5th_adj", "rows", rows)
        out.append(_grads(lambda u, v: AG.Hadvuv5th.apply(u, v, c, L, dom), xs, ybars))
    for a, b in zip(*out):
        assert _relmax(b, a) <= 1e-13


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("kchunk", [1, 2, 3])
@pytest.mark.parametrize("tile", [1, 2])
@pytest.mark.parametrize("shape", [(40, 36, 5), (67, 9, 3), (130, 21, 2), (64, 32, 4), (3, 2, 2), (200, 40, 2)],
                         ids=lambda s: "x".join(map(str, s)))
def test_hadvuv5th_adjoint_tile_matches_two_pass(shape, tile, kchunk, precision, knobs):
    """The tiled hadvuv5th adjoint (stage 1 of a tile and its 3-point halo
    in LDS, then the two-pass form's gathers), forced through
    the tile knob (FP64: 64x16 tiles; FP32: 64x8 and 64x16) with
    1-3 levels per workgroup, and the two-pass form give the same bits: the
    same expressions, the same terms in the same order.  Shapes with edge
    tiles narrower than a tile, several tiles each way, a 3x2 domain, and
    tiles 3 or more points from every edge (200x40: the gathers without
    edge tests)."""
    import torch

    from oeb.generators import cast

    L = Layout(*shape, h=3, hk=1)
    host = cast("hadvuv5th", generate("hadvuv5th", L, "baseline", 53), precision)
    c = _dev("hadvuv5th", L, host, skip=("uin", "vin"))
    dom = (L.nx, L.ny, L.nz)
    xs = [_t(host["uin"]), _t(host["vin"])]
    g = torch.Generator(device="cuda").manual_seed(12)
    ybars = [torch.randn(x.shape, generator=g, device="cuda", dtype=x.dtype) for x in xs]
    print(f"[gpu] hadvuv5th adjoint tile {tile} vs two-pass {shape} kchunk={kchunk} {precision}", flush=True)
    out = []
    knobs.set("hadvuv5th_adj", "form", 0)
    knobs.set("hadvuv5th_adj", "kchunk", kchunk)
    for t in (0, tile):
        knobs.set("hadvuv5th_adj", "tile", t)
        out.append(_grads(lambda u, v: AG.Hadvuv5th.apply(u, v, c, L, dom), xs, ybars))
    for a, b in zip(*out):
        assert torch.equal(a, b)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("kchunk", [1, 3])
@pytest.mark.parametrize("tile", [1, 2])
@pytest.mark.parametrize("shape", [(40, 36, 5), (67, 9, 3), (130, 21, 2), (64, 32, 4), (3, 2, 2), (200, 40, 2)],
                         ids=lambda s: "x".join(map(str, s)))
def test_hdiffsmag_adjoint_tile_matches_two_pass(shape, tile, kchunk, precision, knobs):
    """The tiled hdiffsmag adjoint (stage 1, then T_bar and S_bar, in LDS,
    then the two-pass form's gathers), forced through the tile knob
    with 64x8 and 64x16 tiles and 1 or 3 levels per workgroup, and the
    two-pass form give the same bits: the same expressions, terms and order.
    Shapes with edge tiles narrower than a tile, several tiles each way and
    a 3x2 domain."""
    import torch

    from oeb.generators import cast

    L = Layout(*shape, h=3, hk=1)
    host = cast("hdiffsmag", generate("hdiffsmag", L, "baseline", 59), precision)
    c = _dev("hdiffsmag", L, host, skip=("uin", "vin"))
    dom = (L.nx, L.ny, L.nz)
    xs = [_t(host["uin"]), _t(host["vin"])]
    g = torch.Generator(device="cuda").manual_seed(14)
    ybars = [torch.randn(x.shape, generator=g, device="cuda", dtype=x.dtype) for x in xs]
    print(f"[gpu] hdiffsmag adjoint tile {tile} vs two-pass {shape} kchunk={kchunk} {precision}", flush=True)
    out = []
    knobs.set("hdiffsmag_adj", "form", 0)
    knobs.set("hdiffsmag_adj", "kchunk", kchunk)
    for t in (0, tile):
        knobs.set("hdiffsmag_adj", "tile", t)
        out.append(_grads(lambda u, v: AG.Hdiffsmag.apply(u, v, c, L, dom), xs, ybars))
    for a, b in zip(*out):
        assert torch.equal(a, b)


@pytest.mark.parametrize("rows", [1, 3, 64])
@pytest.mark.parametrize("shape", [(40, 36, 5), (67, 9, 3)], ids=lambda s: "x".join(map(str, s)))
def test_hdiffsmag_adjoint_strip_matches_two_pass(shape, rows, knobs):
    """The fused row-march hdiffsmag adjoint and the two-pass one (gu, gv,
    wu, wv through memory), forced through the form knob, give the
    same doubles, at row chunks of one row, of a number that does not divide
    ny + 2, and of more than ny + 2."""
    import torch

    L = Layout(*shape, h=3, hk=1)
    host = generate("hdiffsmag", L, "baseline", 47)
    c = _dev("hdiffsmag", L, host, skip=("uin", "vin"))
    dom = (L.nx, L.ny, L.nz)
    xs = [_t(host["uin"]), _t(host["vin"])]
    g = torch.Generator(device="cuda").manual_seed(8)
    ybars = [torch.randn(x.shape, generator=g, device="cuda", dtype=x.dtype) for x in xs]
    print(f"[gpu] hdiffsmag adjoint strip vs two-pass {shape} rows={rows}", flush=True)
    out = []
    knobs.set("hdiffsmag_adj", "tile", 0)
    for strip in (0, 1):
        knobs.set("hdiffsmag_adj", "form", strip)
        knobs.set("hdiffsmag_adj", "rows", rows)
        out.append(_grads(lambda u, v: AG.Hdiffsmag.apply(u, v, c, L, dom), xs, ybars))
    for a, b in zip(*out):
        assert torch.equal(a, b)


@pytest.mark.parametrize("shape", [(40, 36, 5), (67, 9, 3), (130, 21, 2)], ids=lambda s: "x".join(map(str, s)))
def test_fvtp2d_adjoint_tiled_matches_per_point(shape, knobs):
    """The tiled transposed PPM gathers of the fvtp2d adjoint (the products
    of each face formed once in LDS) and the per-point ones, forced through
    the tile knob, agree to 1e-14 of the largest cotangent: the same
    products summed in the same order (a skipped face adds an exact zero)."""
    import torch

    L = Layout(*shape, h=3, hk=1)
    host = generate("fvtp2d", L, "baseline", 49)
    c = _dev("fvtp2d", L, host, skip=("q",))
    dom = (L.nx, L.ny, L.nz)
    q = _t(host["q"])
    ys = AG.Fvtp2d.apply(q, c, L, dom)
    g = torch.Generator(device="cuda").manual_seed(10)
    ybars = [torch.randn(y.shape, generator=g, device="cuda", dtype=y.dtype) for y in ys]
    print(f"[gpu] fvtp2d adjoint tiled vs per point {shape}", flush=True)
    out = []
    for tiled in (0, 1, 2):
        knobs.set("fvtp2d_adj", "tile", tiled)
        out.append(_grads(lambda v: AG.Fvtp2d.apply(v, c, L, dom), [q], ybars)[0])
    assert _relmax(out[1], out[0]) <= 1e-14
    # Tile 2 (the default) issues every face's loads first, at clamped
    # addresses: the same products in the same order as tile 1.
    assert torch.equal(out[2], out[1])


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(64, 16, 2), (67, 9, 3), (130, 21, 2), (256, 256, 4)],
                         ids=lambda s: "x".join(map(str, s)) if isinstance(s, tuple) else s)
def test_fvtp2d_adjoint_hoisted_tiles_bit_exact(shape, precision, knobs):
    """The hoisted tiles (tile knob 2) against the per-face tiles (tile knob
    1) bit for bit, in both precisions, on shapes with whole and partial
    tiles, through the adjoint entry point with every cotangent nonzero."""
    import torch

    L = Layout(*shape, h=3, hk=1)
    host = cast("fvtp2d", generate("fvtp2d", L, "baseline", 61), precision)
    c = _dev("fvtp2d", L, host, skip=("q",))
    dom = (L.nx, L.ny, L.nz)
    q = _t(host["q"])
    ys = AG.Fvtp2d.apply(q, c, L, dom)
    g = torch.Generator(device="cuda").manual_seed(12)
    ybars = [torch.randn(y.shape, generator=g, device="cuda", dtype=torch.float64).to(y.dtype) for y in ys]
    print(f"[gpu] fvtp2d adjoint hoisted vs per-face tiles {shape} {precision}", flush=True)
    out = []
    for tile in (1, 2):
        knobs.set("fvtp2d_adj", "tile", tile)
        out.append(_grads(lambda v: AG.Fvtp2d.apply(v, c, L, dom), [q], ybars)[0])
    assert torch.isfinite(out[1]).all()
    assert torch.equal(out[1], out[0])
