"""Seeded synthetic inputs for the ten stencils.

Every generator writes every allocated point, i/j halo, pitch padding and
k-halo planes included.  The reference never fills its k-halo planes, yet
fastwavesuv, p_grad_c and nh_p_grad read them (kernels/fastwavesuv.h:312,
kernels/p_grad_c.h:17-19, kernels/nh_p_grad.h:123).  Outputs are filled with
NaN, so a point a kernel fails to write, or writes outside its region, shows
up.

Families:
  baseline   the synthetic distributions of the benchmark workloads:
             tracer = 3 random-phase sinusoids (wavenumbers 1-4, amplitude 1)
             + U(-1e-3, 1e-3); wind = N(0, 10 m/s) smoothed by a 5-point box
             filter; pressure = 1e5 exp(-z/8000) (z = 250 m per level) times
             (1 + N(0, 0.01)) per column; thickness = level difference of that
             profile; geopotential = 9.81 z + 9.81 N(0, 1); hhl = monotone
             z-based half-level heights over a smooth terrain; Courant numbers
             ~ U(-0.4, 0.4); metric / coefficient fields ~ U(0.01, 0.1);
             crlat* = cos(latitude) at ny evenly spaced latitudes in
             [-80, 80] degrees (the halo continues the spacing).
  refdriver  the reference drivers' own fillMath inputs (include/util.h:150-188
             with the per-field parameters of src/<name>.cpp), extended to the
             k-halo planes by the same formula.
  stress     baseline with overrides that drive rarely taken branches, e.g.
             hdiffsmag's acrlat0 ~ 10^U(0, 4), so both clamps of the
             Smagorinsky coefficient are hit.
"""
from __future__ import annotations

import math

import numpy as np

from .layout import Layout, Layout1D
from .spec import KERNELS, ORDER

# --------------------------------------------------------------- recipes
_F1 = (1.1, 2.0, 1.5, 2.8, 2.0, 4.1)
_F2 = (8.0, 9.4, 1.5, 1.7, 2.0, 3.5)
_F3 = (5.0, 8.0, 1.5, 7.1, 2.0, 4.3)
_F4 = (3.2, 7.0, 2.5, 6.1, 3.0, 2.3)
_F5 = (4.5, 5.0, 2.5, 2.1, 3.0, 2.3)
_F6 = (1.0, 3.3, 1.5, 1.5, 2.0, 4.0)
_F7 = (4.0, 1.7, 1.5, 6.3, 2.0, 1.4)


def _fill(p):
    return ("fill",) + p


# Per kernel: input name -> recipe, for the baseline family.
BASELINE = {
    "laplace": {"in": ("tracer",)},
    "hdiffsa": {"in": ("tracer",), "mask": ("coef",), "crlato": ("coslat",), "crlatu": ("coslat",)},
    "hdiffsmag": {"uin": ("wind",), "vin": ("wind",), "mask": ("coef",), "crlavo": ("coslat",),
                  "crlavu": ("coslat",), "crlato": ("coslat",), "crlatu": ("coslat",), "acrlat0": ("coef1",)},
    "hadvuv": {"uin": ("wind",), "vin": ("wind",), "acrlat0": ("coef1",), "acrlat1": ("coef1",),
               "tgrlatda0": ("coef1",), "tgrlatda1": ("coef1",)},
    "fastwavesuv": {"uin": ("wind",), "vin": ("wind",), "utens": ("coef",), "vtens": ("coef",),
                    "wgtfac": ("coef",), "ppuv": ("pressure",), "hhl": ("hhl",), "rho": ("coef",),
                    "fx": ("coef1",)},
    "uvbke": {"uc": ("wind",), "vc": ("wind",), "cosa": ("coef",), "rsina": ("coef",)},
    "p_grad_c": {"uin": ("wind",), "vin": ("wind",), "rdxc": ("coef",), "rdyc": ("coef",),
                 "delpc": ("thickness",), "gz": ("geopot",), "pkc": ("pressure",)},
    "nh_p_grad": {"uin": ("wind",), "vin": ("wind",), "rdx": ("coef",), "rdy": ("coef",), "gz": ("geopot",),
                  "pp": ("pressure",), "pk3": ("pressure",), "wk1": ("thickness",)},
    "fvtp2d": {"q": ("tracer",), "crx": ("courant",), "cry": ("courant",), "ra_x": ("coef",), "ra_y": ("coef",),
               "xfx": ("coef",), "yfx": ("coef",), "area": ("coef",)},
}
BASELINE["hadvuv5th"] = dict(BASELINE["hadvuv"])

# The reference drivers' fillMath parameters (src/<name>.cpp).
REFDRIVER = {
    "laplace": {"in": _fill(_F1)},  # src/laplace.cpp:37
    "hdiffsa": {"in": _fill(_F1), "crlato": _fill(_F2), "crlatu": _fill(_F2), "mask": ("const", 0.025)},  # :44-47
    "hdiffsmag": {"uin": _fill(_F1), "vin": _fill(_F2), "crlavo": _fill(_F2), "crlavu": _fill(_F3),
                  "crlato": _fill(_F3), "crlatu": _fill(_F4), "acrlat0": _fill(_F5),
                  "mask": ("const", 0.025)},  # src/hdiffsmag.cpp:54-63
    "hadvuv": {"uin": _fill(_F1), "vin": _fill(_F2), "acrlat0": _fill(_F2), "acrlat1": _fill(_F3),
               "tgrlatda0": _fill(_F3), "tgrlatda1": _fill(_F4)},  # src/hadvuv.cpp:56-62
    "hadvuv5th": {"uin": _fill(_F6), "vin": _fill(_F1), "acrlat0": _fill(_F1), "acrlat1": _fill(_F2),
                  "tgrlatda0": _fill(_F2), "tgrlatda1": _fill(_F3)},  # src/hadvuv5th.cpp:55-61
    "fastwavesuv": {"uin": _fill(_F6), "utens": _fill(_F1), "vin": _fill(_F1), "vtens": _fill(_F2),
                    "ppuv": _fill(_F2), "wgtfac": _fill(_F3), "hhl": _fill(_F3), "rho": _fill(_F4),
                    "fx": _fill(_F5)},  # src/fastwavesuv.cpp:55-63
    "uvbke": {"uc": _fill(_F6), "vc": _fill(_F1), "cosa": _fill(_F7), "rsina": _fill(_F2)},  # src/uvbke.cpp:46-49
    "p_grad_c": {"uin": _fill(_F6), "vin": _fill(_F1), "rdxc": _fill(_F1), "rdyc": _fill(_F2), "delpc": _fill(_F2),
                 "gz": _fill(_F3), "pkc": _fill(_F3)},  # src/p_grad_c.cpp:49-55
    "nh_p_grad": {"uin": _fill(_F6), "vin": _fill(_F1), "rdx": _fill(_F1), "rdy": _fill(_F2), "gz": _fill(_F2),
                  "pp": _fill(_F3), "pk3": _fill(_F3), "wk1": _fill(_F4)},  # src/nh_p_grad.cpp:51-58
    "fvtp2d": {"q": _fill(_F6), "crx": _fill(_F1), "cry": _fill(_F1), "ra_x": _fill(_F2), "ra_y": _fill(_F2),
               "xfx": _fill(_F3), "yfx": _fill(_F3), "area": _fill(_F4)},  # src/fvtp2d.cpp:62-69
}

STRESS = {k: dict(v) for k, v in BASELINE.items()}
STRESS["hdiffsmag"]["acrlat0"] = ("logu1", 0.0, 4.0)

FAMILIES = {"baseline": BASELINE, "refdriver": REFDRIVER, "stress": STRESS}


# ------------------------------------------------------------- 3-D makers
def _tracer(L: Layout, rng: np.random.Generator) -> np.ndarray:
    I, J = L.ij_grid()
    x = I / L.nx
    y = J / L.ny
    kx = rng.integers(1, 5, size=3)
    ky = rng.integers(1, 5, size=3)
    phase = rng.uniform(0.0, 2.0 * math.pi, size=(3, L.levels))
    out = np.zeros((L.levels, L.rows, L.sj))
    for c in range(3):
        arg = 2.0 * math.pi * (kx[c] * x + ky[c] * y)
        out += np.sin(arg[None, :, :] + phase[c][:, None, None])
    out += rng.uniform(-1e-3, 1e-3, size=out.shape)
    return out


def _wind(L: Layout, rng: np.random.Generator) -> np.ndarray:
    n = rng.normal(0.0, 10.0, size=(L.levels, L.rows, L.sj))
    return (n + np.roll(n, 1, 2) + np.roll(n, -1, 2) + np.roll(n, 1, 1) + np.roll(n, -1, 1)) / 5.0


def _z(L: Layout) -> np.ndarray:
    return 250.0 * L.k_levels().astype(np.float64)


def _pressure_profile(L: Layout, rng: np.random.Generator, dz_levels: int = 0) -> np.ndarray:
    eps = rng.normal(0.0, 0.01, size=(L.rows, L.sj))
    z = _z(L) + 250.0 * dz_levels
    prof = 1e5 * np.exp(-z / 8000.0)
    return prof[:, None, None] * (1.0 + eps)[None, :, :]


def _thickness(L: Layout, rng: np.random.Generator) -> np.ndarray:
    eps = rng.normal(0.0, 0.01, size=(L.rows, L.sj))
    z = _z(L)
    prof = 1e5 * (np.exp(-z / 8000.0) - np.exp(-(z + 250.0) / 8000.0))
    return prof[:, None, None] * (1.0 + eps)[None, :, :]


def _geopot(L: Layout, rng: np.random.Generator) -> np.ndarray:
    base = 9.81 * _z(L)
    return base[:, None, None] + 9.81 * rng.normal(0.0, 1.0, size=(L.levels, L.rows, L.sj))


def _hhl(L: Layout, rng: np.random.Generator) -> np.ndarray:
    I, J = L.ij_grid()
    p1, p2 = rng.uniform(0.0, 2.0 * math.pi, size=2)
    t = 0.5 + 0.25 * (np.sin(2.0 * math.pi * I / L.nx + p1) + np.sin(2.0 * math.pi * J / L.ny + p2))
    k = L.k_levels().astype(np.float64)
    return 250.0 * k[:, None, None] * (1.0 + 0.1 * t)[None] + 500.0 * t[None]


def _fill3(L: Layout, a, b, c, d, e, f) -> np.ndarray:
    """fillMath for a 3-D field, include/util.h:150-164, over every allocated point."""
    pi = math.acos(-1.0)
    I, J = L.ij_grid()
    dx = 1.0 / float(L.nx + 2 * L.h)
    dy = 1.0 / float(L.ny + 2 * L.h)
    x = dx * I.astype(np.float64)
    y = dy * J.astype(np.float64)
    plane = a * (b + np.cos(pi * (x + c * y)) + np.sin(d * pi * (x + e * y))) / f
    k = L.k_levels().astype(np.float64)
    return k[:, None, None] * 10e-3 + plane[None, :, :]


def make3(recipe: tuple, L: Layout, rng: np.random.Generator) -> np.ndarray:
    """A flat float64 array of L.size elements built by `recipe`."""
    kind = recipe[0]
    shape = (L.levels, L.rows, L.sj)
    if kind == "tracer":
        a = _tracer(L, rng)
    elif kind == "coef":
        a = rng.uniform(0.01, 0.1, size=shape)
    elif kind == "wind":
        a = _wind(L, rng)
    elif kind == "pressure":
        a = _pressure_profile(L, rng)
    elif kind == "thickness":
        a = _thickness(L, rng)
    elif kind == "geopot":
        a = _geopot(L, rng)
    elif kind == "hhl":
        a = _hhl(L, rng)
    elif kind == "courant":
        a = rng.uniform(-0.4, 0.4, size=shape)
    elif kind == "fill":
        a = _fill3(L, *recipe[1:])
    elif kind == "const":
        a = np.full(shape, float(recipe[1]))
    else:
        raise ValueError(f"unknown 3-D recipe {recipe!r}")
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


def make1(recipe: tuple, L1: Layout1D, n_dom: int, h: int, rng: np.random.Generator) -> np.ndarray:
    """A 1-D float64 array over j (halo included) built by `recipe`."""
    kind = recipe[0]
    j = L1.coords().astype(np.float64)
    if kind == "coslat":
        step = 160.0 / (n_dom - 1) if n_dom > 1 else 0.0
        lat = -80.0 + step * j
        a = np.cos(np.deg2rad(lat))
    elif kind == "coef1":
        a = rng.uniform(0.01, 0.1, size=j.shape)
    elif kind == "logu1":
        a = 10.0 ** rng.uniform(recipe[1], recipe[2], size=j.shape)
    elif kind == "fill":
        # fillMath for a 1-D field, include/util.h:178-188.
        a_, b, c, d, e, f = recipe[1:]
        pi = math.acos(-1.0)
        x = (1.0 / float(n_dom + 2 * h)) * j
        a = a_ * (b + np.cos(pi * (c * x)) + np.sin(d * pi * (e * x))) / f
    else:
        raise ValueError(f"unknown 1-D recipe {recipe!r}")
    return np.ascontiguousarray(a, dtype=np.float64)


def generate(kernel: str, L: Layout, family: str = "baseline", seed: int = 0) -> dict[str, np.ndarray]:
    """Host arrays for every input and output of `kernel` (scratch arguments are
    omitted).  Input `name` at argument position p is seeded by
    SeedSequence([seed, kernel index, p]), so each field has its own stream."""
    spec = KERNELS[kernel]
    recipes = FAMILIES[family][kernel]
    kidx = ORDER.index(kernel)
    L1 = Layout1D(L.ny, L.h)
    out: dict[str, np.ndarray] = {}
    for pos, arg in enumerate(spec.args):
        if arg.kind == "scratch":
            continue
        if arg.kind == "out":
            out[arg.name] = np.full(L.size, np.nan)
            continue
        rng = np.random.default_rng(np.random.SeedSequence([seed, kidx, pos]))
        if arg.kind == "in":
            out[arg.name] = make3(recipes[arg.name], L, rng)
        else:
            out[arg.name] = make1(recipes[arg.name], L1, L.ny, L.h, rng)
    return out


def cast(kernel: str, host: dict, precision: str, islands: bool = True) -> dict[str, np.ndarray]:
    """Copies of `host` in the element types of the given form: unchanged
    float64 for "f64"; for "f32" every array rounded to float32 except the
    FP64 islands of the FP32 form (spec.ISLANDS), unless islands=False (the
    reference's float build, where every array is float32)."""
    spec = KERNELS[kernel]
    out = {}
    for name, arr in host.items():
        keep = precision == "f64" or (islands and spec.arg_precision(name, precision) == "f64")
        out[name] = arr.astype(np.float64 if keep else np.float32)
    return out


# ------------------------------------------------------- on-device generation
def generate_device(kernel: str, L: Layout, seed: int = 0, device="cuda") -> dict:
    """The baseline family's fields of `kernel` generated on the device with
    torch, as flat float64 tensors of L.size (3-D) or Layout1D size (1-D)
    elements; outputs NaN.  Same recipes and distributions as make3 / make1,
    from torch's generator instead of numpy's, so the values differ from
    generate(); used where host generation of large domains would dominate
    (the 512x512x80 benchmark block)."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed * 1000 + ORDER.index(kernel))
    f64 = torch.float64
    shape = (L.levels, L.rows, L.sj)
    i = torch.arange(L.sj, device=device, dtype=f64) - L.lead
    j = torch.arange(L.rows, device=device, dtype=f64) - L.h
    J, I = torch.meshgrid(j, i, indexing="ij")
    k = torch.arange(L.levels, device=device, dtype=f64) - L.hk
    z = 250.0 * k

    def uni(lo, hi, shp=shape):
        return torch.rand(shp, generator=g, device=device, dtype=f64) * (hi - lo) + lo

    def normal(std, shp=shape):
        return torch.randn(shp, generator=g, device=device, dtype=f64) * std

    def tracer():
        out = torch.zeros(shape, device=device, dtype=f64)
        for _ in range(3):
            kx, ky = (torch.randint(1, 5, (2,), generator=g, device=device)).tolist()
            phase = uni(0.0, 2 * math.pi, (L.levels,))
            arg = 2 * math.pi * (kx * I / L.nx + ky * J / L.ny)
            out += torch.sin(arg[None] + phase[:, None, None])
        return out + uni(-1e-3, 1e-3)

    def wind():
        n = normal(10.0)
        return (n + n.roll(1, 2) + n.roll(-1, 2) + n.roll(1, 1) + n.roll(-1, 1)) / 5.0

    def pressure():
        return (1e5 * torch.exp(-z / 8000.0))[:, None, None] * (1.0 + normal(0.01, shape[1:]))[None]

    def thickness():
        prof = 1e5 * (torch.exp(-z / 8000.0) - torch.exp(-(z + 250.0) / 8000.0))
        return prof[:, None, None] * (1.0 + normal(0.01, shape[1:]))[None]

    def geopot():
        return (9.81 * z)[:, None, None] + 9.81 * normal(1.0)

    def hhl():
        p1, p2 = uni(0.0, 2 * math.pi, (2,)).tolist()
        t = 0.5 + 0.25 * (torch.sin(2 * math.pi * I / L.nx + p1) + torch.sin(2 * math.pi * J / L.ny + p2))
        return 250.0 * k[:, None, None] * (1.0 + 0.1 * t)[None] + 500.0 * t[None]

    L1 = Layout1D(L.ny, L.h)
    j1 = torch.arange(L1.size, device=device, dtype=f64) - L1.offset

    def coslat():
        step = 160.0 / (L.ny - 1) if L.ny > 1 else 0.0
        return torch.cos(torch.deg2rad(-80.0 + step * j1))

    makers3 = {"tracer": tracer, "coef": lambda: uni(0.01, 0.1), "wind": wind, "pressure": pressure,
               "thickness": thickness, "geopot": geopot, "hhl": hhl, "courant": lambda: uni(-0.4, 0.4)}
    makers1 = {"coslat": coslat, "coef1": lambda: uni(0.01, 0.1, (L1.size,))}
    out = {}
    for arg in KERNELS[kernel].args:
        if arg.kind == "scratch":
            continue
        if arg.kind == "out":
            out[arg.name] = torch.full((L.size,), float("nan"), device=device, dtype=f64)
            continue
        recipe = BASELINE[kernel][arg.name][0]
        out[arg.name] = (makers3 if arg.kind == "in" else makers1)[recipe]().reshape(-1).contiguous()
    return out
