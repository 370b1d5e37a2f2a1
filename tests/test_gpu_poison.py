"""Stencil reach on the GPU: every input point outside the box the stencil
reads is set to NaN before the call, and every output must still equal the
oracle's output on the unpoisoned inputs bit for bit on its region (so
finite).  A kernel that reads one point beyond its box, or uses a value it
loaded from outside it, turns that output NaN.

The boxes are those of the reference semantics (kernels/<name>.h, the
_openmp variants and fvtp2d's DEFAULT), the same the C-ABI validates each
array against.  The reference leaves halos uninitialised (include/util.h:138)
and never initialises the k-halo planes although fastwavesuv reads planes -1
and nz (kernels/fastwavesuv.h:10-11), so what lies outside each box is
exactly what a kernel must not depend on.  1-D arrays are poisoned outside
their index range too.  Both precisions, both forms (tiles and row march)
of the stencils that have two, and shapes with partial tiles and strips."""
import numpy as np
import pytest

from common import bit_diffs, copy, inputs, out_region
from gpu_util import run_gpu
from oeb.generators import cast
from oeb.layout import Layout1D
from oeb.spec import KERNELS, ORDER
from oracle import binding as B

# (i0, i1, j0, j1, k0, k1) offsets: box [i0, nx + i1) x [j0, ny + j1) x [k0, nz + k1).
IN0 = (0, 0, 0, 0, 0, 0)
BOXES = {
    "laplace": {"in": (-1, 1, -1, 1, 0, 0)},
    "hdiffsa": {"in": (-2, 2, -2, 2, 0, 0), "mask": IN0},
    "hdiffsmag": {"uin": (-1, 1, -1, 1, 0, 0), "vin": (-1, 1, -1, 1, 0, 0), "mask": IN0},
    "hadvuv": {"uin": (-2, 2, -2, 2, 0, 0), "vin": (-2, 2, -2, 2, 0, 0)},
    "hadvuv5th": {"uin": (-3, 3, -3, 3, 0, 0), "vin": (-3, 3, -3, 3, 0, 0)},
    "fastwavesuv": {"uin": IN0, "vin": IN0, "utens": IN0, "vtens": IN0, "wgtfac": (0, 1, 0, 1, 0, 1),
                    "ppuv": (0, 1, 0, 1, -1, 1), "hhl": (0, 1, 0, 1, 0, 1), "rho": (0, 1, 0, 1, 0, 0)},
    "uvbke": {"uc": (0, 0, -1, 0, 0, 0), "vc": (-1, 0, 0, 0, 0, 0), "cosa": IN0, "rsina": IN0},
    "p_grad_c": {"uin": IN0, "vin": IN0, "rdxc": IN0, "rdyc": IN0, "delpc": (-1, 0, -1, 0, 0, 0),
                 "gz": (-1, 0, -1, 0, 0, 1), "pkc": (-1, 0, -1, 0, 0, 1)},
    "nh_p_grad": {"uin": IN0, "vin": IN0, "rdx": IN0, "rdy": IN0, "gz": (0, 1, 0, 1, 0, 1),
                  "pp": (0, 1, 0, 1, 0, 1), "pk3": (0, 1, 0, 1, 0, 1), "wk1": (0, 1, 0, 1, 0, 0)},
    "fvtp2d": {"q": (-3, 3, -3, 3, 0, 0), "crx": (0, 1, -3, 2, 0, 0), "cry": (-3, 2, 0, 1, 0, 0),
               "ra_x": (0, 0, -3, 2, 0, 0), "ra_y": (-3, 2, 0, 0, 0, 0), "xfx": (0, 1, -3, 2, 0, 0),
               "yfx": (-3, 2, 0, 1, 0, 0), "area": (-3, 2, -3, 2, 0, 0)},
}
# 1-D arrays: index range [j0, ny + j1).
BOXES_1D = {
    "hdiffsa": {"crlato": (-1, 1), "crlatu": (-1, 1)},
    "hdiffsmag": {"crlavo": (0, 0), "crlavu": (0, 0), "crlato": (0, 0), "crlatu": (0, 0), "acrlat0": (-1, 1)},
    "hadvuv": {n: (0, 0) for n in ("acrlat0", "acrlat1", "tgrlatda0", "tgrlatda1")},
    "hadvuv5th": {n: (0, 0) for n in ("acrlat0", "acrlat1", "tgrlatda0", "tgrlatda1")},
    "fastwavesuv": {"fx": (0, 0)},
}
TWO_FORMS = {"laplace", "hdiffsa", "hdiffsmag", "fvtp2d"}


def test_boxes_cover_every_input():
    for name in ORDER:
        spec = KERNELS[name]
        assert set(BOXES[name]) == set(spec.inputs3), name
        assert set(BOXES_1D.get(name, {})) == set(spec.inputs1), name


def poison(name, host, L):
    """Copy of `host` with NaN at every input point outside the read boxes."""
    out = copy(host)
    for a, (i0, i1, j0, j1, k0, k1) in BOXES[name].items():
        v = L.view(out[a])
        keep = np.zeros(v.shape, dtype=bool)
        keep[k0 + L.hk: L.nz + k1 + L.hk, j0 + L.h: L.ny + j1 + L.h, i0 + L.lead: L.nx + i1 + L.lead] = True
        v[~keep] = np.nan
    L1 = Layout1D(L.ny, L.h)
    for a, (j0, j1) in BOXES_1D.get(name, {}).items():
        keep = np.zeros(L1.size, dtype=bool)
        keep[j0 + L1.offset: L.ny + j1 + L1.offset] = True
        out[a][~keep] = np.nan
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["default", "other"])
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(67, 9, 3), (130, 36, 4), (256, 20, 2)], ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("name", ORDER)
def test_poisoned_halo(name, shape, precision, form, knobs):
    if form == "other":
        if name not in TWO_FORMS:
            pytest.skip("one form only")
        # The form the stencil does not take by default in this precision.
        default_strip = {"laplace": False, "fvtp2d": precision == "f32"}.get(name, True)
        knobs.set(name, "form", 0 if default_strip else 1)
    L, host = inputs(name, *shape, "baseline", seed=41)
    host = cast(name, host, precision)
    print(f"[gpu] poisoned halo {name} {shape} {precision} {form}", flush=True)
    ref = copy(host)
    B.run_oracle(name, ref, L, precision=precision)
    gpu = run_gpu(name, poison(name, host, L), L)
    for o in KERNELS[name].outputs:
        x, y = out_region(name, o, L, gpu[o]), out_region(name, o, L, ref[o])
        assert np.isfinite(x).all(), f"{name}.{o}: NaN from a point outside the read boxes"
        assert bit_diffs(x, y) == 0, f"{name}.{o}: differs from the oracle on unpoisoned inputs"
