"""Stencil reach on the GPU: every input point outside the box the stencil
reads is set to NaN before the call, and every output must still equal the
oracle's output on the unpoisoned inputs bit for bit on its region (so
finite).  A kernel that reads one point beyond its box, or uses a value it
loaded from outside it, turns that output NaN.

The boxes (tests/reach.py) are those of the reference semantics; what lies
outside them the reference leaves uninitialised.  1-D arrays are poisoned
outside their index range too.  Both precisions, both forms (tiles and row
march) of the stencils that have two, and shapes with partial tiles and
strips."""
import numpy as np
import pytest

from common import bit_diffs, copy, inputs, out_region
from gpu_util import run_gpu
from oeb.generators import cast
from oeb.spec import KERNELS, ORDER
from oracle import binding as B
from reach import BOXES, BOXES_1D, keep_mask, keep_mask_1d

TWO_FORMS = {"laplace", "hdiffsa", "hdiffsmag", "fvtp2d"}


def test_boxes_cover_every_input():
    for name in ORDER:
        spec = KERNELS[name]
        assert set(BOXES[name]) == set(spec.inputs3), name
        assert set(BOXES_1D.get(name, {})) == set(spec.inputs1), name


def poison(name, host, L):
    """Copy of `host` with NaN at every input point outside the read boxes."""
    out = copy(host)
    for a, box in BOXES[name].items():
        L.view(out[a])[~keep_mask(L, box)] = np.nan
    for a, box in BOXES_1D.get(name, {}).items():
        out[a][~keep_mask_1d(L, box)] = np.nan
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
        default_strip = {"laplace": False, "fvtp2d": False}.get(name, True)
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


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [13, 14, 15, 16, 17])
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(67, 9, 3), (130, 36, 4), (256, 20, 2)], ids=lambda s: "x".join(map(str, s)))
def test_poisoned_halo_fvtp2d_hoisted(shape, precision, tile, knobs):
    """fvtp2d's hoisted tiles (tile knobs 13-17) load each region's operands
    at addresses clamped into the region; with NaN outside the read boxes,
    every output still equals the oracle's bit for bit."""
    knobs.set("fvtp2d", "form", 0)
    knobs.set("fvtp2d", "tile", tile)
    L, host = inputs("fvtp2d", *shape, "baseline", seed=47)
    host = cast("fvtp2d", host, precision)
    print(f"[gpu] poisoned halo fvtp2d hoisted tile={tile} {shape} {precision}", flush=True)
    ref = copy(host)
    B.run_oracle("fvtp2d", ref, L, precision=precision)
    gpu = run_gpu("fvtp2d", poison("fvtp2d", host, L), L)
    for o in KERNELS["fvtp2d"].outputs:
        x, y = out_region("fvtp2d", o, L, gpu[o]), out_region("fvtp2d", o, L, ref[o])
        assert np.isfinite(x).all(), f"fvtp2d.{o}: NaN from a point outside the read boxes"
        assert bit_diffs(x, y) == 0, f"fvtp2d.{o}: differs from the oracle on unpoisoned inputs"


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(66, 9, 3), (130, 36, 4), (256, 20, 2)], ids=lambda s: "x".join(map(str, s)))
def test_poisoned_halo_fastwavesuv_tall(shape, precision, knobs):
    """fastwavesuv's 8-row column-pair tiles (tile knob 1) read nothing
    outside the read boxes: with NaN there, every output still equals the
    oracle's bit for bit."""
    knobs.set("fastwavesuv", "tile", 1)
    L, host = inputs("fastwavesuv", *shape, "baseline", seed=49)
    host = cast("fastwavesuv", host, precision)
    print(f"[gpu] poisoned halo fastwavesuv tall {shape} {precision}", flush=True)
    ref = copy(host)
    B.run_oracle("fastwavesuv", ref, L, precision=precision)
    gpu = run_gpu("fastwavesuv", poison("fastwavesuv", host, L), L)
    for o in KERNELS["fastwavesuv"].outputs:
        x, y = out_region("fastwavesuv", o, L, gpu[o]), out_region("fastwavesuv", o, L, ref[o])
        assert np.isfinite(x).all(), f"fastwavesuv.{o}: NaN from a point outside the read boxes"
        assert bit_diffs(x, y) == 0, f"fastwavesuv.{o}: differs from the oracle on unpoisoned inputs"


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [0, 1])
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(67, 9, 3), (130, 36, 4), (256, 20, 2)], ids=lambda s: "x".join(map(str, s)))
def test_poisoned_halo_fvtp2d_lds(shape, precision, tile, knobs):
    """fvtp2d's LDS-staged tiles (form knob 2) load every input region with
    its rows and columns clamped into the array's box; with NaN outside the
    read boxes, every output still equals the oracle's bit for bit."""
    knobs.set("fvtp2d", "form", 2)
    knobs.set("fvtp2d", "tile", tile)
    L, host = inputs("fvtp2d", *shape, "baseline", seed=43)
    host = cast("fvtp2d", host, precision)
    print(f"[gpu] poisoned halo fvtp2d lds tile={tile} {shape} {precision}", flush=True)
    ref = copy(host)
    B.run_oracle("fvtp2d", ref, L, precision=precision)
    gpu = run_gpu("fvtp2d", poison("fvtp2d", host, L), L)
    for o in KERNELS["fvtp2d"].outputs:
        x, y = out_region("fvtp2d", o, L, gpu[o]), out_region("fvtp2d", o, L, ref[o])
        assert np.isfinite(x).all(), f"fvtp2d.{o}: NaN from a point outside the read boxes"
        assert bit_diffs(x, y) == 0, f"fvtp2d.{o}: differs from the oracle on unpoisoned inputs"
