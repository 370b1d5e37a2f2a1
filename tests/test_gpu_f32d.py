"""GPU parity of the deviation-island FP32 form (oeb_fastwavesuv_f32d,
oeb_p_grad_c_f32d, oeb_nh_p_grad_f32d) through the C-ABI on cuda:0.

Bar: bit for bit equal to the oracle's FP32 form (oracle_<name>_f32) run on
the islands the entry point reads, base(k) + dev rebuilt in FP64 on the host
(oeb.islands.reconstruct, the same IEEE sum); points outside each output's
region keep their NaN fill.  At 256x256x80 and 512x512x80 the error against
the FP64 oracle on the unrounded inputs is bounded like that of the
FP64-island form (relative max-norm and RMS <= 1e-6)."""
import numpy as np
import pytest

from common import bit_diffs, copy, inputs, out_region, outside_mask, rel_maxnorm, rel_rms
from gpu_util import to_device
from oeb import stencils
from oeb.generators import cast
from oeb.islands import DEVIATION_STENCILS, reconstruct, split_host
from oeb.spec import KERNELS
from oeb.storage import DeviationIsland, Storage1D, Storage3D, level_layout
from oracle import binding as B

SHAPES = [(12, 12, 5, 16), (20, 20, 3, 1), (24, 17, 4, 16), (65, 3, 2, 1), (1, 1, 1, 16), (64, 64, 60, 16)]


def _run(name, host32, L):
    """Splits the islands of `host32` (float64), runs the _f32d entry point,
    and checks it against the FP32 oracle on the rebuilt islands.  Returns
    the GPU outputs."""
    import torch

    rebuilt = copy(host32)
    arrays = {}
    for a in KERNELS[name].args:
        if a.kind == "scratch":
            continue
        if a.name in DEVIATION_STENCILS[name]:
            d, base = split_host(L, host32[a.name])
            rebuilt[a.name] = reconstruct(L, d, base)
            arrays[a.name] = DeviationIsland(Storage3D.to_device(L, d), Storage1D.to_device(level_layout(L), base))
    for k, v in to_device(name, host32, L).items():
        arrays.setdefault(k, v)
    stencils.run(name, arrays, domain=(L.nx, L.ny, L.nz))
    torch.cuda.synchronize()
    gpu = {o: arrays[o].to_host() for o in KERNELS[name].outputs}
    ref = copy(rebuilt)
    B.run_oracle(name, ref, L, precision="f32")
    for o in KERNELS[name].outputs:
        x, y = out_region(name, o, L, gpu[o]), out_region(name, o, L, ref[o])
        assert x.dtype == np.float32 and np.isfinite(y).all()
        nd = bit_diffs(x, y)
        assert nd == 0, f"{name}.{o}: {nd} values differ from the FP32 oracle (rel {rel_maxnorm(x, y):.2e})"
        outside = gpu[o].reshape(L.levels, L.rows, L.sj)[outside_mask(name, o, L)]
        assert np.isnan(outside).all(), f"{name}.{o}: wrote outside its region"
    return gpu


@pytest.mark.gpu
@pytest.mark.parametrize("pairs", [0, 1])
@pytest.mark.parametrize("family", ["baseline", "refdriver"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s[:3])) + f"a{s[3]}")
@pytest.mark.parametrize("name", sorted(DEVIATION_STENCILS))
def test_f32d_parity_small(name, shape, family, pairs, knobs):
    knobs.set(name, "vector", pairs)
    nx, ny, nz, align = shape
    L, host = inputs(name, nx, ny, nz, family, seed=21, align=align)
    print(f"[gpu] f32d {name} {family} {nx}x{ny}x{nz} align={align} pairs={pairs}", flush=True)
    _run(name, cast(name, host, "f32"), L)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [256, 512])
@pytest.mark.parametrize("name", sorted(DEVIATION_STENCILS))
def test_f32d_against_f64(name, n):
    L, host = inputs(name, n, n, 80, "baseline", seed=2)
    print(f"[gpu] f32d {name} {n}x{n}x80", flush=True)
    gpu = _run(name, cast(name, host, "f32"), L)
    exact = copy(host)
    B.run_oracle(name, exact, L)
    for o in KERNELS[name].outputs:
        x, y = out_region(name, o, L, gpu[o]), out_region(name, o, L, exact[o])
        err, rms = rel_maxnorm(x, y), rel_rms(x, y)
        print(f"  {name}.{o} {n}^2: FP32 deviation islands vs FP64 rel max {err:.2e} rms {rms:.2e}", flush=True)
        assert err <= 1e-6 and rms <= 1e-6, (name, o, err, rms)


@pytest.mark.gpu
@pytest.mark.parametrize("kchunk", [1, 3, 12])
@pytest.mark.parametrize("tile", [0, 1])
@pytest.mark.parametrize("shape", [(64, 36, 7), (130, 17, 5)], ids=lambda s: "x".join(map(str, s)))
def test_f32d_fastwavesuv_tile_rows(shape, tile, kchunk, knobs):
    """fastwavesuv's deviation-island form in column-pair tiles of 4 and 8
    rows (tile knobs 0 and 1), with level chunks that divide nz, do not
    divide it, and exceed it, agrees bit for bit with the FP32 oracle."""
    knobs.set("fastwavesuv", "tile", tile)
    knobs.set("fastwavesuv", "kchunk", kchunk)
    L, host = inputs("fastwavesuv", *shape, "baseline", seed=33)
    print(f"[gpu] f32d fastwavesuv tile={tile} kchunk={kchunk} {shape}", flush=True)
    _run("fastwavesuv", cast("fastwavesuv", host, "f32"), L)


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [0, 1])
@pytest.mark.parametrize("pairs", [0, 1])
@pytest.mark.parametrize("shape", [(67, 9, 3), (130, 36, 4)], ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("name", sorted(DEVIATION_STENCILS))
def test_f32d_poisoned_halo(name, shape, pairs, tile, knobs):
    """Every deviation and input point outside the stencil's read box, and
    each base profile's level -1 where the island's box starts at level 0,
    set to NaN: the outputs still equal the FP32 oracle on the unpoisoned
    inputs bit for bit (fastwavesuv also in 8-row tiles, tile knob 1)."""
    import torch

    from oeb.layout import Layout1D
    from reach import BOXES, BOXES_1D, keep_mask, keep_mask_1d

    if tile and name != "fastwavesuv":
        pytest.skip("tile rows: fastwavesuv only")
    knobs.set(name, "vector", pairs)
    knobs.set(name, "tile", tile)
    L, host = inputs(name, *shape, "baseline", seed=43)
    host32 = cast(name, host, "f32")
    rebuilt = copy(host32)
    arrays = {}
    for a in KERNELS[name].args:
        if a.kind == "scratch":
            continue
        if a.name in DEVIATION_STENCILS[name]:
            d, base = split_host(L, host32[a.name])
            rebuilt[a.name] = reconstruct(L, d, base)
            L.view(d)[~keep_mask(L, BOXES[name][a.name])] = np.nan
            base = base.copy()
            if BOXES[name][a.name][4] >= 0:
                base[0] = np.nan  # level -1, outside the island's levels
            arrays[a.name] = DeviationIsland(Storage3D.to_device(L, d), Storage1D.to_device(level_layout(L), base))
        elif a.kind == "in1":
            h = host32[a.name].copy()
            h[~keep_mask_1d(L, BOXES_1D[name][a.name])] = np.nan
            arrays[a.name] = Storage1D.to_device(Layout1D(L.ny, L.h), h)
        else:
            h = host32[a.name].copy()
            if a.kind == "in":
                L.view(h)[~keep_mask(L, BOXES[name][a.name])] = np.nan
            arrays[a.name] = Storage3D.to_device(L, h)
    print(f"[gpu] f32d poisoned halo {name} {shape} pairs={pairs}", flush=True)
    stencils.run(name, arrays, domain=(L.nx, L.ny, L.nz))
    torch.cuda.synchronize()
    ref = copy(rebuilt)
    B.run_oracle(name, ref, L, precision="f32")
    for o in KERNELS[name].outputs:
        x, y = out_region(name, o, L, arrays[o].to_host()), out_region(name, o, L, ref[o])
        assert np.isfinite(x).all(), f"{name}.{o}: NaN from a point outside the read boxes"
        assert bit_diffs(x, y) == 0
