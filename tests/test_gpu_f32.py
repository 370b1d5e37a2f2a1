"""GPU parity of the FP32 form (oeb_<name>_f32) through the C-ABI on cuda:0.

Bar: bit for bit equal to the oracle's FP32 form (oracle_<name>_f32) on the
same float32 inputs (FP64 islands float64), with no contraction on either
side.  Points outside each output's region keep their NaN fill.

At 256x256x80 and at the FP32 workload's own size, 512x512x80, the error
against the FP64 oracle on the unrounded inputs is also bounded, per output,
with a margin over what was measured (F32_BOUNDS): relative max-norm and RMS
<= 1e-6 for most outputs.  Two effects of the inputs, not of the
arithmetic, need wider bounds, so the 1e-5 target is not met by these
outputs at these sizes:
  laplace.out        rounding `in` to float32 (u * max|in| ~ 6e-8) against
                     an output ~100x (256^2) to ~400x (512^2) smaller than its
                     input (the Laplacian of a smooth field scales with the
                     grid spacing squared): 4.7e-5 max measured at 256^2,
                     1.2e-4 at 512^2.
  hdiffsa, fvtp2d    rounding flips a limiter decision (hdiffsa's flux
                     limiter; fvtp2d's almq*br < 0) at isolated points, where
                     the output jumps: up to 4e-2 max at ~1e-4 of the points.
                     These outputs are bounded in RMS and in the fraction of
                     points off by more than 1e-5 of the output's max-norm.
"""
import numpy as np
import pytest

from common import bit_diffs, copy, inputs, out_region, outside_mask, rel_maxnorm, rel_rms
from gpu_util import run_gpu
from oeb.generators import cast
from oeb.spec import KERNELS, ORDER
from oracle import binding as B

# (kernel, output) -> (relative max-norm, relative RMS, fraction of points off
# by more than 1e-5 of max|FP64|) bounds at 256x256x80.
DEFAULT_BOUND = (1e-6, 1e-6, 0.0)
LIMITER_BOUND = (1e-1, 1e-3, 1e-3)
F32_BOUNDS = {("laplace", "out"): (2e-4, 1e-4, 1.0), ("hdiffsa", "out"): LIMITER_BOUND}
F32_BOUNDS.update({("fvtp2d", o): LIMITER_BOUND for o in KERNELS["fvtp2d"].outputs})
F32_BOUNDS_512 = {**F32_BOUNDS, ("laplace", "out"): (5e-4, 3e-4, 1.0)}
SMALL = [(12, 12, 5, 16), (20, 20, 3, 1), (24, 17, 4, 16), (65, 3, 2, 1), (1, 1, 1, 16), (64, 64, 60, 16)]


def _check_exact(name, L, host32):
    ref = copy(host32)
    B.run_oracle(name, ref, L, precision="f32")
    gpu = run_gpu(name, copy(host32), L)
    for o in KERNELS[name].outputs:
        x, y = out_region(name, o, L, gpu[o]), out_region(name, o, L, ref[o])
        assert x.dtype == np.float32
        assert np.isfinite(y).all(), f"{name}.{o}: oracle produced non-finite values"
        nd = bit_diffs(x, y)
        assert nd == 0, f"{name}.{o}: {nd} values differ from the FP32 oracle (rel {rel_maxnorm(x, y):.2e})"
        outside = gpu[o].reshape(L.levels, L.rows, L.sj)[outside_mask(name, o, L)]
        assert np.isnan(outside).all(), f"{name}.{o}: wrote {np.count_nonzero(~np.isnan(outside))} points outside"
    return gpu


@pytest.mark.gpu
@pytest.mark.parametrize("family", ["baseline", "refdriver", "stress"])
@pytest.mark.parametrize("shape", SMALL, ids=lambda s: "x".join(map(str, s[:3])) + f"a{s[3]}")
@pytest.mark.parametrize("name", ORDER)
def test_f32_parity_small(name, shape, family):
    nx, ny, nz, align = shape
    L, host = inputs(name, nx, ny, nz, family, seed=12, align=align)
    print(f"[gpu] f32 {name} {family} {nx}x{ny}x{nz} align={align}", flush=True)
    _check_exact(name, L, cast(name, host, "f32"))


@pytest.mark.gpu
@pytest.mark.parametrize("n", [256, 512])
@pytest.mark.parametrize("name", ORDER)
def test_f32_against_f64(name, n):
    """n x n x 80: exact against the FP32 oracle, and within the bounds of
    F32_BOUNDS (256) or F32_BOUNDS_512 of the FP64 oracle."""
    L, host = inputs(name, n, n, 80, "baseline", seed=2)
    print(f"[gpu] f32 {name} {n}x{n}x80", flush=True)
    gpu = _check_exact(name, L, cast(name, host, "f32"))
    exact = copy(host)
    B.run_oracle(name, exact, L)
    bounds = F32_BOUNDS if n == 256 else F32_BOUNDS_512
    for o in KERNELS[name].outputs:
        x, y = out_region(name, o, L, gpu[o]), out_region(name, o, L, exact[o])
        err, rms = rel_maxnorm(x, y), rel_rms(x, y)
        frac = float(np.mean(np.abs(x.astype(np.float64) - y) > 1e-5 * np.max(np.abs(y))))
        print(f"  {name}.{o} {n}^2: FP32 vs FP64 rel max {err:.2e} rms {rms:.2e} frac>1e-5 {frac:.1e}", flush=True)
        bmax, brms, bfrac = bounds.get((name, o), DEFAULT_BOUND)
        assert err <= bmax and rms <= brms and frac <= bfrac, (
            f"{name}.{o}: max {err:.2e} (bound {bmax:.0e}), rms {rms:.2e} ({brms:.0e}), frac {frac:.1e} ({bfrac:.0e})")


@pytest.mark.gpu
@pytest.mark.parametrize("nx", [64, 66, 67])
@pytest.mark.parametrize("pairs", [0, 1])
@pytest.mark.parametrize("name", ["laplace", "uvbke", "fastwavesuv", "p_grad_c", "nh_p_grad"])
def test_f32_vector_path_switch(name, pairs, nx, knobs):
    """The vector paths of the FP32 form (4-wide for laplace and uvbke,
    column pairs with FP64 island pairs for the others), forced on or off
    with the vector knob, and nx where they do not apply (66 for the 4-wide
    paths, 67 for all), agree bit for bit with the FP32 oracle."""
    knobs.set(name, "vector", pairs)
    L, host = inputs(name, nx, 36, 7, "baseline", seed=9)
    print(f"[gpu] f32 {name} pairs={pairs} nx={nx}", flush=True)
    _check_exact(name, L, cast(name, host, "f32"))


@pytest.mark.gpu
@pytest.mark.parametrize("rows", [1, 5, 64])
@pytest.mark.parametrize("strip", [0, 1])
@pytest.mark.parametrize("shape", [(64, 36, 3), (1030, 9, 2)], ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("name", ["hdiffsa", "hdiffsmag", "fvtp2d"])
def test_f32_strip_switches(name, shape, strip, rows, knobs):
    """The FP32 row-march (strip) and tile paths, forced through
    the form knob, with row chunks of one row, of a number that does not
    divide ny, and of more than ny, agree bit for bit with the FP32 oracle."""
    knobs.set(name, "form", strip)
    knobs.set(name, "rows", rows)
    L, host = inputs(name, *shape, "baseline", seed=14)
    print(f"[gpu] f32 {name} strip={strip} rows={rows} {shape}", flush=True)
    _check_exact(name, L, cast(name, host, "f32"))


@pytest.mark.gpu
@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("lanes", [64, 128, 512, 1024])
def test_f32_fvtp2d_strip_shapes(lanes, depth, knobs):
    """fvtp2d's FP32 strips of at most 64 .. 1024 lanes (one to five strips
    across nx = 300), with one and two rows loaded ahead, agree bit for bit
    with the FP32 oracle."""
    knobs.set("fvtp2d", "form", 1)
    knobs.set("fvtp2d", "lanes", lanes)
    knobs.set("fvtp2d", "depth", depth)
    L, host = inputs("fvtp2d", 300, 23, 2, "stress", seed=15)
    print(f"[gpu] f32 fvtp2d strip lanes={lanes} depth={depth}", flush=True)
    _check_exact("fvtp2d", L, cast("fvtp2d", host, "f32"))


@pytest.mark.gpu
@pytest.mark.parametrize("kchunk", [1, 3, 12])
@pytest.mark.parametrize("tile", [0, 1])
@pytest.mark.parametrize("shape", [(64, 36, 7), (130, 17, 5), (2, 1, 1)], ids=lambda s: "x".join(map(str, s)))
def test_f32_fastwavesuv_tile_rows(shape, tile, kchunk, knobs):
    """fastwavesuv's FP32 column-pair tiles of 4 and 8 rows (tile knobs 0
    and 1), on whole and partial tiles, with level chunks that divide nz, do
    not divide it, and exceed it, agree bit for bit with the FP32 oracle."""
    knobs.set("fastwavesuv", "tile", tile)
    knobs.set("fastwavesuv", "kchunk", kchunk)
    L, host = inputs("fastwavesuv", *shape, "stress", seed=29)
    print(f"[gpu] f32 fastwavesuv tile={tile} kchunk={kchunk} {shape}", flush=True)
    _check_exact("fastwavesuv", L, cast("fastwavesuv", host, "f32"))


@pytest.mark.gpu
@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("tile", [13, 14, 15, 16, 17])
@pytest.mark.parametrize("shape", [(64, 36, 3), (67, 9, 2), (130, 17, 4), (128, 16, 8), (1, 1, 1)],
                         ids=lambda s: "x".join(map(str, s)))
def test_f32_fvtp2d_hoisted_tiles(shape, tile, order, knobs):
    """fvtp2d's hoisted tiles in FP32 (tile knobs 13-17; 14 is the
    default), in two tile orders, on whole and partial tiles and one-point
    domains, agree bit for bit with the FP32 oracle."""
    knobs.set("fvtp2d", "form", 0)
    knobs.set("fvtp2d", "tile", tile)
    knobs.set("fvtp2d", "order", order)
    L, host = inputs("fvtp2d", *shape, "stress", seed=21)
    print(f"[gpu] f32 fvtp2d hoisted tile={tile} order={order} {shape}", flush=True)
    _check_exact("fvtp2d", L, cast("fvtp2d", host, "f32"))


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [0, 1, 2])
@pytest.mark.parametrize("shape", [(64, 36, 3), (67, 9, 2), (130, 17, 4)], ids=lambda s: "x".join(map(str, s)))
def test_f32_fvtp2d_lds_tiles(shape, tile, knobs):
    """fvtp2d's LDS-staged tiles (form knob 2) in FP32, each tile shape, on
    whole and partial tiles, agree bit for bit with the FP32 oracle."""
    knobs.set("fvtp2d", "form", 2)
    knobs.set("fvtp2d", "tile", tile)
    L, host = inputs("fvtp2d", *shape, "stress", seed=19)
    print(f"[gpu] f32 fvtp2d lds tile={tile} {shape}", flush=True)
    _check_exact("fvtp2d", L, cast("fvtp2d", host, "f32"))


@pytest.mark.gpu
@pytest.mark.parametrize("xcd", [0, 1])
@pytest.mark.parametrize("shape", [(64, 36, 8), (200, 9, 4), (256, 20, 6)], ids=lambda s: "x".join(map(str, s)))
def test_f32_laplace_xcd_interleave(shape, xcd, knobs):
    """FP32 laplace with and without the level-chunk interleave agrees bit
    for bit with the FP32 oracle."""
    knobs.set("laplace", "order", xcd)
    L, host = inputs("laplace", *shape, "baseline", seed=23)
    print(f"[gpu] f32 laplace xcd={xcd} {shape}", flush=True)
    _check_exact("laplace", L, cast("laplace", host, "f32"))
