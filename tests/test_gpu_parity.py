"""GPU parity: every stencil through the C-ABI on cuda:0 against the C oracle
on the same seeded inputs.

Bar: relative max-norm (max|gpu - oracle| / max|oracle|) <= 1e-12 on each
output's valid region.  Both sides compile without FMA contraction, so the
results are expected to agree bit for bit.  test_bit_exact asserts that.
Points outside each output's region must keep their NaN fill, which shows
that the kernel writes nothing outside its region.
"""
import numpy as np
import pytest

from common import bit_diffs, copy, inputs, out_region, outside_mask, rel_maxnorm
from gpu_util import run_gpu
from oeb.spec import KERNELS, ORDER
from oracle import binding as B

TOL = 1e-12

# (nx, ny, nz, align): square, reference layout (align 1), rectangular, one
# row, one point, and the 64x64x60 shape of the hdiffsa benchmark.
SMALL = [(12, 12, 5, 16), (20, 20, 3, 1), (24, 17, 4, 16), (65, 3, 2, 1), (1, 1, 1, 16), (64, 64, 60, 16)]


def _check(name, L, host, gpu, tol=TOL, require_exact=False):
    ref = copy(host)
    B.run_oracle(name, ref, L)
    report = []
    for o in KERNELS[name].outputs:
        x = out_region(name, o, L, gpu[o])
        y = out_region(name, o, L, ref[o])
        assert np.isfinite(y).all(), f"{name}.{o}: oracle produced non-finite values"
        assert np.isfinite(x).all(), f"{name}.{o}: GPU left unwritten (NaN) points in its region"
        err = rel_maxnorm(x, y)
        nd = bit_diffs(x, y)
        report.append((o, err, nd))
        assert err <= tol, f"{name}.{o}: rel max-norm {err:.3e} > {tol:.0e} ({nd} values differ)"
        if require_exact:
            assert nd == 0, f"{name}.{o}: {nd} values differ from the oracle (rel {err:.3e})"
        outside = gpu[o].reshape(L.levels, L.rows, L.sj)[outside_mask(name, o, L)]
        assert np.isnan(outside).all(), f"{name}.{o}: wrote {np.count_nonzero(~np.isnan(outside))} points outside"
    return report


@pytest.mark.gpu
@pytest.mark.parametrize("family", ["baseline", "refdriver", "stress"])
@pytest.mark.parametrize("shape", SMALL, ids=lambda s: "x".join(map(str, s[:3])) + f"a{s[3]}")
@pytest.mark.parametrize("name", ORDER)
def test_parity_small(name, shape, family):
    nx, ny, nz, align = shape
    L, host = inputs(name, nx, ny, nz, family, seed=11, align=align)
    print(f"[gpu] {name} {family} {nx}x{ny}x{nz} align={align}", flush=True)
    gpu = run_gpu(name, copy(host), L)
    _check(name, L, host, gpu)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ORDER)
def test_bit_exact(name):
    """Without contraction on either side, the GPU reproduces the oracle's
    doubles exactly (division and sqrt are correctly rounded on both)."""
    L, host = inputs(name, 48, 40, 6, "baseline", seed=5)
    print(f"[gpu] bit-exact {name}", flush=True)
    _check(name, L, host, run_gpu(name, copy(host), L), require_exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ORDER)
def test_parity_256(name):
    """The all-stencil workload shape: 256x256x80, halo 3, k-halo 1."""
    L, host = inputs(name, 256, 256, 80, "baseline", seed=2)
    print(f"[gpu] {name} 256x256x80", flush=True)
    _check(name, L, host, run_gpu(name, copy(host), L))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ORDER)
def test_parity_512(name):
    """The 512x512x80 shape of the bandwidth target."""
    L, host = inputs(name, 512, 512, 80, "baseline", seed=3)
    print(f"[gpu] {name} 512x512x80", flush=True)
    _check(name, L, host, run_gpu(name, copy(host), L))


PAIRED = ["laplace", "uvbke", "fastwavesuv", "p_grad_c", "nh_p_grad"]


@pytest.mark.gpu
@pytest.mark.parametrize("kchunk", [1, 3, 80])
@pytest.mark.parametrize("pairs", [0, 1])
@pytest.mark.parametrize("name", PAIRED)
def test_path_switches(name, pairs, kchunk, knobs):
    """Both the scalar and the paired (16-byte) path, forced through
    the vector knob, at k-chunks that divide nz, do not divide it, and exceed
    it, agree bit for bit with the oracle."""
    knobs.set(name, "vector", pairs)
    knobs.set(name, "kchunk", kchunk)
    L, host = inputs(name, 64, 36, 7, "baseline", seed=9)
    print(f"[gpu] {name} pairs={pairs} kchunk={kchunk}", flush=True)
    _check(name, L, host, run_gpu(name, copy(host), L), require_exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("kchunk", [1, 3, 16])
@pytest.mark.parametrize("tile", [0, 1])
@pytest.mark.parametrize("family", ["baseline", "stress"])
@pytest.mark.parametrize("shape", [(64, 36, 7), (130, 17, 5), (2, 1, 1), (300, 8, 3)],
                         ids=lambda s: "x".join(map(str, s)))
def test_fastwavesuv_tile_rows(shape, family, tile, kchunk, knobs):
    """fastwavesuv's column-pair tiles of 4 rows (tile knob 0) and of 8 rows
    (tile knob 1, the default on levels of more than 256x256 points), on
    whole and partial tiles in i and j and a one-pair domain, with level
    chunks that divide nz, do not divide it, and exceed it, agree bit for bit
    with the oracle."""
    knobs.set("fastwavesuv", "tile", tile)
    knobs.set("fastwavesuv", "kchunk", kchunk)
    L, host = inputs("fastwavesuv", *shape, family, seed=27)
    print(f"[gpu] fastwavesuv tile={tile} kchunk={kchunk} {family} {shape}", flush=True)
    _check("fastwavesuv", L, host, run_gpu("fastwavesuv", copy(host), L), require_exact=True)


STRIPPED = ["laplace", "hdiffsa", "hdiffsmag", "fvtp2d"]


@pytest.mark.gpu
@pytest.mark.parametrize("rows", [1, 5, 64])
@pytest.mark.parametrize("strip", [0, 1])
@pytest.mark.parametrize("shape", [(64, 36, 3), (1030, 9, 2), (2100, 5, 2)], ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("name", STRIPPED)
def test_strip_switches(name, shape, strip, rows, knobs):
    """The row-march (strip) path and the tile path, forced through
    the form knob, with row chunks of one row, of a number that does not
    divide ny, and of more than ny, and with nx wide enough for two or three
    strips, agree bit for bit with the oracle."""
    knobs.set(name, "form", strip)
    knobs.set(name, "rows", rows)
    L, host = inputs(name, *shape, "baseline", seed=13)
    print(f"[gpu] {name} strip={strip} rows={rows} {shape}", flush=True)
    _check(name, L, host, run_gpu(name, copy(host), L), require_exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("family", ["baseline", "stress"])
@pytest.mark.parametrize("tile", [8, 13, 14, 15, 16, 17])
@pytest.mark.parametrize("shape", [(64, 16, 2), (64, 36, 3), (67, 9, 2), (130, 17, 4), (1, 1, 1), (256, 256, 3)],
                         ids=lambda s: "x".join(map(str, s)))
def test_fvtp2d_hoisted_tiles(shape, tile, family, knobs):
    """fvtp2d's register-staged tiles: the per-element loop (tile knob 8) and
    the hoisted tiles (13: each region's loads at the start of its phase, 14:
    one region ahead, the default, 15: 64 x 8 tiles, 16: 128 x 16 tiles of
    1024 lanes, 17: 128 x 8 tiles, each one region ahead),
    whose loads go to addresses clamped into each region, on whole and
    partial tiles and one-point domains, agree bit for bit with the oracle."""
    knobs.set("fvtp2d", "form", 0)
    knobs.set("fvtp2d", "tile", tile)
    L, host = inputs("fvtp2d", *shape, family, seed=19)
    print(f"[gpu] fvtp2d tile={tile} {family} {shape}", flush=True)
    _check("fvtp2d", L, host, run_gpu("fvtp2d", copy(host), L), require_exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("kchunk", [1, 3])
@pytest.mark.parametrize("tile", [0, 1, 2])
@pytest.mark.parametrize("shape", [(64, 36, 3), (67, 9, 2), (130, 17, 4), (1, 1, 1)],
                         ids=lambda s: "x".join(map(str, s)))
def test_fvtp2d_lds_tiles(shape, tile, kchunk, knobs):
    """fvtp2d's LDS-staged tiles (form knob 2: every input region of a level
    loaded straight into LDS) in each tile shape, on whole and partial tiles
    and one-point domains, with level chunks that divide nz and that do not,
    agree bit for bit with the oracle."""
    knobs.set("fvtp2d", "form", 2)
    knobs.set("fvtp2d", "tile", tile)
    knobs.set("fvtp2d", "kchunk", kchunk)
    L, host = inputs("fvtp2d", *shape, "baseline", seed=17)
    print(f"[gpu] fvtp2d lds tile={tile} kchunk={kchunk} {shape}", flush=True)
    _check("fvtp2d", L, host, run_gpu("fvtp2d", copy(host), L), require_exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("xcd", [0, 1])
@pytest.mark.parametrize("shape", [(64, 36, 8), (200, 9, 4), (256, 20, 6), (66, 17, 2)],
                         ids=lambda s: "x".join(map(str, s)))
def test_laplace_xcd_interleave(shape, xcd, knobs):
    """laplace with level chunks interleaved between tile rows (the order knob)
    and without, on shapes where the interleave applies (8 or 4 chunks per
    group) and where it does not (chunks not a multiple of the group, or the
    scalar path at nx = 66), agrees bit for bit with the oracle."""
    knobs.set("laplace", "order", xcd)
    L, host = inputs("laplace", *shape, "baseline", seed=22)
    print(f"[gpu] laplace xcd={xcd} {shape}", flush=True)
    _check("laplace", L, host, run_gpu("laplace", copy(host), L), require_exact=True)

@pytest.mark.gpu
@pytest.mark.parametrize("kchunk", [1, 3, 8, 80])
@pytest.mark.parametrize("name", ["hadvuv", "hadvuv5th"])
def test_hadv_level_chunks(name, kchunk, knobs):
    """hadvuv and hadvuv5th at level chunks that divide nz, do not divide
    it, and exceed it (the next level's rows are loaded during the current
    one) agree bit for bit with the oracle."""
    knobs.set(name, "kchunk", kchunk)
    L, host = inputs(name, 130, 36, 11, "baseline", seed=25)
    print(f"[gpu] {name} kchunk={kchunk}", flush=True)
    _check(name, L, host, run_gpu(name, copy(host), L), require_exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("order", [0, 1, 2, 3])
@pytest.mark.parametrize("name", ["fvtp2d", "hadvuv", "hadvuv5th", "fastwavesuv", "p_grad_c", "nh_p_grad", "uvbke",
                                  "laplace"])
@pytest.mark.parametrize("shape", [(200, 64, 8), (130, 32, 16), (67, 9, 3), (64, 40, 64)],
                         ids=lambda s: "x".join(map(str, s)))
def test_tile_row_order(shape, name, order, knobs):
    """The tiled stencils in every tile order of the order knob: rows in order
    (0), level chunks interleaved between tile rows (1, Tiles::xcd), the tiles
    of a row on one XCD (2, Tiles::irow) and whole level chunks per XCD (3,
    Tiles::klev), on shapes where each order applies (8, 16 and 64 level
    chunks; 1 to 4 tiles per row) and where it does not, agree bit for bit
    with the oracle."""
    knobs.set(name, "order", order)
    knobs.set("fvtp2d", "form", 0)
    L, host = inputs(name, *shape, "baseline", seed=24)
    print(f"[gpu] {name} tile order {order} {shape}", flush=True)
    _check(name, L, host, run_gpu(name, copy(host), L), require_exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("lanes", [64, 128, 512, 1024])
@pytest.mark.parametrize("family", ["baseline", "stress"])
def test_fvtp2d_strip_shapes(family, lanes, depth, knobs):
    """fvtp2d's strips of at most 64 .. 1024 lanes (one to five strips across
    nx = 300), with one and two rows loaded ahead, agree bit for bit with the
    oracle."""
    knobs.set("fvtp2d", "form", 1)
    knobs.set("fvtp2d", "lanes", lanes)
    knobs.set("fvtp2d", "depth", depth)
    L, host = inputs("fvtp2d", 300, 23, 2, family, seed=15)
    print(f"[gpu] fvtp2d strip {family} lanes={lanes} depth={depth}", flush=True)
    _check("fvtp2d", L, host, run_gpu("fvtp2d", copy(host), L), require_exact=True)
