"""Pins the oracle's FP32 form (oracle_<name>_f32).

- Stencils without FP64 islands, and fvtp2d's fx2 / fy2: bit for bit equal to
  the reference built with ElementType = float (oracle/_ref/libref_f32.so,
  where it is built) and to the committed FP32 fixtures made from it.
- fastwavesuv, p_grad_c, nh_p_grad (FP64 islands): against the FP64 oracle,
  which the FP64 tests pin to the reference.  The bar, 5e-7 relative
  max-norm, is about 3x the measured error.  A dropped term or a wrong sign
  misses it by orders of magnitude.  On the physical inputs the reference's
  float build, which has no islands, is at least 5x further from FP64.
- fvtp2d's compensated divergences q_i and q_j: within 1 ulp of the float
  nearest the exact value (q*area + xa*ya - xb*yb) / ra, formed in double
  from the same float inputs.  Plain float rounding is off by up to
  thousands of ulps where the terms cancel.  fx1 / fy1 then stay within
  5e-7 of FP64.
- Every stencil: within 1e-5 relative max-norm of the FP64 oracle on the
  same (unrounded) inputs.  That is the FP32 target error.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from common import bit_diffs, copy, inputs, out_region, rel_maxnorm, rel_rms
from golden.make_golden import F32_CASES, F32_EXACT, fixture_name_f32, sha
from oeb.generators import cast, generate
from oeb.layout import Layout
from oeb.spec import ISLANDS, KERNELS, ORDER
from oracle import binding as B

GOLDEN = Path(__file__).resolve().parent / "golden"
SHAPES = [(12, 12, 5, 16), (20, 20, 3, 1), (33, 33, 4, 16)]
FAMILIES = ["baseline", "refdriver", "stress"]
ISLAND_TOL = 5e-7
F32_TARGET = 1e-5


def _f32(name, host, L):
    h = cast(name, host, "f32")
    B.run_oracle(name, h, L, precision="f32")
    return h


def _f64(name, host, L):
    h = copy(host)
    B.run_oracle(name, h, L)
    return h


@pytest.mark.parametrize("case", [(k,) + c for k in F32_EXACT for c in F32_CASES], ids=lambda c: c[0])
def test_f32_oracle_matches_float_golden(case):
    kernel, family, nx, ny, nz, align, seed = case
    f = GOLDEN / fixture_name_f32(kernel, family, nx, ny, nz)
    assert f.exists(), f"missing fixture {f.name}: run tests/golden/make_golden.py"
    data = np.load(f, allow_pickle=False)
    meta = json.loads(str(data["meta"]))
    assert meta["element_type"] == "float"
    L = Layout(nx, ny, nz, h=meta["h"], hk=meta["hk"], align=meta["align"])
    host = cast(kernel, generate(kernel, L, family, seed), "f32", islands=False)
    for name, digest in meta["input_sha256"].items():
        assert sha(host[name]) == digest, f"generator output for {kernel}.{name} changed; regenerate the fixtures"
    a = copy(host)
    B.run_oracle(kernel, a, L, precision="f32")
    for o in F32_EXACT[kernel]:
        x = out_region(kernel, o, L, a[o])
        assert x.dtype == np.float32
        assert bit_diffs(x, data[o]) == 0, f"{kernel}.{o}: FP32 oracle differs from the float reference"


@pytest.mark.skipif(not B.ref_available("f32"), reason="oracle/_ref/libref_f32.so not built (no reference tree)")
@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s[:3])))
@pytest.mark.parametrize("kernel", list(F32_EXACT))
def test_f32_oracle_matches_float_reference(kernel, shape, family):
    nx, ny, nz, align = shape
    L, host = inputs(kernel, nx, ny, nz, family, seed=3, align=align)
    ours = _f32(kernel, host, L)
    ref = cast(kernel, host, "f32", islands=False)
    B.run_ref(kernel, ref, L, threads=1, precision="f32")
    for o in F32_EXACT[kernel]:
        x, y = out_region(kernel, o, L, ours[o]), out_region(kernel, o, L, ref[o])
        assert bit_diffs(x, y) == 0, f"{kernel}.{o}: rel {rel_maxnorm(x, y):.2e}"


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("kernel", sorted(ISLANDS))
def test_f32_islands_track_f64(kernel, family):
    L, host = inputs(kernel, 40, 40, 6, family, seed=5)
    ours, exact = _f32(kernel, host, L), _f64(kernel, host, L)
    ref = None
    # The reference drivers' fillMath inputs are O(1): there, nothing cancels
    # and the islands change little.  The physical (baseline) pressures do.
    if family != "refdriver" and B.ref_available("f32"):
        ref = cast(kernel, host, "f32", islands=False)
        B.run_ref(kernel, ref, L, threads=1, precision="f32")
    for o in KERNELS[kernel].outputs:
        x, y = out_region(kernel, o, L, ours[o]), out_region(kernel, o, L, exact[o])
        err = rel_maxnorm(x, y)
        assert err <= ISLAND_TOL, f"{kernel}.{o}: FP32-with-islands rel max-norm {err:.2e} > {ISLAND_TOL:.0e}"
        if ref is not None:
            plain = rel_maxnorm(out_region(kernel, o, L, ref[o]), y)
            assert plain >= 5 * err, f"{kernel}.{o}: islands {err:.2e} vs plain float {plain:.2e}"


def _shift(L, a, di, dj):
    """a(i+di, j+dj) as a flat array over the allocation (wraps at the edges,
    which the region extraction never reaches)."""
    v = a.reshape(L.levels, L.rows, L.sj)
    return np.roll(np.roll(v, -dj, axis=1), -di, axis=2).ravel()


@pytest.mark.parametrize("family", FAMILIES)
def test_fvtp2d_f32_divergence_within_one_ulp(family):
    L, host = inputs("fvtp2d", 40, 36, 5, family, seed=4)
    h = _f32("fvtp2d", host, L)
    for out, (xa, ya, ra, di, dj) in {"q_i": ("yfx", "fy2", "ra_y", 0, 1), "q_j": ("xfx", "fx2", "ra_x", 1, 0)}.items():
        g = lambda name, i=0, j=0: out_region("fvtp2d", out, L, _shift(L, h[name], i, j)).astype(np.float64)  # noqa
        num = g("q") * g("area") + g(xa) * g(ya) - g(xa, di, dj) * g(ya, di, dj)
        want = (num / g(ra)).astype(np.float32)
        got = out_region("fvtp2d", out, L, h[out])
        ulps = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
        assert ulps.max() <= 1, f"{out}: {int(ulps.max())} ulps from the correctly rounded divergence"


@pytest.mark.parametrize("family", FAMILIES)
def test_fvtp2d_f32_fluxes_track_f64(family):
    L, host = inputs("fvtp2d", 40, 40, 5, family, seed=6)
    ours, exact = _f32("fvtp2d", host, L), _f64("fvtp2d", host, L)
    for o in ("fx1", "fy1"):
        x, y = out_region("fvtp2d", o, L, ours[o]), out_region("fvtp2d", o, L, exact[o])
        assert rel_maxnorm(x, y) <= ISLAND_TOL, f"fvtp2d.{o}: {rel_maxnorm(x, y):.2e}"


@pytest.mark.parametrize("kernel", ORDER)
def test_f32_error_budget(kernel):
    """FP32 form against FP64 on the same FP64 inputs: <= 1e-5 relative
    max-norm on every output (input rounding included)."""
    L, host = inputs(kernel, 64, 64, 8, "baseline", seed=8)
    ours, exact = _f32(kernel, host, L), _f64(kernel, host, L)
    for o in KERNELS[kernel].outputs:
        x, y = out_region(kernel, o, L, ours[o]), out_region(kernel, o, L, exact[o])
        err, rms = rel_maxnorm(x, y), rel_rms(x, y)
        assert err <= F32_TARGET, f"{kernel}.{o}: rel max {err:.2e} (rms {rms:.2e})"
