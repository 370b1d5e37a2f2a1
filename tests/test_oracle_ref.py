"""The oracle against the reference's own kernels, compiled in place from the
reference sources (oracle/_ref/libref.so, built by oracle/Makefile).

The oracle must reproduce the reference's oracle variant (OpenMP; DEFAULT
for fvtp2d) bit for bit on every output region, for several shapes, both
layouts and all input families.  The other reference variants are checked on
the regions where they are well defined.  The ring where hdiffsmag DEFAULT
and p_grad_c DEFAULT read unwritten scratch is excluded.
"""
import numpy as np
import pytest

from common import bit_diffs, copy, inputs, out_region, rel_maxnorm
from oeb.spec import KERNELS, ORDER
from oracle import binding as B

pytestmark = pytest.mark.skipif(not B.ref_available(), reason="oracle/_ref/libref.so not built (no reference tree)")

SHAPES = [(12, 12, 5, 16), (20, 20, 3, 1), (7, 7, 2, 16), (33, 33, 4, 1)]


@pytest.mark.parametrize("family", ["baseline", "refdriver", "stress"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}x{s[2]}a{s[3]}")
@pytest.mark.parametrize("name", ORDER)
def test_oracle_equals_reference(name, shape, family):
    nx, ny, nz, align = shape
    L, host = inputs(name, nx, ny, nz, family, seed=3, align=align)
    a, b = copy(host), copy(host)
    B.run_oracle(name, a, L)
    B.run_ref(name, b, L, threads=2)
    for o in KERNELS[name].outputs:
        x, y = out_region(name, o, L, a[o]), out_region(name, o, L, b[o])
        assert np.isfinite(y).all()
        assert bit_diffs(x, y) == 0, f"{name}.{o}: oracle differs from the reference (rel {rel_maxnorm(x, y):.2e})"


def test_reference_constants_match_registry():
    """The driver constants the reference build uses equal the registry's
    defaults (src/uvbke.cpp:11-12, src/fvtp2d.cpp:11-12)."""
    lib = B.ref_lib()
    assert lib.ref_uvbke_dt5() == dict(KERNELS["uvbke"].scalars)["dt5"]
    assert lib.ref_fvtp2d_p1() == dict(KERNELS["fvtp2d"].scalars)["p1"]
    assert lib.ref_fvtp2d_p2() == dict(KERNELS["fvtp2d"].scalars)["p2"]


# Variants that reproduce the oracle bit for bit on the full output region.
EXACT_VARIANTS = {
    "laplace": [B.DEFAULT, B.FULL_FUSION, B.PARTIAL_FUSION],
    "hdiffsa": [B.DEFAULT, B.FULL_FUSION, B.PARTIAL_FUSION],
    "hadvuv": [B.DEFAULT, B.FULL_FUSION, B.PARTIAL_FUSION],
    "hadvuv5th": [B.DEFAULT, B.FULL_FUSION, B.PARTIAL_FUSION],
    "fastwavesuv": [B.DEFAULT, B.FULL_FUSION, B.PARTIAL_FUSION],
    "uvbke": [B.DEFAULT, B.FULL_FUSION, B.PARTIAL_FUSION],
    "nh_p_grad": [B.DEFAULT, B.FULL_FUSION, B.PARTIAL_FUSION],
    "p_grad_c": [B.FULL_FUSION, B.PARTIAL_FUSION],
}


@pytest.mark.parametrize("name", sorted(EXACT_VARIANTS))
def test_other_variants_exact(name):
    L, host = inputs(name, 14, 14, 4, "baseline", seed=8)
    a = copy(host)
    B.run_oracle(name, a, L)
    for v in EXACT_VARIANTS[name]:
        b = copy(host)
        B.run_ref(name, b, L, variant=v, threads=1)
        for o in KERNELS[name].outputs:
            assert bit_diffs(out_region(name, o, L, a[o]), out_region(name, o, L, b[o])) == 0, (name, v, o)


def test_p_grad_c_default_differs_only_on_faces():
    """p_grad_c DEFAULT copies delpc to scratch on [0,N)^2 only and then reads
    wk(i-1) / wk(j-1) (kernels/p_grad_c.h:8-12, 17, 21).  It agrees with the
    oracle on i >= 1 (uout) and j >= 1 (vout), and not on the faces."""
    L, host = inputs("p_grad_c", 12, 12, 3, "baseline", seed=4)
    a, b = copy(host), copy(host)
    B.run_oracle("p_grad_c", a, L)
    B.run_ref("p_grad_c", b, L, variant=B.DEFAULT, threads=1)
    assert bit_diffs(L.region(a["uout"], 1, 12, 0, 12), L.region(b["uout"], 1, 12, 0, 12)) == 0
    assert bit_diffs(L.region(a["vout"], 0, 12, 1, 12), L.region(b["vout"], 0, 12, 1, 12)) == 0
    assert bit_diffs(L.region(a["uout"], 0, 1, 0, 12), L.region(b["uout"], 0, 1, 0, 12)) > 0


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_hdiffsmag_variants_inner_region(variant):
    """hdiffsmag DEFAULT reads T_sqr_s / S_sqr_uv outside the region it wrote,
    so only [1,N-1)^2 is defined (kernels/hdiffsmag.h:41).  The fused
    variants associate (du * acrlat0) * eddlon instead of du * (acrlat0 *
    eddlon) (:11 vs :213).  All agree with the oracle to rounding."""
    L, host = inputs("hdiffsmag", 16, 16, 3, "stress", seed=6)
    a, b = copy(host), copy(host)
    B.run_oracle("hdiffsmag", a, L)
    B.run_ref("hdiffsmag", b, L, variant=variant, threads=1)
    for o in ("uout", "vout"):
        x, y = L.region(a[o], 1, 15, 1, 15), L.region(b[o], 1, 15, 1, 15)
        assert rel_maxnorm(y, x) <= 1e-13, (variant, o)
