"""Deviation islands on the CPU (oeb.islands): the split into an FP64
per-level base and FP32 deviations, its rebuilt values, and what it buys in
the FP32 forms of p_grad_c and nh_p_grad, measured with the oracle."""
import numpy as np
import pytest

from common import copy, inputs, out_region, rel_maxnorm
from oeb.generators import cast
from oeb.islands import DEVIATION_STENCILS, reconstruct, split_host
from oeb.layout import Layout
from oeb.spec import KERNELS
from oracle import binding as B


def test_split_rebuilds_within_the_deviation_rounding():
    """base(k) is the level's mean; rebuilt values differ from the field by
    the float32 rounding of the deviation (<= 2^-24 |dev|) plus the FP64
    rounding of the sum (<= 2^-53 |value|)."""
    L = Layout(20, 14, 5, h=3, hk=1, align=16)
    rng = np.random.default_rng(3)
    x = (1e5 * np.exp(-np.arange(L.levels) / 10.0))[:, None, None] * (1 + 0.01 * rng.standard_normal(
        (L.levels, L.rows, L.sj)))
    x = x.reshape(-1)
    dev, base = split_host(L, x)
    assert dev.dtype == np.float32 and base.dtype == np.float64 and base.shape == (L.levels,)
    assert np.allclose(base, L.view(x).mean(axis=(1, 2)), rtol=0, atol=0)
    back = reconstruct(L, dev, base)
    d = L.view(x) - base[:, None, None]
    bound = 2.0 ** -24 * np.abs(d) + 2.0 ** -52 * np.abs(L.view(x))
    assert (np.abs(L.view(back) - L.view(x)) <= bound).all()
    # the rebuilt field is ~1e-2 x 6e-8 off, where FP32 storage of the value is 6e-8 off
    assert rel_maxnorm(back, x) < 2e-9 < rel_maxnorm(x.astype(np.float32), x)


def test_level_constant_field_splits_exactly():
    L = Layout(9, 7, 4, h=3, hk=1, align=1)
    prof = 9.81 * 250.0 * np.arange(-1, L.nz + 1)
    x = np.broadcast_to(prof[:, None, None], (L.levels, L.rows, L.sj)).reshape(-1).copy()
    dev, base = split_host(L, x)
    assert not dev.any() and np.array_equal(base, prof)
    assert np.array_equal(reconstruct(L, dev, base), x)


def _f32_error(name, L, host, islands):
    """Relative max-norm error of the FP32 oracle against the FP64 oracle, with
    the islands given as FP64 arrays ("f64"), rebuilt from deviations about the
    level mean ("mean"), or stored in FP32 (deviations about a zero base)."""
    h32 = cast(name, host, "f32")
    for a in DEVIATION_STENCILS[name]:
        if islands == "mean":
            h32[a] = reconstruct(L, *split_host(L, host[a]))
        elif islands == "zero":
            h32[a] = host[a].astype(np.float32).astype(np.float64)
    exact = copy(host)
    B.run_oracle(name, exact, L)
    got = copy(h32)
    B.run_oracle(name, got, L, precision="f32")
    return max(rel_maxnorm(out_region(name, o, L, got[o]), out_region(name, o, L, exact[o]))
               for o in KERNELS[name].outputs)


@pytest.mark.parametrize("name", sorted(DEVIATION_STENCILS))
def test_deviation_islands_keep_the_fp64_island_accuracy(name):
    """With deviations about the level mean the FP32 form stays within 1e-6
    of the FP64 oracle, as with FP64 islands (48x40x12, measured 1.2e-7 to
    1.5e-7 for both), while FP32 storage of the island values themselves (a
    zero base) is off by 1.8e-6 to 2.6e-6: 13-20x more, past the 1e-6 bound,
    from the cancellation in the level and neighbour differences."""
    L, host = inputs(name, 48, 40, 12, "baseline", seed=4)
    e64 = _f32_error(name, L, host, "f64")
    emean = _f32_error(name, L, host, "mean")
    ezero = _f32_error(name, L, host, "zero")
    print(f"{name}: FP64 islands {e64:.2e}, deviations {emean:.2e}, FP32 values {ezero:.2e}")
    assert e64 <= 1e-6 and emean <= 1e-6
    assert emean <= 1.2 * e64
    assert ezero > 1e-6 and ezero > 8 * emean
