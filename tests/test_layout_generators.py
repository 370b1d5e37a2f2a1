"""Layout arithmetic and the seeded input generators (CPU)."""
import numpy as np
import pytest

from oeb.generators import FAMILIES, generate
from oeb.layout import Layout, Layout1D
from oeb.spec import KERNELS, ORDER


def test_reference_layout_matches_util_h():
    """align = 1 reproduces allocateStorage (include/util.h:121-141) with the k
    halo: sx = nx + 2h, sk = sx * sy, offset = h + h sx + hk sk."""
    L = Layout(64, 64, 60, h=4, hk=4, align=1)
    assert L.sj == 72 and L.sk == 72 * 72
    assert L.offset == 4 * (1 + 72 + 72 * 72)
    assert L.sizes() == (68, 72, 72) and L.strides() == (72 * 72, 72, 1)


@pytest.mark.parametrize("nx", [1, 7, 64, 256, 500])
def test_aligned_rows(nx):
    L = Layout(nx, 5, 3, h=3, hk=1, align=16)
    assert L.sj % 16 == 0 and L.lead % 16 == 0
    assert L.lead >= L.h and L.sj - L.lead - nx >= L.h
    for j in range(-3, 8):
        assert (L.offset + j * L.sj + 0 * L.sk) % 16 == 0


def test_region_and_index_agree():
    L = Layout(6, 5, 2, h=3, hk=1, align=16)
    a = np.arange(L.size, dtype=np.float64)
    r = L.region(a, -3, 8, -2, 6, -1, 3)
    assert r[0, 0, 0] == L.index(-3, -2, -1)
    assert r[-1, -1, -1] == L.index(7, 5, 2)


def test_one_d_layout():
    L1 = Layout1D(10, 3)
    assert L1.size == 16 and L1.offset == 3 and L1.coords()[0] == -3


@pytest.mark.parametrize("family", sorted(FAMILIES))
@pytest.mark.parametrize("name", ORDER)
def test_generators_fill_everything_deterministically(name, family):
    L = Layout(9, 7, 3, h=3, hk=1, align=16)
    a = generate(name, L, family, seed=5)
    b = generate(name, L, family, seed=5)
    c = generate(name, L, family, seed=6)
    spec = KERNELS[name]
    for arg in spec.args:
        if arg.kind == "scratch":
            assert arg.name not in a
        elif arg.kind == "out":
            assert np.isnan(a[arg.name]).all()
        else:
            assert np.isfinite(a[arg.name]).all(), arg.name  # halos, padding and k-halo included
            assert np.array_equal(a[arg.name], b[arg.name])
    if family != "refdriver":  # refdriver inputs are the reference's deterministic formulas
        assert any(not np.array_equal(a[x.name], c[x.name]) for x in spec.args if x.kind == "in")


def test_baseline_distributions():
    L = Layout(64, 64, 20, h=3, hk=1, align=16)
    fv = generate("fvtp2d", L, "baseline", 0)
    assert -0.4 <= fv["crx"].min() and fv["crx"].max() <= 0.4
    assert 0.01 <= fv["area"].min() and fv["area"].max() <= 0.1
    assert np.abs(fv["q"]).max() <= 3.0 + 1e-3
    nh = generate("nh_p_grad", L, "baseline", 0)
    pk = L.view(nh["pk3"])
    assert (np.diff(pk, axis=0) < 0).all(), "pressure must decrease monotonically with k (no division poles)"
    assert (nh["wk1"] > 0).all()
    fw = generate("fastwavesuv", L, "baseline", 0)
    assert (np.diff(L.view(fw["hhl"]), axis=0) > 0).all()
    hd = generate("hdiffsa", L, "baseline", 0)
    L1 = Layout1D(64, 3)
    lat = np.rad2deg(np.arccos(hd["crlato"][L1.offset:L1.offset + 64]))
    assert abs(lat[0] - 80) < 1e-9 and abs(lat[-1] - 80) < 1e-9  # cos is even: both ends at |80| degrees


def test_refdriver_matches_fillmath_formula():
    """fillMath (include/util.h:150-164): k * 10e-3 + a (b + cos(pi (x + c y)) + sin(d pi (x + e y))) / f
    with x = i / (N + 2h), y = j / (N + 2h)."""
    L = Layout(12, 12, 3, h=3, hk=1, align=1)
    g = generate("laplace", L, "refdriver", 0)["in"]
    a, b, c, d, e, f = 1.1, 2.0, 1.5, 2.8, 2.0, 4.1
    i, j, k = 5, -2, 2
    x, y = i / 18.0, j / 18.0
    want = k * 10e-3 + a * (b + np.cos(np.pi * (x + c * y)) + np.sin(d * np.pi * (x + e * y))) / f
    assert abs(g[L.index(i, j, k)] - want) < 1e-15


@pytest.mark.parametrize("name", ORDER)
def test_device_generator_matches_host_distributions(name):
    """generate_device (torch, here on the CPU) follows the host generator's
    recipes: same fields and sizes, NaN outputs, finite inputs everywhere,
    the same distributions (3-D fields: 1st, 50th and 99th percentiles and
    the standard deviation within 10% of the host field's 1-99 percentile
    spread; 1-D fields: the host field's range), and the
    structural properties the stencils rely on."""
    import torch

    from oeb.generators import generate_device

    L = Layout(64, 48, 12, h=3, hk=1, align=16)
    host = generate(name, L, "baseline", seed=2)
    dev = generate_device(name, L, seed=2, device="cpu")
    assert set(dev) == set(host)
    for arg in KERNELS[name].args:
        if arg.kind == "scratch":
            continue
        a = dev[arg.name].numpy()
        assert a.dtype == np.float64 and a.shape == host[arg.name].shape, arg.name
        if arg.kind == "out":
            assert np.isnan(a).all()
            continue
        h = host[arg.name]
        assert np.isfinite(a).all(), arg.name
        if arg.kind == "in1":  # a few dozen samples: the range only
            span = h.max() - h.min()
            assert h.min() - 0.1 * span <= a.min() and a.max() <= h.max() + 0.1 * span, arg.name
            continue
        qa, qh = np.percentile(a, [1, 50, 99]), np.percentile(h, [1, 50, 99])
        spread = max(qh[2] - qh[0], 1e-12)
        assert np.abs(qa - qh).max() <= 0.1 * spread, (arg.name, qa, qh)
        assert abs(a.std() - h.std()) <= 0.1 * spread, (arg.name, a.std(), h.std())
    if name == "nh_p_grad":
        assert (np.diff(L.view(dev["pk3"].numpy()), axis=0) < 0).all()
        assert (dev["wk1"] > 0).all()
    if name == "fastwavesuv":
        assert (np.diff(L.view(dev["hhl"].numpy()), axis=0) > 0).all()
    again = generate_device(name, L, seed=2, device="cpu")
    assert all(torch.equal(dev[k], again[k]) or dev[k].isnan().all() for k in dev)
