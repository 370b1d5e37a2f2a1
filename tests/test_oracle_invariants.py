"""The oracle against closed forms and invariants the mathematics fixes.

These pin the oracle independently of the reference build.  Each case is
chosen so that a plausible transcription mistake fails it: a dropped term,
a wrong sign or index, swapped operands (crlato/crlatu, crx/cry, xfx/yfx)
or a wrong output region.  Values are exact binary fractions wherever the
answer is asserted exactly.
"""
import numpy as np
import pytest

from common import copy, inputs, out_region, rel_maxnorm
from oeb.generators import generate
from oeb.layout import Layout, Layout1D
from oeb.spec import KERNELS
from oracle import binding as B


def _grid(L: Layout):
    """Domain-relative (k, j, i) index arrays over the full allocation."""
    I, J = L.ij_grid()
    K = L.k_levels()[:, None, None]
    return np.broadcast_to(K, (L.levels, L.rows, L.sj)), np.broadcast_to(J, (L.levels, L.rows, L.sj)), \
        np.broadcast_to(I, (L.levels, L.rows, L.sj))


def _flat(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


# ------------------------------------------------------------------ laplace
def test_laplace_quadratic_is_exact():
    """in = i^2 + 3 j^2 + 5 k  ->  out = 2 + 6 = 8 everywhere (exact in FP)."""
    L = Layout(10, 9, 3)
    K, J, I = _grid(L)
    h = {"in": _flat(I * I + 3 * J * J + 5 * K), "out": np.full(L.size, np.nan)}
    B.run_oracle("laplace", h, L)
    assert (L.interior(h["out"]) == 8.0).all()


def test_laplace_transpose_symmetry():
    """laplace commutes with the i<->j transpose bit for bit (the sums are
    pairwise commutative)."""
    L = Layout(11, 11, 2, align=1)
    _, h = inputs("laplace", 11, 11, 2, seed=4, align=1)
    t = copy(h)
    t["in"] = _flat(L.view(h["in"]).transpose(0, 2, 1))
    B.run_oracle("laplace", h, L)
    B.run_oracle("laplace", t, L)
    assert np.array_equal(L.interior(h["out"]), L.interior(t["out"]).transpose(0, 2, 1))


# ------------------------------------------------------------------ hdiffsa
def _hdiffsa_case(nx, ny, nz, in_fn, crlato, crlatu, mask):
    L = Layout(nx, ny, nz)
    K, J, I = _grid(L)
    L1 = Layout1D(ny, 3)
    h = {"out": np.full(L.size, np.nan), "in": _flat(in_fn(I, J, K)), "mask": np.full(L.size, mask),
         "crlato": np.full(L1.size, crlato), "crlatu": np.full(L1.size, crlatu)}
    B.run_oracle("hdiffsa", h, L)
    return L, h, I, J


def test_hdiffsa_alternating_in_i():
    """in = (-1)^i: lap = -4(-1)^i, flx = 8(-1)^i (anti-gradient, not limited),
    fly = 0, out = (-1)^i (1 - 16 mask).  mask = 1/32 -> out = (-1)^i / 2."""
    L, h, I, J = _hdiffsa_case(8, 6, 2, lambda i, j, k: np.where(i % 2 == 0, 1.0, -1.0), 1.0, 1.0, 1.0 / 32)
    want = np.where(L.interior(h["in"]) > 0, 0.5, -0.5)
    assert np.array_equal(L.interior(h["out"]), want)


def test_hdiffsa_alternating_in_j_uses_crlato_for_fly():
    """in = (-1)^j, crlato = 2, crlatu = 1: lap = -6(-1)^j, fly = crlato * 12
    (-1)^j = 24 (-1)^j, out = (-1)^j (1 - 48 mask).  mask = 1/64 -> out =
    (-1)^j / 4.  Swapping crlato and crlatu anywhere changes the answer."""
    L, h, I, J = _hdiffsa_case(6, 8, 2, lambda i, j, k: np.where(j % 2 == 0, 1.0, -1.0), 2.0, 1.0, 1.0 / 64)
    want = np.where(L.interior(h["in"]) > 0, 0.25, -0.25)
    assert np.array_equal(L.interior(h["out"]), want)


def test_hdiffsa_limiter_and_constant_fields():
    """A constant field or mask = 0 leaves out == in exactly; a cubic in i
    makes every flux down the gradient, so the limiter zeroes all of them."""
    for fn, mask in [(lambda i, j, k: 0 * i + 3.25, 0.05), (lambda i, j, k: np.sin(i + 2.0 * j), 0.0),
                     (lambda i, j, k: (i + 0.0) ** 3, 0.05)]:
        L, h, I, J = _hdiffsa_case(7, 7, 2, fn, 1.0, 1.0, mask)
        assert np.array_equal(L.interior(h["out"]), L.interior(h["in"]))


# ---------------------------------------------------------------- hdiffsmag
def test_hdiffsmag_uniform_wind_unchanged():
    L, h = inputs("hdiffsmag", 9, 9, 2, "stress", seed=1)
    h["uin"][:] = 3.5
    h["vin"][:] = -1.25
    B.run_oracle("hdiffsmag", h, L)
    assert (L.interior(h["uout"]) == 3.5).all() and (L.interior(h["vout"]) == -1.25).all()


def test_hdiffsmag_quadratic_u_known_answer():
    """uin = a i^2, vin = 0, acrlat0 = c: T(p) = a (2p - 1) c eddlon, S = 0,
    lapu = 2a, lapv = 0.  uout = uin + clamp(tau sqrt((T(i+1)^2 + T(i)^2)/2)
    - w mask) * 2a, and vout = vin.  a, c are chosen so that the coefficient
    crosses both clamps across i."""
    nx, ny, nz = 40, 4, 1
    L = Layout(nx, ny, nz)
    K, J, I = _grid(L)
    L1 = Layout1D(ny, 3)
    a, c, mask = 0.5, 2730.0, 5.0  # |T| runs from ~1 to ~80; w * mask = 0.05
    sc = dict(KERNELS["hdiffsmag"].scalars)
    h = {"uout": np.full(L.size, np.nan), "vout": np.full(L.size, np.nan), "uin": _flat(a * I * I),
         "vin": np.zeros(L.size), "mask": np.full(L.size, mask)}
    for n in ("crlavo", "crlavu", "crlato", "crlatu"):
        h[n] = np.ones(L1.size)
    h["acrlat0"] = np.full(L1.size, c)
    B.run_oracle("hdiffsmag", h, L)
    i = np.arange(nx, dtype=np.float64)
    t0 = a * (2 * i - 1) * c * sc["eddlon"]
    t1 = a * (2 * (i + 1) - 1) * c * sc["eddlon"]
    smag = sc["tau_smag"] * np.sqrt(0.5 * (t1 * t1 + t0 * t0)) - sc["weight_smag"] * mask
    smag = np.minimum(0.5, np.maximum(0.0, smag))
    assert (smag == 0).any() and (smag == 0.5).any() and ((smag > 0) & (smag < 0.5)).any()
    want = a * i * i + smag * 2 * a
    got = L.interior(h["uout"])[0]
    assert rel_maxnorm(got, np.broadcast_to(want, got.shape)) < 1e-14
    assert (L.interior(h["vout"]) == 0).all()


# ------------------------------------------------------------ hadvuv(5th)
@pytest.mark.parametrize("name", ["hadvuv", "hadvuv5th"])
def test_hadv_constant_field_metric_term_only(name):
    """For constant u = v = c the upwind stencils' coefficients sum to zero
    (3rd order: -1/6 + 1 - 1/2 - 1/3; 5th: 1/30 - 1/4 + 1 - 1/3 - 1/2 + 1/20),
    so only the metric terms remain: uout = tgrlatda0 c^2, vout = -tgrlatda1 c^2
    (to rounding of the folded coefficients)."""
    L, h = inputs(name, 10, 8, 2, seed=2)
    c = 1.5
    h["uin"][:] = c
    h["vin"][:] = c
    B.run_oracle(name, h, L)
    L1 = Layout1D(8, 3)
    tg0 = h["tgrlatda0"][L1.offset:L1.offset + 8][None, :, None]
    tg1 = h["tgrlatda1"][L1.offset:L1.offset + 8][None, :, None]
    assert rel_maxnorm(L.interior(h["uout"]), np.broadcast_to(tg0 * c * c, (2, 8, 10))) < 1e-12
    assert rel_maxnorm(L.interior(h["vout"]), np.broadcast_to(-tg1 * c * c, (2, 8, 10))) < 1e-12


@pytest.mark.parametrize("name,order", [("hadvuv", 3), ("hadvuv5th", 5)])
def test_hadv_upwind_on_linear_field(name, order):
    """On u = x (linear in i), v = 0, with tgrlatda = 0, the x-stencil is a
    consistent derivative: both upwind forms give ua * (-1) * (du/di), i.e.
    uout = eddlat * ua * (-1) with ua = acrlat0 * u (u > 0: first branch,
    u < 0: second).  vout = 0 (the v stencil sees v = 0)."""
    nx, ny = 12, 6
    L = Layout(nx, ny, 1)
    K, J, I = _grid(L)
    L1 = Layout1D(ny, 3)
    sc = dict(KERNELS[name].scalars)
    h = {"uout": np.full(L.size, np.nan), "vout": np.full(L.size, np.nan), "uin": _flat(I - 5.5 + 0.0 * J),
         "vin": np.zeros(L.size), "acrlat0": np.full(L1.size, 2.0), "acrlat1": np.full(L1.size, 1.0),
         "tgrlatda0": np.zeros(L1.size), "tgrlatda1": np.zeros(L1.size)}
    B.run_oracle(name, h, L)
    u = np.arange(nx) - 5.5
    ua = 2.0 * u  # acrlat0 * (1/3)(u-1 + u + u+1) = 2 u
    want = sc["eddlat"] * ua * -1.0
    got = L.interior(h["uout"])[0]
    assert rel_maxnorm(got, np.broadcast_to(want, got.shape)) < 1e-13
    assert np.abs(L.interior(h["vout"])).max() == 0.0


# -------------------------------------------------------------- fastwavesuv
def test_fastwavesuv_flat_pressure_gives_tendency_only():
    """ppuv uniform (zero, so the interpolation wgt p + (1 - wgt) p is exact)
    -> every pressure gradient is 0 -> uout = uin + dt utens exactly."""
    L, h = inputs("fastwavesuv", 9, 7, 3, seed=3)
    h["ppuv"][:] = 0.0
    B.run_oracle("fastwavesuv", h, L)
    dt = dict(KERNELS["fastwavesuv"].scalars)["dt"]
    assert np.array_equal(L.interior(h["uout"]), L.interior(h["uin"]) + dt * L.interior(h["utens"]))
    assert np.array_equal(L.interior(h["vout"]), L.interior(h["vin"]) + dt * L.interior(h["vtens"]))


def test_fastwavesuv_horizontal_gradient_only():
    """ppuv = 2 i (k-independent), wgtfac = 1/2, flat hhl per level:
    ppgk differences in k vanish (ppgc = 0), ppgu = 2, ppgv = 0, so
    uout = uin + dt (utens - 2 * 2 fx / (rho(i+1) + rho)), vout = uin + dt vtens."""
    nx, ny, nz = 8, 6, 2
    L = Layout(nx, ny, nz)
    K, J, I = _grid(L)
    L1 = Layout1D(ny, 3)
    rho = _flat(1.0 + 0.25 * ((I + J) % 3))
    h = {"uout": np.full(L.size, np.nan), "vout": np.full(L.size, np.nan), "uin": np.full(L.size, 0.5),
         "vin": np.full(L.size, -0.5), "utens": np.full(L.size, 0.125), "vtens": np.full(L.size, 0.25),
         "wgtfac": np.full(L.size, 0.5), "ppuv": _flat(2.0 * I + 0.0 * K), "hhl": _flat(100.0 * K + 0.0 * I),
         "rho": rho, "fx": np.full(L1.size, 0.75)}
    B.run_oracle("fastwavesuv", h, L)
    dt = dict(KERNELS["fastwavesuv"].scalars)["dt"]
    r = L.view(rho)
    rc = r[L.hk:L.hk + nz, 3:3 + ny, L.lead:L.lead + nx]
    re = r[L.hk:L.hk + nz, 3:3 + ny, L.lead + 1:L.lead + nx + 1]
    want_u = 0.5 + dt * (0.125 - 2.0 * (2.0 * 0.75 / (re + rc)))
    assert rel_maxnorm(L.interior(h["uout"]), want_u) < 1e-15
    assert (L.interior(h["vout"]) == -0.5 + dt * 0.25).all()


# -------------------------------------------------------------------- uvbke
def test_uvbke_linearity_and_known_answer():
    """ub, vb are linear in (uc, vc) for fixed cosa, rsina; and for uc = 1,
    vc = 0: ub = dt5 * 2 * rsina, vb = -dt5 * 2 * cosa * rsina."""
    L, h = inputs("uvbke", 10, 9, 2, seed=4)
    dt5 = dict(KERNELS["uvbke"].scalars)["dt5"]
    a = copy(h)
    a["uc"][:] = 1.0
    a["vc"][:] = 0.0
    B.run_oracle("uvbke", a, L)
    rs, ca = L.interior(h["rsina"]), L.interior(h["cosa"])
    assert np.array_equal(L.interior(a["ub"]), dt5 * 2.0 * rs)
    assert np.array_equal(L.interior(a["vb"]), (dt5 * (0.0 - 2.0 * ca)) * rs)
    s1, s2, s12 = copy(h), copy(h), copy(h)
    s2["uc"], s2["vc"] = h["uc"][::-1].copy(), h["vc"][::-1].copy()
    s12["uc"], s12["vc"] = h["uc"] + s2["uc"], h["vc"] + s2["vc"]
    for s in (s1, s2, s12):
        B.run_oracle("uvbke", s, L)
    for o in ("ub", "vb"):
        assert rel_maxnorm(L.interior(s12[o]), L.interior(s1[o]) + L.interior(s2[o])) < 1e-14


# ----------------------------------------------------------------- p_grad_c
def test_p_grad_c_hydrostatic_column_is_unforced():
    """gz and pkc functions of k only -> both cross products cancel exactly:
    (gz(k+1) - gz(k)) (pkc(k+1) - pkc(k)) + (gz(k) - gz(k+1)) (pkc(k+1) - pkc(k)) = 0."""
    L, h = inputs("p_grad_c", 8, 8, 3, seed=5)
    K, J, I = _grid(L)
    h["gz"] = _flat(9.81 * 250.0 * K + 0.0 * I)
    h["pkc"] = _flat(1e5 * np.exp(-K / 32.0) + 0.0 * I)
    B.run_oracle("p_grad_c", h, L)
    assert np.array_equal(L.interior(h["uout"]), L.interior(h["uin"]))
    assert np.array_equal(L.interior(h["vout"]), L.interior(h["vin"]))


def test_p_grad_c_known_answer():
    """gz = i + 2k, pkc = k (integers): the x bracket is (-1)(1) + (1)(1) ... =
    ((gz(i-1,k+1) - gz)(pkc(k+1) - pkc(i-1)) + (gz(i-1) - gz(k+1))(pkc(i-1,k+1) - pkc))
    = (1)(1) + (-3)(1) = -2; uout = uin + dt2 rdxc / (delpc(i-1) + delpc) * (-2)."""
    L, h = inputs("p_grad_c", 6, 5, 2, seed=6)
    K, J, I = _grid(L)
    h["gz"] = _flat(I + 2.0 * K + 0.0 * J)
    h["pkc"] = _flat(K + 0.0 * I)
    h["delpc"][:] = 0.5
    h["rdxc"][:] = 0.25
    B.run_oracle("p_grad_c", h, L)
    dt2 = dict(KERNELS["p_grad_c"].scalars)["dt2"]
    assert np.array_equal(L.interior(h["uout"]), L.interior(h["uin"]) + ((dt2 * 0.25) / 1.0) * -2.0)
    # along j gz has no j dependence: bracket = (2)(1) + (-2)(1) = 0
    assert np.array_equal(L.interior(h["vout"]), L.interior(h["vin"]))


# ---------------------------------------------------------------- nh_p_grad
def test_nh_p_grad_pp_term_equals_du_when_pp_is_pk3():
    """With pp = pk3 and wk1 = pk3(k+1) - pk3(k), the pp term has exactly the
    form of du (kernels/nh_p_grad.h:18-25): uout = ((uin + du) + du) * rdx.
    With uin = 0, rdx = 1 this gives 2 du; the general case must follow."""
    L, h = inputs("nh_p_grad", 9, 8, 3, seed=7)
    h["pp"] = h["pk3"].copy()
    pk = L.view(h["pk3"])
    wk1 = np.empty_like(pk)
    wk1[:-1] = pk[1:] - pk[:-1]
    wk1[-1] = wk1[-2]
    h["wk1"] = _flat(wk1)
    z = copy(h)
    z["uin"][:] = 0.0
    z["vin"][:] = 0.0
    z["rdx"][:] = 1.0
    z["rdy"][:] = 1.0
    B.run_oracle("nh_p_grad", z, L)
    B.run_oracle("nh_p_grad", h, L)
    du = L.interior(z["uout"]) / 2.0
    dv = L.interior(z["vout"]) / 2.0
    assert np.array_equal(L.interior(h["uout"]), ((L.interior(h["uin"]) + du) + du) * L.interior(h["rdx"]))
    assert np.array_equal(L.interior(h["vout"]), ((L.interior(h["vin"]) + dv) + dv) * L.interior(h["rdy"]))


# ------------------------------------------------------------------- fvtp2d
def test_fvtp2d_constant_tracer():
    """q = 1, area = ra = 1, xfx = yfx = 1/2: every PPM flux equals the upwind
    value (al - q = almq = br = 0.333332 > 0, so smt5 = 0), q_i = q_j = 1
    exactly, fx1 = fy1 = (0.5 (1 + 1)) * 1/2 = 1/2, fx2 = fy2 = 1 on their
    regions."""
    L, h = inputs("fvtp2d", 8, 7, 2, seed=8)
    h["q"][:] = 1.0
    for n in ("area", "ra_x", "ra_y"):
        h[n][:] = 1.0
    h["xfx"][:] = 0.5
    h["yfx"][:] = 0.5
    B.run_oracle("fvtp2d", h, L)
    for o, v in (("q_i", 1.0), ("q_j", 1.0), ("fx2", 1.0), ("fy2", 1.0), ("fx1", 0.5), ("fy1", 0.5)):
        assert (out_region("fvtp2d", o, L, h[o]) == v).all(), o


def test_fvtp2d_transpose_symmetry():
    """Transposing i<->j and swapping crx/cry, xfx/yfx, ra_x/ra_y maps q_i<->q_j,
    fx1<->fy1, fx2<->fy2 (transposed), bit for bit, regions included."""
    n = 9
    L, h = inputs("fvtp2d", n, n, 2, seed=9, align=1)
    t = copy(h)

    def tr(a):
        return _flat(L.view(a).transpose(0, 2, 1))

    t["q"], t["area"] = tr(h["q"]), tr(h["area"])
    t["crx"], t["cry"] = tr(h["cry"]), tr(h["crx"])
    t["xfx"], t["yfx"] = tr(h["yfx"]), tr(h["xfx"])
    t["ra_x"], t["ra_y"] = tr(h["ra_y"]), tr(h["ra_x"])
    B.run_oracle("fvtp2d", h, L)
    B.run_oracle("fvtp2d", t, L)
    for a, b in (("q_i", "q_j"), ("fx1", "fy1"), ("fx2", "fy2")):
        x = L.view(h[a])
        y = L.view(t[b]).transpose(0, 2, 1)
        assert np.array_equal(np.isnan(x), np.isnan(y)), (a, b)
        ok = ~np.isnan(x)
        assert np.array_equal(x[ok], y[ok]), (a, b)


def test_fvtp2d_branches_exercised():
    """The baseline inputs drive both Courant branches and the smt5 limiter."""
    L, h = inputs("fvtp2d", 16, 16, 2, seed=10)
    cry = out_region("fvtp2d", "fy1", L, h["cry"])
    assert (cry > 0).any() and (cry <= 0).any()
    q = L.view(h["q"])[L.hk]
    d = np.diff(q[3:-3], axis=1)
    assert ((d[:, 1:] * d[:, :-1]) < 0).any()  # local extrema exist -> smt5 = 1 somewhere
