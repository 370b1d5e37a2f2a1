"""ATen compositions of the differentiated stencils (tests only): the
reference expressions written with torch slicing and torch.where, so that
torch.autograd gives the frozen-branch gradients the HIP adjoints are
checked against.

Fields are flat tensors of a Layout (the allocation).  W(a, di, dj, dk)
is a(i+di, j+dj, k+dk) over a box of (i, j) and all levels.  Outputs are
flat tensors, zero outside their region.
"""
from __future__ import annotations

import torch

from oeb.layout import Layout

R = 1.0 / 6371.229e3  # EARTH_RADIUS_RECIP, include/util.h:17-18


class Box:
    def __init__(self, L: Layout, i0, i1, j0, j1):
        self.L, self.i0, self.i1, self.j0, self.j1 = L, i0, i1, j0, j1

    def w(self, a, di=0, dj=0, dk=0):
        L = self.L
        v = a.view(L.levels, L.rows, L.sj)
        return v[L.hk + dk: L.hk + dk + L.nz, L.h + self.j0 + dj: L.h + self.j1 + dj,
                 L.lead + self.i0 + di: L.lead + self.i1 + di]

    def c1(self, a, dj=0):
        """1-D array over j (Layout1D with offset h) on this box's rows."""
        L = self.L
        return a[L.h + self.j0 + dj: L.h + self.j1 + dj][None, :, None]

    def put(self, x):
        L = self.L
        out = torch.zeros((L.levels, L.rows, L.sj), dtype=x.dtype, device=x.device)
        out[L.hk: L.hk + L.nz, L.h + self.j0: L.h + self.j1, L.lead + self.i0: L.lead + self.i1] = x
        return out.reshape(-1)


def interior(L):
    return Box(L, 0, L.nx, 0, L.ny)


# ------------------------------------------------------------------- hdiffsa
def hdiffsa(x, mask, lo, lu, L: Layout):
    """kernels/hdiffsa.h:141-169."""
    nx, ny = L.nx, L.ny
    B = Box(L, -1, nx + 1, -1, ny + 1)
    lap = (B.w(x, -1, 0) + B.w(x, 1, 0) - 2.0 * B.w(x) + B.c1(lo) * (B.w(x, 0, 1) - B.w(x))
           + B.c1(lu) * (B.w(x, 0, -1) - B.w(x)))
    # lap[:, j+1, i+1] = lap(i, j)
    Bx = Box(L, -1, nx, 0, ny)  # flx(i, j)
    raw = lap[:, 1:ny + 1, 1:nx + 2] - lap[:, 1:ny + 1, 0:nx + 1]
    flx = torch.where(raw * (Bx.w(x, 1, 0) - Bx.w(x)) > 0, torch.zeros_like(raw), raw)
    By = Box(L, 0, nx, -1, ny)  # fly(i, j)
    raw = By.c1(lo) * (lap[:, 1:ny + 2, 1:nx + 1] - lap[:, 0:ny + 1, 1:nx + 1])
    fly = torch.where(raw * (By.w(x, 0, 1) - By.w(x)) > 0, torch.zeros_like(raw), raw)
    I = interior(L)
    out = I.w(x) + (flx[:, :, 0:nx] - flx[:, :, 1:nx + 1] + fly[:, 0:ny, :] - fly[:, 1:ny + 1, :]) * I.w(mask)
    return I.put(out)


# ----------------------------------------------------------------- hadvuv5th
def _s5(f, B: Box, plus, along_i):
    d = (lambda k: B.w(f, k, 0)) if along_i else (lambda k: B.w(f, 0, k))
    if plus:
        return (1.0 / 30.0 * d(-3) + -1.0 / 4.0 * d(-2) + d(-1) + -1.0 / 3.0 * d(0) + -1.0 / 2.0 * d(1)
                + 1.0 / 20.0 * d(2))
    return (1.0 / 20.0 * d(-2) + -1.0 / 2.0 * d(-1) + -1.0 / 3.0 * d(0) + d(1) + -1.0 / 4.0 * d(2)
            + 1.0 / 30.0 * d(3))


def _upwind5(f, B, ua, va, eddlat, eddlon):
    x = torch.where(ua > 0, ua * _s5(f, B, True, True), -ua * _s5(f, B, False, True))
    y = torch.where(va > 0, va * _s5(f, B, True, False), -va * _s5(f, B, False, False))
    return eddlat * x + eddlon * y


def hadvuv5th(u, v, a0, a1, t0, t1, L: Layout, eddlat, eddlon):
    """kernels/hadvuv5th.h:177-201."""
    B = interior(L)
    U = (1.0 / 3.0) * (B.w(u, -1, 0) + B.w(u) + B.w(u, 1, 0))
    Vu = 0.25 * (B.w(v, 1, 0) + B.w(v, 1, -1) + B.w(v) + B.w(v, 0, -1))
    uo = _upwind5(u, B, B.c1(a0) * U, R * Vu, eddlat, eddlon) + B.c1(t0) * B.w(u) * Vu
    Uv = 0.25 * (B.w(u, -1, 0) + B.w(u) + B.w(u, 0, 1) + B.w(u, -1, 1))
    V = 1.0 / 3.0 * (B.w(v, 0, -1) + B.w(v) + B.w(v, 0, 1))
    vo = _upwind5(v, B, B.c1(a1) * Uv, R * V, eddlat, eddlon) - B.c1(t1) * Uv * Uv
    return B.put(uo), B.put(vo)


# --------------------------------------------------------------- fastwavesuv
def fastwavesuv(u, v, pp, utens, vtens, wgtfac, hhl, rho, fx, L: Layout, edadlat, dt):
    """kernels/fastwavesuv.h:304-341."""
    B = interior(L)

    def gk(di, dj, dk):
        return B.w(wgtfac, di, dj, dk) * B.w(pp, di, dj, dk) + (1.0 - B.w(wgtfac, di, dj, dk)) * B.w(pp, di, dj, dk - 1)

    gc0 = gk(0, 0, 1) - gk(0, 0, 0)
    gci = gk(1, 0, 1) - gk(1, 0, 0)
    gcj = gk(0, 1, 1) - gk(0, 1, 0)
    h = lambda di, dj, dk: B.w(hhl, di, dj, dk)  # noqa: E731
    ppgu = (B.w(pp, 1, 0) - B.w(pp)) + (gci + gc0) * 0.5 * ((h(0, 0, 1) + h(0, 0, 0)) - (h(1, 0, 1) + h(1, 0, 0))) / (
        (h(0, 0, 1) - h(0, 0, 0)) + (h(1, 0, 1) - h(1, 0, 0)))
    ppgv = (B.w(pp, 0, 1) - B.w(pp)) + (gcj + gc0) * 0.5 * ((h(0, 0, 1) + h(0, 0, 0)) - (h(0, 1, 1) + h(0, 1, 0))) / (
        (h(0, 0, 1) - h(0, 0, 0)) + (h(0, 1, 1) - h(0, 1, 0)))
    uo = B.w(u) + dt * (B.w(utens) - ppgu * (2.0 * B.c1(fx) / (B.w(rho, 1, 0) + B.w(rho))))
    vo = B.w(v) + dt * (B.w(vtens) - ppgv * (2.0 * edadlat / (B.w(rho, 0, 1) + B.w(rho))))
    return B.put(uo), B.put(vo)


# ----------------------------------------------------------------- hdiffsmag
def hdiffsmag(u, v, mask, vo, vu, lo, lu, ac, L: Layout, eddlat, eddlon, tau, weight):
    """kernels/hdiffsmag.h:205-263 (clamps as std::min / std::max)."""
    B = interior(L)

    def T_(di, dj):
        return (B.w(v, di, dj - 1) - B.w(v, di, dj)) * eddlat * R - (B.w(u, di - 1, dj) - B.w(u, di, dj)) * B.c1(
            ac, dj) * eddlon

    def S_(di, dj):
        return (B.w(u, di, dj + 1) - B.w(u, di, dj)) * eddlat * R + (B.w(v, di + 1, dj) - B.w(v, di, dj)) * B.c1(
            ac, dj) * eddlon

    t0, tip, tjp = T_(0, 0), T_(1, 0), T_(0, 1)
    s0, sim, sjm = S_(0, 0), S_(-1, 0), S_(0, -1)
    hw = weight * B.w(mask)

    def clamp(x):
        x = torch.where(0.0 < x, x, torch.zeros_like(x))
        return torch.where(x < 0.5, x, torch.full_like(x, 0.5))

    smu = clamp(tau * torch.sqrt(0.5 * (tip * tip + t0 * t0) + 0.5 * (sjm * sjm + s0 * s0)) - hw)
    lapu = (B.w(u, 1, 0) + B.w(u, -1, 0) - 2.0 * B.w(u) + B.c1(lo) * (B.w(u, 0, 1) - B.w(u))
            + B.c1(lu) * (B.w(u, 0, -1) - B.w(u)))
    smv = clamp(tau * torch.sqrt(0.5 * (tjp * tjp + t0 * t0) + 0.5 * (sim * sim + s0 * s0)) - hw)
    lapv = (B.w(v, 1, 0) + B.w(v, -1, 0) - 2.0 * B.w(v) + B.c1(vo) * (B.w(v, 0, 1) - B.w(v))
            + B.c1(vu) * (B.w(v, 0, -1) - B.w(v)))
    return B.put(B.w(u) + smu * lapu), B.put(B.w(v) + smv * lapv)


# -------------------------------------------------------------------- fvtp2d
def _ppm(f, c, p1, p2, along_i, B: Box):
    """PPM flux at the faces of box B along one axis (kernels/fvtp2d.h:12-34)."""
    d = (lambda a, k: B.w(a, k, 0)) if along_i else (lambda a, k: B.w(a, 0, k))
    al = lambda k: (p1 * (d(f, k - 1) + d(f, k))) + (p2 * (d(f, k - 2) + d(f, k + 1)))  # noqa: E731
    almq_m, br_m = al(-1) - d(f, -1), al(0) - d(f, -1)
    almq_0, br_0 = al(0) - d(f, 0), al(1) - d(f, 0)
    b0_m, b0_0 = almq_m + br_m, almq_0 + br_0
    one, zero = torch.ones_like(br_m), torch.zeros_like(br_m)
    smt_m = torch.where(almq_m * br_m < 0, one, zero)
    smt_0 = torch.where(almq_0 * br_0 < 0, one, zero)
    smt = smt_m + smt_0 * torch.where(smt_m == 0, one, zero)
    cm = B.w(c)
    up = d(f, -1) + ((1.0 - cm) * (br_m - cm * b0_m)) * smt
    dn = d(f, 0) + ((1.0 + cm) * (almq_0 + cm * b0_0)) * smt
    return torch.where(cm > 0, up, dn)


def fvtp2d(q, crx, cry, ra_x, ra_y, xfx, yfx, area, L: Layout, p1, p2):
    """kernels/fvtp2d.h:4-177 (DEFAULT), with the regions of each stage."""
    nx, ny = L.nx, L.ny
    Bfy2 = Box(L, -3, nx + 2, 0, ny + 1)
    fy2 = Bfy2.put(_ppm(q, cry, p1, p2, False, Bfy2))
    Bqi = Box(L, -3, nx + 2, 0, ny)
    q_i = Bqi.put((((Bqi.w(q) * Bqi.w(area)) + Bqi.w(yfx) * Bqi.w(fy2)) - Bqi.w(yfx, 0, 1) * Bqi.w(fy2, 0, 1))
                  / Bqi.w(ra_y))
    I = interior(L)
    fx1p = _ppm(q_i, crx, p1, p2, True, I)
    Bfx2 = Box(L, 0, nx + 1, -3, ny + 2)
    fx2 = Bfx2.put(_ppm(q, crx, p1, p2, True, Bfx2))
    Bqj = Box(L, 0, nx, -3, ny + 2)
    q_j = Bqj.put((((Bqj.w(q) * Bqj.w(area)) + Bqj.w(xfx) * Bqj.w(fx2)) - Bqj.w(xfx, 1, 0) * Bqj.w(fx2, 1, 0))
                  / Bqj.w(ra_x))
    fy1p = _ppm(q_j, cry, p1, p2, False, I)
    fx1 = I.put((0.5 * (fx1p + I.w(fx2))) * I.w(xfx))
    fy1 = I.put((0.5 * (fy1p + I.w(fy2))) * I.w(yfx))
    return q_i, q_j, fx1, fx2, fy1, fy2
