"""The adjoint workload: a time step hdiffsa -> hadvuv5th -> fvtp2d ->
fastwavesuv on the state (q, u, v), differentiated end to end with the
hand-written HIP adjoints (oeb.autograd) and checkpointed every
`segment` steps, so that a 100-step chain keeps only the segment inputs plus
one segment's intermediates in memory.

The reference couples nothing (each kernel writes only its own outputs), so
the step defines the coupling, in torch (differentiated by autograd):
  q1       = hdiffsa(q)                                  diffusion of the tracer
  au, av   = hadvuv5th(u, v)                             advective tendencies
  u1, v1   = u + tau au, v + tau av
  fluxes   = fvtp2d(q1)
  q2       = q1 - kappa (fx1(i+1) - fx1 + fy1(j+1) - fy1)  flux divergence
  u2, v2   = fastwavesuv(u1, v1; ppuv)                   pressure-gradient update
Before each stencil the halo its inputs are read on is refilled periodically
(an index gather, whose adjoint folds the halo cotangents back onto the
interior).

The generated inputs are sized for one application of each stencil; repeated
100 times they are not all stable, so the chain fixes three of them:
  * hdiffsa's mask is scaled by MASK_SCALE: the fourth-order diffusion with
    the generated mask (up to 0.1) amplifies the grid-scale mode when applied
    every step;
  * fvtp2d transports with a smooth, divergence-free Courant field
    crx(j) = 0.3 cos(2 pi j / ny), cry(i) = 0.2 sin(2 pi i / nx), with
    xfx = crx, yfx = cry and unit ra_x, ra_y, area, so that
    q2 = q1 - (fx1(i+1) - fx1 + fy1(j+1) - fy1) is the conservative
    finite-volume update of that field (random Courant numbers are a
    compressive flow whose divergence grows q geometrically);
  * fastwavesuv runs with dt = FAST_DT instead of the benchmark's 10.
With these q stays of order 1 and u, v of order 10^3 over 100 steps, in both
precisions.
"""
from __future__ import annotations

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from . import autograd as AG
from .generators import cast, generate
from .layout import Layout, Layout1D
from .spec import KERNELS
from .storage import Storage1D, Storage3D

TAU = 1e-3  # advective tendency step
KAPPA = 1.0  # flux-divergence weight (1: the conservative update on unit areas)
FAST_DT = 1e-3  # fastwavesuv dt (the benchmark's 10 makes the state blow up)
MASK_SCALE = 0.1  # hdiffsa mask scale
CRX_AMP, CRY_AMP = 0.3, 0.2  # Courant amplitudes of the transport field


def transport_fields(L: Layout) -> dict:
    """fvtp2d's Courant, flux and area fields for the chain: the
    divergence-free field of the module docstring, on every level and over the
    whole allocation (halo included, so no refill is needed)."""
    i, j = L.ij_grid()
    cx = CRX_AMP * np.cos(2 * np.pi * j / L.ny)
    cy = CRY_AMP * np.sin(2 * np.pi * i / L.nx)
    planes = lambda a: np.broadcast_to(a, (L.levels,) + a.shape).reshape(-1).copy()  # noqa: E731
    crx, cry = planes(cx), planes(cy)
    one = np.ones_like(crx)
    return {"crx": crx, "cry": cry, "xfx": crx.copy(), "yfx": cry.copy(), "ra_x": one, "ra_y": one.copy(),
            "area": one.copy()}


def periodic_index(L: Layout, h: int) -> np.ndarray:
    """Source index of every allocated point after a periodic refill of the
    width-h ring on levels [0, nz): the ring takes the wrapped interior point,
    everything else keeps its own index."""
    idx = np.arange(L.size, dtype=np.int64).reshape(L.levels, L.rows, L.sj)
    ii = np.arange(-h, L.nx + h)
    jj = np.arange(-h, L.ny + h)
    si, sj = ii % L.nx, jj % L.ny
    for k in range(L.nz):
        plane = idx[L.hk + k]
        src = plane[np.ix_(sj + L.h, si + L.lead)].copy()
        plane[L.h - h: L.h + L.ny + h, L.lead - h: L.lead + L.nx + h] = src
    return idx.reshape(-1)


class AdjointChain:
    """The step above on one Layout, FP64 or FP32 (with FP64 islands)."""

    def __init__(self, L: Layout, precision: str = "f64", seed: int = 0, device="cuda"):
        self.L, self.precision, self.domain = L, precision, (L.nx, L.ny, L.nz)
        L1 = Layout1D(L.ny, L.h)
        self.idx = {h: torch.from_numpy(periodic_index(L, h)).to(device) for h in (1, 2, 3)}

        def consts(name, skip, override=None):
            host = generate(name, L, "baseline", seed)
            host.update(override or {})
            host = cast(name, host, precision)
            out = {}
            for a in KERNELS[name].args:
                if a.kind in ("in", "in1") and a.name not in skip:
                    out[a.name] = (Storage1D.to_device(L1, host[a.name], device) if a.kind == "in1"
                                   else Storage3D.to_device(L, host[a.name], device))
            return host, out

        gq = generate("hdiffsa", L, "baseline", seed)
        hq, self.c_hdiffsa = consts("hdiffsa", ("in",), {"mask": MASK_SCALE * gq["mask"]})
        hw, self.c_hadv = consts("hadvuv5th", ("uin", "vin"))
        _, self.c_fv = consts("fvtp2d", ("q",), transport_fields(L))
        hf, self.c_fw = consts("fastwavesuv", ("uin", "vin", "ppuv"))
        self.ppuv = torch.from_numpy(hf["ppuv"]).to(device)
        self.state0 = tuple(torch.from_numpy(np.ascontiguousarray(hq["in"] if n == "q" else hw[n])).to(device)
                            for n in ("q", "uin", "vin"))

    def fill(self, x, h):
        return x[self.idx[h]]

    def step(self, q, u, v):
        L, dom = self.L, self.domain
        q1 = AG.Hdiffsa.apply(self.fill(q, 2), self.c_hdiffsa, L, dom)
        au, av = AG.Hadvuv5th.apply(self.fill(u, 3), self.fill(v, 3), self.c_hadv, L, dom)
        u1, v1 = u + TAU * au, v + TAU * av
        _, _, fx1, _, fy1, _ = AG.Fvtp2d.apply(self.fill(q1, 3), self.c_fv, L, dom)
        fx1, fy1 = self.fill(fx1, 1), self.fill(fy1, 1)
        div = (fx1.roll(-1) - fx1) + (fy1.roll(-L.sj) - fy1)
        q2 = q1 - KAPPA * div
        u2, v2 = AG.Fastwavesuv.apply(u1, v1, self.ppuv, self.c_fw, L, dom, {"dt": FAST_DT})
        # Keep the interior: outputs are zero outside it, and the next step's
        # fills rebuild the halos.
        m = self.interior_mask
        return q2 * m, u2 * m, v2 * m

    @property
    def interior_mask(self):
        if not hasattr(self, "_mask"):
            L = self.L
            m = torch.zeros((L.levels, L.rows, L.sj), dtype=self.state0[0].dtype, device=self.state0[0].device)
            m[L.hk: L.hk + L.nz, L.h: L.h + L.ny, L.lead: L.lead + L.nx] = 1
            self._mask = m.reshape(-1)
        return self._mask

    def run(self, steps: int, state=None, segment: int = 0):
        """`steps` steps from `state` (default: the generated one).  With
        segment > 0 the steps are checkpointed in segments of that many."""
        q, u, v = state if state is not None else self.state0
        m = self.interior_mask
        q, u, v = q * m, u * m, v * m

        def seg(n):
            def f(q, u, v):
                for _ in range(n):
                    q, u, v = self.step(q, u, v)
                return q, u, v
            return f

        done = 0
        while done < steps:
            n = min(segment, steps - done) if segment > 0 else steps - done
            if segment > 0 and torch.is_grad_enabled():
                q, u, v = checkpoint(seg(n), q, u, v, use_reentrant=False)
            else:
                q, u, v = seg(n)(q, u, v)
            done += n
        return q, u, v
