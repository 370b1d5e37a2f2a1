"""The checkpointed adjoint chain (oeb.adjoint_chain): its gradient through
the HIP adjoints against the gradient of the same chain through the ATen
compositions (torch.autograd), with and without checkpointing, FP64; and, on
the CPU through the ATen compositions, that its constants keep the state
bounded over 100 steps in both precisions."""
import numpy as np
import pytest

import aten_ref as A
from oeb import adjoint_chain as AC
from oeb.layout import Layout


class _Aten:
    """Stand-ins for the oeb.autograd functions, built from aten_ref."""

    class Hdiffsa:
        @staticmethod
        def apply(x, c, L, dom, scalars=None):
            return A.hdiffsa(x, c["mask"].buf, c["crlato"].buf, c["crlatu"].buf, L)

    class Hadvuv5th:
        @staticmethod
        def apply(u, v, c, L, dom, scalars=None):
            from oeb.spec import KERNELS

            sc = dict(KERNELS["hadvuv5th"].scalars)
            return A.hadvuv5th(u, v, c["acrlat0"].buf, c["acrlat1"].buf, c["tgrlatda0"].buf, c["tgrlatda1"].buf, L,
                               sc["eddlat"], sc["eddlon"])

    class Fvtp2d:
        @staticmethod
        def apply(q, c, L, dom, scalars=None):
            from oeb.spec import KERNELS

            sc = dict(KERNELS["fvtp2d"].scalars)
            return A.fvtp2d(q, *[c[n].buf for n in ("crx", "cry", "ra_x", "ra_y", "xfx", "yfx", "area")], L,
                            sc["p1"], sc["p2"])

    class Fastwavesuv:
        @staticmethod
        def apply(u, v, pp, c, L, dom, scalars=None):
            from oeb.spec import KERNELS

            sc = dict(KERNELS["fastwavesuv"].scalars)
            sc.update(scalars or {})
            uo, vo = A.fastwavesuv(u, v, pp, *[c[n].buf for n in ("utens", "vtens", "wgtfac", "hhl", "rho", "fx")],
                                   L, sc["edadlat"], sc["dt"])
            return uo.to(u.dtype), vo.to(v.dtype)


def _use_aten(monkeypatch):
    for name in ("Hdiffsa", "Hadvuv5th", "Fvtp2d", "Fastwavesuv"):
        monkeypatch.setattr(AC.AG, name, getattr(_Aten, name))


def _grad(chain, steps, segment, ybar):
    import torch

    x = [t.clone().requires_grad_(True) for t in chain.state0]
    out = chain.run(steps, x, segment=segment)
    loss = sum((o * yb).sum() for o, yb in zip(out, ybar))
    return [g.detach() for g in torch.autograd.grad(loss, x)], [o.detach() for o in out]


@pytest.mark.gpu
def test_adjoint_chain_matches_aten(monkeypatch):
    import torch

    L = Layout(32, 28, 4, h=3, hk=1)
    chain = AC.AdjointChain(L, "f64", seed=3)
    g = torch.Generator(device="cuda").manual_seed(9)
    ybar = [torch.randn(t.shape, generator=g, device="cuda", dtype=t.dtype) * chain.interior_mask
            for t in chain.state0]
    print("[gpu] adjoint chain vs aten", flush=True)
    ours, out = _grad(chain, 6, 0, ybar)
    ours_ckpt, out_ckpt = _grad(chain, 6, 2, ybar)
    _use_aten(monkeypatch)
    theirs, out_ref = _grad(chain, 6, 0, ybar)
    for a, b in zip(out, out_ref):
        assert float((a - b).abs().max() / b.abs().max()) <= 1e-12
    for a, b, c in zip(ours, ours_ckpt, theirs):
        assert torch.equal(a, b), "checkpointing changed the gradient"
        err = float((a - c).abs().max() / c.abs().max())
        assert err <= 1e-10, f"chain gradient vs ATen autograd: rel max {err:.2e}"


def test_transport_field_is_divergence_free():
    """The chain's Courant field: periodic in the domain, and its discrete
    divergence crx(i+1) - crx(i) + cry(j+1) - cry(j) vanishes."""
    L = Layout(32, 28, 3, h=3, hk=1)
    f = AC.transport_fields(L)
    crx = f["crx"].reshape(L.levels, L.rows, L.sj)
    cry = f["cry"].reshape(L.levels, L.rows, L.sj)
    i0, j0 = L.lead, L.h
    cx = crx[L.hk, j0:j0 + L.ny + 1, i0:i0 + L.nx + 1]
    cy = cry[L.hk, j0:j0 + L.ny + 1, i0:i0 + L.nx + 1]
    div = (cx[:-1, 1:] - cx[:-1, :-1]) + (cy[1:, :-1] - cy[:-1, :-1])
    assert np.abs(div).max() == 0.0
    assert np.allclose(cx[:, 0], cx[:, -1]) and np.allclose(cy[0, :], cy[-1, :])
    assert np.array_equal(f["xfx"], f["crx"]) and np.array_equal(f["yfx"], f["cry"])
    assert abs(crx).max() < 1.0 and abs(cry).max() < 1.0


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_adjoint_chain_state_bounded(monkeypatch, precision):
    """100 steps of the chain through the ATen compositions on the CPU: the
    state stays finite and q of order one (the generated q is in [0, 1) plus
    smooth terms); u and v grow slowly under the advective tendencies."""
    import torch

    _use_aten(monkeypatch)
    L = Layout(24, 20, 3, h=3, hk=1)
    chain = AC.AdjointChain(L, precision, seed=0, device="cpu")
    with torch.no_grad():
        q, u, v = chain.run(100, segment=0)
    for t in (q, u, v):
        assert bool(torch.isfinite(t).all())
    assert float(q.abs().max()) < 10.0
    assert float(u.abs().max()) < 1e4 and float(v.abs().max()) < 1e4
