"""The checkpointed adjoint chain (oeb.adjoint_chain): its gradient through
the HIP adjoints against the gradient of the same chain through the ATen
compositions (torch.autograd), with and without checkpointing, FP64."""
import pytest

import aten_ref as A
from oeb import adjoint_chain as AC
from oeb.layout import Layout

pytestmark = pytest.mark.gpu


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
            return A.fastwavesuv(u, v, pp, *[c[n].buf for n in ("utens", "vtens", "wgtfac", "hhl", "rho", "fx")], L,
                                 sc["edadlat"], sc["dt"])


def _grad(chain, steps, segment, ybar):
    import torch

    x = [t.clone().requires_grad_(True) for t in chain.state0]
    out = chain.run(steps, x, segment=segment)
    loss = sum((o * yb).sum() for o, yb in zip(out, ybar))
    return [g.detach() for g in torch.autograd.grad(loss, x)], [o.detach() for o in out]


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
    for name in ("Hdiffsa", "Hadvuv5th", "Fvtp2d", "Fastwavesuv"):
        monkeypatch.setattr(AC.AG, name, getattr(_Aten, name))
    theirs, out_ref = _grad(chain, 6, 0, ybar)
    for a, b in zip(out, out_ref):
        assert float((a - b).abs().max() / b.abs().max()) <= 1e-12
    for a, b, c in zip(ours, ours_ckpt, theirs):
        assert torch.equal(a, b), "checkpointing changed the gradient"
        err = float((a - c).abs().max() / c.abs().max())
        assert err <= 1e-10, f"chain gradient vs ATen autograd: rel max {err:.2e}"
