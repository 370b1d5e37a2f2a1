"""The checkpointed adjoint chain (oeb.adjoint_chain): its gradient through
the HIP adjoints against the gradient of the same chain through the ATen
compositions (torch.autograd), with and without checkpointing: 6 steps at
32x28x4 in FP64, and the workload itself, 100 steps at 256x256x80
checkpointed every 10, in FP64 and FP32, with directional central
differences of the FP64 chain along perturbations of 1000 seeded points of
each state field; and,
on the CPU through the ATen compositions, that its constants keep the state
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


# The 100-step chain at 256x256x80.  FP64: the HIP adjoints and the ATen
# compositions evaluate the same expressions, so the states and gradients
# agree to rounding (bounds with a margin over the measured values).  FP32:
# both chains run in FP32 (FP64 islands), each with its own rounding order
# (the HIP FP32 forms are those of the reference's float build, fvtp2d's flux
# divergence compensated); over 100 steps the two trajectories, and their
# gradients, drift apart by accumulated FP32 rounding.
# Measured: FP64 states equal, gradients 3e-15; FP32 q 2.7e-3 (state) and
# 2.2e-3 (gradient), u and v 1e-6.
CHAIN_BOUNDS = {"f64": (1e-12, 1e-12), "f32": (1e-2, 1e-2)}  # (state, gradient) relative max-norm


def _chain_directional_differences(chain, L, grads, ybar, steps=100):
    """Directional central differences of J = sum(ybar * state_steps) along
    perturbations of 1000 seeded interior points of q, of u and of v, one
    field at a time, against the adjoint's <grad, direction>.

    Along u and v the chain is smooth at the scale of eps = 1e-6 (upwind
    switches are kinks of |ua|-type, where both branches meet): measured
    agreement 3e-8 and 2e-7, bound 1e-6.  Along q the perturbation reaches
    hdiffsa's flux limiter and fvtp2d's PPM monotonicity switches, which are
    discontinuous; the number of points whose branch flips inside +-eps grows
    with eps, so the difference quotient converges to the adjoint only to
    first order (one step: 1e-3 off at eps 1e-5, 2e-4 at 1e-6, 4e-9 at
    1e-7).  q is therefore checked at eps 1e-7, where 100 steps measured
    1.7e-5, bound 1e-4."""
    import torch

    rng = np.random.default_rng(12)
    for f, (name, eps, tol) in enumerate((("q", 1e-7, 1e-4), ("u", 1e-6, 1e-6), ("v", 1e-6, 1e-6))):
        d = torch.zeros_like(chain.state0[f])
        pts = rng.choice(L.nx * L.ny * L.nz, size=1000, replace=False)
        k, rest = np.divmod(pts, L.nx * L.ny)
        j, i = np.divmod(rest, L.nx)
        d[torch.from_numpy(L.index(i, j, k).astype(np.int64)).to("cuda")] = torch.from_numpy(
            rng.standard_normal(1000)).to("cuda")

        def state(sign):
            x = list(chain.state0)
            x[f] = x[f] + sign * eps * d
            with torch.no_grad():
                return chain.run(steps, x)

        # J is ~1e6 while its directional derivative is O(1): difference the
        # states first, then sum, so that the rounding of J does not swamp
        # the quotient.
        sp, sm = state(1.0), state(-1.0)
        fd = sum(float((yb * (a - b)).sum()) for a, b, yb in zip(sp, sm, ybar)) / (2 * eps)
        adj = float((grads[f] * d).sum())
        print(f"  directional derivative along {name}: adjoint {adj:.12e}, central difference {fd:.12e}, "
              f"rel {abs(adj - fd) / abs(adj):.2e}", flush=True)
        assert abs(adj - fd) <= tol * abs(adj), (name, adj, fd)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.gpu
def test_adjoint_chain_100_steps_256(monkeypatch, precision):
    import torch

    L = Layout(256, 256, 80, h=3, hk=1)
    chain = AC.AdjointChain(L, precision, seed=3)
    g = torch.Generator(device="cuda").manual_seed(11)
    ybar = [torch.randn(t.shape, generator=g, device="cuda", dtype=t.dtype) * chain.interior_mask
            for t in chain.state0]
    print(f"[gpu] adjoint chain 100 steps 256x256x80 {precision}", flush=True)
    ours, out = _grad(chain, 100, 10, ybar)
    for t in out + ours:
        assert bool(torch.isfinite(t).all())
    if precision == "f64":
        _chain_directional_differences(chain, L, ours, ybar)
    _use_aten(monkeypatch)
    # The ATen chain checkpointed every 2 steps: the gradient is the same
    # (checkpointing changes no bit), the recomputed segments smaller.
    theirs, out_ref = _grad(chain, 100, 2, ybar)
    bs, bg = CHAIN_BOUNDS[precision]
    errs_s = [float((a - b).abs().max() / b.abs().max()) for a, b in zip(out, out_ref)]
    errs_g = [float((a - c).abs().max() / c.abs().max()) for a, c in zip(ours, theirs)]
    rms_g = [float((a - c).pow(2).mean().sqrt() / c.pow(2).mean().sqrt()) for a, c in zip(ours, theirs)]
    print(f"  state vs ATen chain: rel max {[f'{e:.2e}' for e in errs_s]}", flush=True)
    print(f"  gradient vs ATen autograd: rel max {[f'{e:.2e}' for e in errs_g]} rms {[f'{e:.2e}' for e in rms_g]}",
          flush=True)
    assert max(errs_s) <= bs, f"state vs ATen chain: rel max {max(errs_s):.2e}"
    assert max(errs_g) <= bg, f"chain gradient vs ATen autograd: rel max {max(errs_g):.2e}"


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
