"""The allocation-free fvtp2d adjoint (oeb_fvtp2d_adj_f64_noalloc), through
the C-ABI: its q_bar equals that of oeb_fvtp2d_adj_f64 with caller-passed
work arrays bit for bit, on two calls in a row (the second on a smaller
domain, with the first call's values still in the library's scratch), on a
domain whose levels take several chunks of the scratch, and every point
outside q_bar's region keeps its fill.  A domain whose work planes do not fit
the scratch is rejected before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from oeb import abi, stencils
from oeb.generators import generate
from oeb.layout import Layout
from oeb.spec import KERNELS
from oeb.storage import Storage3D

INPUTS = ("q", "crx", "cry", "ra_x", "ra_y", "xfx", "yfx", "area")
BARS = ("q_i_bar", "q_j_bar", "fx1_bar", "fx2_bar", "fy1_bar", "fy2_bar")
P1, P2 = 0.583333, 0.083333


def _problem(nx, ny, nz, seed):
    """Device inputs, the forward's q_i and q_j, and random cotangents."""
    import torch

    from gpu_util import to_device

    L = Layout(nx, ny, nz, h=3, hk=1)
    dev = to_device("fvtp2d", generate("fvtp2d", L, "baseline", seed), L)
    stencils.run("fvtp2d", dev, domain=(nx, ny, nz))
    g = torch.Generator(device="cuda").manual_seed(seed)
    for b in BARS:
        dev[b] = Storage3D(L, torch.randn(L.size, generator=g, device="cuda", dtype=torch.float64))
    return L, dev


def _nan(L):
    import torch

    return Storage3D(L, torch.full((L.size,), float("nan"), device="cuda", dtype=torch.float64))


def _args(dev, q_bar):
    return [dev[n].ref() for n in INPUTS] + [dev["q_i"].ref(), dev["q_j"].ref()] + \
        [dev[n].ref() for n in BARS] + [q_bar.ref()]


def _with_work(L, dev):
    lib = abi.load()
    q_bar, work = _nan(L), [_nan(L) for _ in range(4)]
    abi.check(lib.oeb_fvtp2d_adj_f64(C.byref(abi.OebDomain(L.nx, L.ny, L.nz)), *_args(dev, q_bar),
                                     *[w.ref() for w in work], P1, P2, stencils._stream(None)))
    return q_bar


def _no_alloc(L, dev):
    lib = abi.load()
    q_bar = _nan(L)
    abi.check(lib.oeb_fvtp2d_adj_f64_noalloc(C.byref(abi.OebDomain(L.nx, L.ny, L.nz)), *_args(dev, q_bar), P1, P2,
                                             stencils._stream(None)))
    return q_bar


@pytest.mark.gpu
def test_no_alloc_matches_work_array_form():
    import torch

    # Larger first, then smaller; then 512 x 512 x 16, whose levels take three
    # chunks of the scratch (7, 7 and 2 levels).
    for shape, seed in (((200, 150, 12), 3), ((40, 36, 5), 4), ((512, 512, 16), 5)):
        print(f"[gpu] fvtp2d adjoint no-alloc {shape}", flush=True)
        L, dev = _problem(*shape, seed)
        want = _with_work(L, dev)
        got = _no_alloc(L, dev)
        torch.cuda.synchronize()
        x, y = got.buf.cpu().numpy(), want.buf.cpu().numpy()
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), shape
        reg = L.region(x, -3, L.nx + 3, -3, L.ny + 3)
        assert np.isfinite(reg).all(), shape
        del dev, want, got
        torch.cuda.empty_cache()


@pytest.mark.gpu
def test_no_alloc_rejects_planes_beyond_the_scratch():
    """2048 x 1030: one level's two work planes need more than the 32 MiB
    scratch; the call fails with OEB_ERR_DOMAIN and q_bar stays untouched."""
    import torch

    L = Layout(2048, 1030, 1, h=3, hk=1)
    print("[gpu] fvtp2d adjoint no-alloc too large", flush=True)
    dev = {n: _nan(L) for n in INPUTS + ("q_i", "q_j") + BARS}
    q_bar = _nan(L)
    lib = abi.load()
    rc = lib.oeb_fvtp2d_adj_f64_noalloc(C.byref(abi.OebDomain(L.nx, L.ny, L.nz)), *_args(dev, q_bar), P1, P2,
                                        stencils._stream(None))
    assert rc == abi.OEB_ERR_DOMAIN
    assert b"scratch" in lib.oeb_last_error()
    torch.cuda.synchronize()
    assert torch.isnan(q_bar.buf).all()
