"""Chains: the halo update (oeb_halo_fill_*, oeb_halo_fill_n_*), the forward
workload chain hadvuv5th -> hdiffsmag -> fastwavesuv -> uvbke, and its
HIP-graph replay.

CPU: the host halo fill against its definition, the host chain (oracle plus
host halo fill) staying bounded, and the argument checks of the multi-field
fill.  GPU: the halo fills and the chain, bit for bit against the host, and a
recorded step replayed several times, with eager work and synchronisation in
between, against eager steps.
"""
import ctypes as C

import numpy as np
import pytest

from chain_host import halo_fill_host, run_host
from common import copy
from oeb import abi
from oeb.chain import Chain, forward_host_fields, forward_program, halo_fill, halo_fill_n, to_device
from oeb.layout import Layout
from oeb.storage import Storage3D

STATE = ("u", "v", "u1", "v1", "u2", "v2", "u3", "v3")


def test_halo_fill_host_definition():
    L = Layout(5, 4, 3, h=3, hk=1, align=1)
    a = np.arange(L.size, dtype=np.float64)
    b = a.copy()
    halo_fill_host(b, L, 2, 1, abi.OEB_HALO_PERIODIC)
    f = lambda arr, i, j, k: arr[L.index(i, j, k)]  # noqa: E731
    assert f(b, -1, 0, 0) == f(a, 4, 0, 0) and f(b, 5, 2, 1) == f(a, 0, 2, 1)
    assert f(b, -2, -2, 2) == f(a, 3, 2, 2) and f(b, 6, 5, 0) == f(a, 1, 1, 0)
    assert f(b, 2, 1, -1) == f(b, 2, 1, 0) and f(b, -2, 5, 3) == f(b, -2, 5, 2)
    assert f(b, -3, 0, 0) == f(a, -3, 0, 0)  # outside the width-2 ring: untouched
    assert np.array_equal(L.interior(b), L.interior(a))
    c = a.copy()
    halo_fill_host(c, L, 3, 0, abi.OEB_HALO_ZERO_GRADIENT)
    assert f(c, -3, -3, 1) == f(a, 0, 0, 1) and f(c, 7, 1, 2) == f(a, 4, 1, 2) and f(c, 2, 6, 0) == f(a, 2, 3, 0)
    assert f(c, 0, 0, -1) == f(a, 0, 0, -1)  # hk = 0: k-halo untouched


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_host_chain_stays_bounded(precision):
    L = Layout(24, 20, 6, h=3, hk=1)
    f = forward_host_fields(L, precision, seed=1)
    run_host(forward_program(), f, L, precision, steps=20)
    for name in ("u", "v"):
        x = L.interior(f[name])
        assert np.isfinite(x).all() and np.max(np.abs(x)) < 1e3


HALO_CASES = [(12, 12, 5, 3, 1, abi.OEB_HALO_PERIODIC), (24, 17, 4, 1, 0, abi.OEB_HALO_ZERO_GRADIENT),
              (7, 9, 3, 3, 1, abi.OEB_HALO_ZERO_GRADIENT), (2, 1, 2, 3, 1, abi.OEB_HALO_PERIODIC)]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("case", HALO_CASES, ids=lambda c: "x".join(map(str, c[:3])) + f"h{c[3]}k{c[4]}m{c[5]}")
def test_halo_fill_gpu(case, precision):
    import torch

    from oeb.storage import Storage3D

    nx, ny, nz, h, hk, mode = case
    L = Layout(nx, ny, nz, h=3, hk=1, align=16)
    rng = np.random.default_rng(3)
    a = rng.standard_normal(L.size).astype(np.float64 if precision == "f64" else np.float32)
    print(f"[gpu] halo {case} {precision}", flush=True)
    d = Storage3D.to_device(L, a)
    halo_fill(d, (nx, ny, nz), h, hk, mode)
    torch.cuda.synchronize()
    b = a.copy()
    halo_fill_host(b, L, h, hk, mode)
    assert np.array_equal(d.to_host().view(np.uint8), b.view(np.uint8))


def test_halo_fill_n_rejects_bad_calls():
    """Validation only (host arrays, nothing launched): the field count, the
    fields' layouts, overlap between fields, halo widths and mode."""
    L = Layout(8, 8, 2, h=3, hk=1, align=1)
    a, b = Storage3D(L, np.zeros(L.size)), Storage3D(L, np.zeros(L.size))
    dom = (8, 8, 2)
    for fs, kw, status in [
        ([a] * 9, {}, abi.OEB_ERR_ARG),
        ([a, a], {}, abi.OEB_ERR_ALIAS),
        ([a, Storage3D(Layout(8, 8, 2, h=3, hk=1, align=16), np.zeros(Layout(8, 8, 2, h=3, hk=1, align=16).size))], {},
         abi.OEB_ERR_LAYOUT),
        ([a, b], {"h": 4}, abi.OEB_ERR_LAYOUT),
        ([a, b], {"h": 1, "hk": 2}, abi.OEB_ERR_LAYOUT),
        ([a, b], {"h": 17}, abi.OEB_ERR_ARG),
        ([a, b], {"h": 1, "mode": 7}, abi.OEB_ERR_ARG),
    ]:
        args = {"h": 3, "hk": 0, "mode": abi.OEB_HALO_PERIODIC}
        args.update(kw)
        with pytest.raises(abi.OebError) as e:
            halo_fill_n(fs, dom, args["h"], args["hk"], args["mode"], check_device=False)
        assert e.value.status == status, (len(fs), kw, e.value.message)
    lib = abi.load()
    ptrs = (abi.P3 * 1)(C.pointer(a.desc))
    assert lib.oeb_halo_fill_n_f64(C.byref(abi.OebDomain(8, 8, 2)), ptrs, 0, 1, 0, 0, None) == abi.OEB_ERR_ARG
    assert lib.oeb_halo_fill_n_f64(C.byref(abi.OebDomain(8, 8, 2)), None, 1, 1, 0, 0, None) == abi.OEB_ERR_NULL


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("case", HALO_CASES + [(33, 20, 7, 2, 2, abi.OEB_HALO_PERIODIC)],
                         ids=lambda c: "x".join(map(str, c[:3])) + f"h{c[3]}k{c[4]}m{c[5]}")
@pytest.mark.parametrize("nfields", [1, 3, 8])
def test_halo_fill_n_gpu(case, precision, nfields):
    """nfields fields refilled in one launch, each bit for bit equal to the
    host fill of that field alone; fields of one layout may sit at any
    offsets of one buffer as long as their boxes do not overlap."""
    import torch

    nx, ny, nz, h, hk, mode = case
    L = Layout(nx, ny, nz, h=3, hk=2, align=16)
    rng = np.random.default_rng(5 + nfields)
    dt = np.float64 if precision == "f64" else np.float32
    hosts = [rng.standard_normal(L.size).astype(dt) for _ in range(nfields)]
    print(f"[gpu] halo_fill_n {case} {precision} x{nfields}", flush=True)
    big = torch.from_numpy(np.concatenate(hosts)).to("cuda")
    fs = [Storage3D(L, big[n * L.size:(n + 1) * L.size]) for n in range(nfields)]
    halo_fill_n(fs, (nx, ny, nz), h, hk, mode)
    torch.cuda.synchronize()
    for n in range(nfields):
        b = hosts[n].copy()
        halo_fill_host(b, L, h, hk, mode)
        assert np.array_equal(fs[n].to_host().view(np.uint8), b.view(np.uint8)), f"field {n}"


def _state(fields):
    return {n: fields[n].to_host().copy() for n in STATE}


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_chain_gpu_matches_host(precision):
    import torch

    L = Layout(40, 36, 6, h=3, hk=1)
    host = forward_host_fields(L, precision, seed=2)
    print(f"[gpu] chain eager {precision}", flush=True)
    ch = Chain(forward_program(), to_device(host, L), (L.nx, L.ny, L.nz))
    for _ in range(3):
        ch.step()
    torch.cuda.synchronize()
    ref = copy(host)
    run_host(forward_program(), ref, L, precision, steps=3)
    got = _state(ch.fields)
    for n in STATE:
        assert np.array_equal(L.interior(got[n]), L.interior(ref[n])), f"chain field {n} differs"


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_chain_graph_replays_match_eager(precision):
    """One recorded step, replayed four times on a side stream with eager
    steps of a twin chain and synchronisations in between; after every replay
    the two states agree bit for bit."""
    import torch

    L = Layout(64, 48, 8, h=3, hk=1)
    host = forward_host_fields(L, precision, seed=4)
    graph_chain = Chain(forward_program(), to_device(host, L), (L.nx, L.ny, L.nz))
    eager_chain = Chain(forward_program(), to_device(host, L), (L.nx, L.ny, L.nz))
    side = torch.cuda.Stream()
    print(f"[gpu] chain graph {precision}", flush=True)
    torch.cuda.synchronize()
    g = graph_chain.capture(side.cuda_stream)
    try:
        for r in range(4):
            g.launch(side.cuda_stream)
            side.synchronize()
            eager_chain.step()
            torch.cuda.synchronize()
            a, b = _state(graph_chain.fields), _state(eager_chain.fields)
            for n in STATE:
                assert np.array_equal(a[n].view(np.uint8), b[n].view(np.uint8)), f"replay {r}: field {n} differs"
    finally:
        g.close()
