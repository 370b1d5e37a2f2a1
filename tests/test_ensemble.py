"""Ensembles: one persistent launch per op for all members (oeb_batch_*),
against one launch per member per op, against the host oracle, and replayed
from a HIP graph.  The batch API's recording rules are checked on the CPU."""
import ctypes as C

import numpy as np
import pytest

from chain_host import run_host
from common import copy
from oeb import abi
from oeb.chain import Chain, forward_host_fields, forward_program, to_device
from oeb.ensemble import Batch, Ensemble, draw_shapes
from oeb.layout import Layout

STATE = ("u", "v", "u1", "v1", "u2", "v2", "u3", "v3")
SHAPES = [(24, 20, 5), (40, 16, 3), (16, 33, 4), (64, 64, 6), (8, 8, 2)]


def test_draw_shapes():
    s = draw_shapes(48, seed=0)
    assert len(s) == 48 and s == draw_shapes(48, seed=0)
    assert {x for x, _, _ in s} <= {96, 128, 192, 256, 384, 512} and {z for _, _, z in s} <= {60, 80, 137}
    assert len(set(s)) > 20


def test_batch_recording_rules():
    lib = abi.load()
    b = C.c_void_p()
    abi.check(lib.oeb_batch_create(C.byref(b)))
    try:
        abi.check(lib.oeb_batch_record(b))
        assert lib.oeb_batch_record(b) == abi.OEB_ERR_ARG  # already recording on this thread
        assert lib.oeb_batch_launch(b, None) == abi.OEB_ERR_ARG  # not stopped
        r, n = C.c_int(), C.c_int()
        abi.check(lib.oeb_batch_info(b, C.byref(r), C.byref(n)))
        assert r.value == 0 and n.value == 0
    finally:
        abi.check(lib.oeb_batch_destroy(b))
    b2 = C.c_void_p()
    abi.check(lib.oeb_batch_create(C.byref(b2)))
    assert lib.oeb_batch_stop(b2) == abi.OEB_ERR_ARG  # not recording
    abi.check(lib.oeb_batch_destroy(b2))


def _host_members(precision, seed=3):
    out = []
    for m, (nx, ny, nz) in enumerate(SHAPES):
        L = Layout(nx, ny, nz, h=3, hk=1)
        out.append((L, forward_host_fields(L, precision, seed=seed + m)))
    return out


def _ensemble(hosts):
    return Ensemble([Chain(forward_program(), to_device(h, L), (L.nx, L.ny, L.nz)) for L, h in hosts])


def _states(ens):
    return [{n: m.fields[n].to_host().copy() for n in STATE} for m in ens.members]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_batch_matches_plain_and_host(precision):
    import torch

    hosts = _host_members(precision)
    plain, batched = _ensemble(hosts), _ensemble(hosts)
    print(f"[gpu] ensemble batch vs plain {precision}", flush=True)
    b = batched.record()
    try:
        records, launches = b.info()
        assert records == len(SHAPES) * (3 + 4) and launches == 7  # 3 two-field refills, 4 stencils
        for _ in range(3):
            plain.step_plain()
            b.launch()
        torch.cuda.synchronize()
    finally:
        b.close()
    sp, sb = _states(plain), _states(batched)
    for m, (L, h) in enumerate(hosts):
        ref = copy(h)
        run_host(forward_program(), ref, L, precision, steps=3)
        for n in STATE:
            assert np.array_equal(sp[m][n].view(np.uint8), sb[m][n].view(np.uint8)), f"member {m}: {n} batch != plain"
            assert np.array_equal(L.interior(sb[m][n]), L.interior(ref[n])), f"member {m}: {n} differs from host"


@pytest.mark.gpu
def test_batch_graph_replays():
    """Two batch launches recorded in one graph, replayed three times with
    eager plain steps and synchronisations in between."""
    import torch

    hosts = _host_members("f32", seed=9)
    plain, batched = _ensemble(hosts), _ensemble(hosts)
    side = torch.cuda.Stream()
    print("[gpu] ensemble graph", flush=True)
    b = batched.record()
    g = batched.capture(b, side.cuda_stream, steps=2)
    try:
        for r in range(3):
            g.launch(side.cuda_stream)
            side.synchronize()
            plain.step_plain()
            plain.step_plain()
            torch.cuda.synchronize()
            for m, (sp, sb) in enumerate(zip(_states(plain), _states(batched))):
                for n in STATE:
                    assert np.array_equal(sp[n].view(np.uint8), sb[n].view(np.uint8)), f"replay {r} member {m}: {n}"
    finally:
        g.close()
        b.close()
