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


@pytest.mark.parametrize("name", ["laplace", "hdiffsa", "p_grad_c", "nh_p_grad", "fvtp2d"])
def test_unbatchable_stencil_refused_while_recording(name):
    """A stencil without a batched form returns OEB_ERR_ARG while this thread
    records a batch (it would otherwise launch at once, before the batch's
    earlier records run), and launches nothing; the refusal comes before any
    descriptor is read, so host descriptors suffice here."""
    from oeb import stencils
    from oeb.layout import Layout1D
    from oeb.spec import KERNELS
    from oeb.storage import Storage1D, Storage3D

    L = Layout(8, 8, 2)
    L1 = Layout1D(8, 3)
    f = {a.name: (Storage1D(L1, np.zeros(L1.size)) if a.kind == "in1" else Storage3D(L, np.zeros(L.size)))
         for a in KERNELS[name].args if a.kind != "scratch"}
    with Batch() as b:
        with pytest.raises(abi.OebError) as e:
            stencils.run(name, f, domain=(8, 8, 2), check_device=False)
        assert e.value.status == abi.OEB_ERR_ARG and "batch" in e.value.message
    assert b.info() == (0, 0)
    b.close()


def test_ensemble_refuses_unbatchable_program():
    from oeb.chain import Call

    ens = Ensemble.__new__(Ensemble)
    ens.program = [Call("hdiffsa", {}), Call("uvbke", {})]
    with pytest.raises(ValueError, match="hdiffsa"):
        ens.record()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_batch_splits_dependent_records(precision):
    """Records of one kernel run concurrently in one launch unless one reads
    or writes what an earlier one writes.  uvbke(u, v -> a, b) then
    uvbke(a, b -> c, d) on one member must run in order (two launches) and
    equal the two calls made one after the other; the same calls on two
    members' disjoint fields merge into one launch."""
    import torch

    from gpu_util import to_device
    from oeb import stencils
    from oeb.generators import cast
    from oeb.storage import Storage3D

    L, host = _inputs_uvbke(precision)
    print(f"[gpu] batch dependency split {precision}", flush=True)

    def fields():
        d = to_device("uvbke", copy(host), L)
        d["c"] = Storage3D(L, torch.full_like(d["ub"].buf, float("nan")))
        d["d"] = Storage3D(L, torch.full_like(d["ub"].buf, float("nan")))
        return d

    def calls(d):
        stencils.uvbke(d["ub"], d["vb"], d["uc"], d["vc"], d["cosa"], d["rsina"], domain=(L.nx, L.ny, L.nz))
        stencils.uvbke(d["c"], d["d"], d["ub"], d["vb"], d["cosa"], d["rsina"], domain=(L.nx, L.ny, L.nz))

    ref, dep = fields(), fields()
    calls(ref)
    with Batch() as b:
        calls(dep)
    assert b.info() == (2, 2)
    b.launch()
    torch.cuda.synchronize()
    b.close()
    for n in ("ub", "vb", "c", "d"):
        x, y = L.interior(dep[n].to_host()), L.interior(ref[n].to_host())
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), n
    m1, m2 = fields(), fields()
    with Batch() as b2:
        for m in (m1, m2):
            stencils.uvbke(m["ub"], m["vb"], m["uc"], m["vc"], m["cosa"], m["rsina"], domain=(L.nx, L.ny, L.nz))
    assert b2.info() == (2, 1)
    b2.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_batch_fastwavesuv_large_level(precision):
    """fastwavesuv on a level of more than 256x256 points (300x260): a plain
    launch takes 128x8 column-pair tiles, a batched launch the 128x4 kind
    that one launch shares with every member of its run.  Recorded into a
    batch, the call equals the plain launch bit for bit, and both equal the
    oracle."""
    import torch

    from common import bit_diffs, inputs, out_region
    from gpu_util import to_device
    from oeb import stencils
    from oeb.generators import cast
    from oeb.spec import KERNELS
    from oracle import binding as B

    L, host = inputs("fastwavesuv", 300, 260, 4, "baseline", seed=37)
    host = cast("fastwavesuv", host, precision)
    print(f"[gpu] batch fastwavesuv 300x260x4 {precision}", flush=True)
    dom = (L.nx, L.ny, L.nz)
    plain = to_device("fastwavesuv", copy(host), L)
    stencils.run("fastwavesuv", plain, domain=dom)
    torch.cuda.synchronize()
    batched = to_device("fastwavesuv", copy(host), L)
    with Batch() as b:
        stencils.run("fastwavesuv", batched, domain=dom)
    assert b.info() == (1, 1)
    b.launch()
    torch.cuda.synchronize()
    b.close()
    ref = copy(host)
    B.run_oracle("fastwavesuv", ref, L, precision=precision)
    for o in KERNELS["fastwavesuv"].outputs:
        x, y = plain[o].to_host(), batched[o].to_host()
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f"{o}: batch != plain"
        nd = bit_diffs(out_region("fastwavesuv", o, L, x), out_region("fastwavesuv", o, L, ref[o]))
        assert nd == 0, f"{o}: {nd} values differ from the oracle"


def _inputs_uvbke(precision):
    from common import inputs
    from oeb.generators import cast

    L, host = inputs("uvbke", 40, 24, 5, "baseline", seed=31)
    return L, cast("uvbke", host, precision)


@pytest.mark.gpu
def test_ensemble_48_member_draw():
    """The ensemble workload itself: 48 members drawn from {96..512}^2 x
    {60, 80, 137} (draw_shapes(48, 0)), FP32 with FP64 islands.  Two steps of
    the batched launches equal two steps of per-member launches bit for bit on
    every member, every state is finite, and the two smallest members equal
    the host chain (the oracle) bit for bit."""
    import torch

    from oeb.ensemble import forward_ensemble

    shapes = draw_shapes(48, 0)
    print(f"[gpu] ensemble 48 members, {sum(x * y * z for x, y, z in shapes)} points", flush=True)
    plain, batched = forward_ensemble(shapes, "f32", seed=0), forward_ensemble(shapes, "f32", seed=0)
    small = sorted(range(48), key=lambda m: shapes[m][0] * shapes[m][1] * shapes[m][2])[:2]
    hosts = {m: {n: f.buf.cpu().numpy().copy() for n, f in plain.members[m].fields.items()} for m in small}
    b = batched.record()
    try:
        records, launches = b.info()
        assert records == 48 * 7 and launches == 7
        for _ in range(2):
            plain.step_plain()
            b.launch()
        torch.cuda.synchronize()
    finally:
        b.close()
    for m, (mp, mb) in enumerate(zip(plain.members, batched.members)):
        for n in STATE:
            x, y = mb.fields[n].buf, mp.fields[n].buf
            assert torch.equal(x.view(torch.int32), y.view(torch.int32)), f"member {m} {shapes[m]}: {n}"
        L = mp.fields["u"].layout
        for n in ("u", "v"):
            assert bool(torch.isfinite(L.interior(mp.fields[n].buf)).all()), f"member {m}: {n} not finite"
    for m in small:
        L = plain.members[m].fields["u"].layout
        ref = hosts[m]
        run_host(forward_program(), ref, L, "f32", steps=2)
        for n in STATE:
            got = plain.members[m].fields[n].buf.cpu().numpy()
            assert np.array_equal(L.interior(got).view(np.uint8), L.interior(ref[n]).view(np.uint8)), \
                f"member {m} {shapes[m]}: {n} differs from the host chain"


@pytest.mark.gpu
def test_autotune_accepts_only_a_faster_step(tmp_path):
    """Ensemble.autotune: every (stencil, member shape) gets a choice; the
    tuned knobs stay in the library's table only when the batched graph step
    measured faster with them, and are all removed otherwise; the batched
    step still equals the plain one either way."""
    import torch

    from oeb import tune

    hosts = _host_members("f32", seed=5)
    ens = _ensemble(hosts)
    side = torch.cuda.Stream()
    tuner = tune.Tuner(path=tmp_path / "tuning.json", reps=2)
    print("[gpu] ensemble autotune", flush=True)
    tune.clear()
    try:
        res = ens.autotune(tuner, "f32", side.cuda_stream, steps=2, reps=2)
        assert set(res) == {"choices", "default_ms", "tuned_ms", "accepted"}
        assert len(res["choices"]) == 4 * len(SHAPES)
        assert res["default_ms"] > 0 and res["tuned_ms"] > 0
        assert res["accepted"] == (res["tuned_ms"] < 0.98 * res["default_ms"])
        for (stencil, dom), v in res["choices"].items():
            for knob in tune.KNOBS:
                got = tune.get_knob(stencil, "f32", dom, knob)
                want = v["knobs"].get(knob, -1) if res["accepted"] else -1
                assert got == want, (stencil, dom, knob, got, want)
        # with the table autotune left, a batched step still equals a plain one
        torch.cuda.synchronize()
        ens2, plain2 = _ensemble(hosts), _ensemble(hosts)
        b = ens2.record()
        try:
            for _ in range(2):
                b.launch()
                plain2.step_plain()
            torch.cuda.synchronize()
        finally:
            b.close()
        for sp, sb in zip(_states(plain2), _states(ens2)):
            for n in STATE:
                assert np.array_equal(sp[n].view(np.uint8), sb[n].view(np.uint8)), n
    finally:
        tune.clear()
