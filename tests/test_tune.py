"""The tuning table (oeb_tune_*) and the autotuner (oeb.tune).

CPU: set / get / clear / remove semantics and argument checks.  GPU: the
autotuner times its candidates and its choice is what the library uses;
results stay bit for bit those of the oracle whatever the choice."""
import numpy as np
import pytest

from common import copy, inputs, out_region
from oeb import abi, tune


def test_tune_table_roundtrip():
    tune.clear()
    dom = (64, 48, 20)
    assert tune.get_choice("hdiffsa", "f64", dom) == (0, -1)
    tune.set_choice("hdiffsa", "f64", dom, 8, -1)
    tune.set_choice("laplace", "f32", dom, 0, 1)
    assert tune.get_choice("hdiffsa", "f64", dom) == (8, -1)
    assert tune.get_choice("hdiffsa", "f32", dom) == (0, -1)  # other element size
    assert tune.get_choice("hdiffsa", "f64", (64, 48, 21)) == (0, -1)  # other shape
    assert tune.get_choice("laplace", "f32", dom) == (0, 1)
    tune.set_choice("hdiffsa", "f64", dom, 0, -1)  # both unset: entry removed
    assert tune.get_choice("hdiffsa", "f64", dom) == (0, -1)
    tune.clear()
    assert tune.get_choice("laplace", "f32", dom) == (0, -1)


def test_knob_table_precedence():
    """The most specific entry wins: this domain and element size, then every
    domain, then both precisions; -1 removes a value."""
    tune.clear()
    dom = (64, 48, 20)
    tune.set_knob("fvtp2d", None, None, "order", 1)
    tune.set_knob("fvtp2d", "f64", None, "order", 2)
    tune.set_knob("fvtp2d", "f64", dom, "order", 3)
    assert tune.get_knob("fvtp2d", "f64", dom, "order") == 3
    assert tune.get_knob("fvtp2d", "f64", None, "order") == 2
    assert tune.get_knob("fvtp2d", None, None, "order") == 1
    assert tune.get_knob("fvtp2d", "f32", dom, "order") == -1  # only the exact key is reported
    tune.set_knob("fvtp2d", "f64", dom, "order", -1)
    assert tune.get_knob("fvtp2d", "f64", dom, "order") == -1
    tune.set_knobs("hdiffsa", "f32", dom, {"form": 1, "rows": 12})
    assert tune.get_knob("hdiffsa", "f32", dom, "rows") == 12
    tune.set_knobs("hdiffsa", "f32", dom, {})
    assert tune.get_knob("hdiffsa", "f32", dom, "rows") == -1
    tune.clear()


@pytest.mark.parametrize("knob,value", [("order", 4), ("lanes", 32), ("depth", 0), ("grid", 0), ("nosuch", 1)])
def test_knob_set_rejects_bad_values(knob, value):
    with pytest.raises(abi.OebError) as e:
        tune.set_knob("laplace", None, None, knob, value)
    assert e.value.status == abi.OEB_ERR_ARG


def test_knob_spaces_name_known_knobs():
    for kernel, branches in tune.SPACES.items():
        for fixed, knobs in branches:
            assert set(fixed) <= set(tune.KNOBS), kernel
            assert all(k in tune.KNOBS and len(v) >= 1 for k, v in knobs), kernel


@pytest.mark.parametrize("kchunk,vector", [(-1, -1), (4097, -1), (4, 2)])
def test_tune_set_rejects_bad_values(kchunk, vector):
    with pytest.raises(abi.OebError) as e:
        tune.set_choice("laplace", "f64", (8, 8, 8), kchunk, vector)
    assert e.value.status == abi.OEB_ERR_ARG


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("name", ["laplace", "hdiffsmag", "nh_p_grad"])
def test_tuner_choice_is_used_and_exact(name, precision, tmp_path):
    from gpu_util import run_gpu, to_device
    from oeb.generators import cast
    from oracle import binding as B

    tune.clear()
    L, host = inputs(name, 64, 40, 12, "baseline", seed=21)
    host = cast(name, host, precision)
    print(f"[gpu] tune {name} {precision}", flush=True)
    t = tune.Tuner(path=tmp_path / "tuning.json", reps=2)
    dev = to_device(name, copy(host), L)
    choice = t.tune(name, dev, (L.nx, L.ny, L.nz), precision)
    t.save()
    assert (tmp_path / "tuning.json").exists()
    for k in tune.KNOBS:
        assert tune.get_knob(name, precision, (L.nx, L.ny, L.nz), k) == choice["knobs"].get(k, -1)
    assert len(choice["candidates"]) >= 4
    gpu = run_gpu(name, copy(host), L)
    ref = copy(host)
    B.run_oracle(name, ref, L, precision=precision)
    for o in gpu:
        x, y = out_region(name, o, L, gpu[o]), out_region(name, o, L, ref[o])
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f"{name}.{o} with knobs {choice['knobs']}"
    tune.clear()
