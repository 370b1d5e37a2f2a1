"""The drivers on the GPU.

- bin/oeb_bench (ElementType double) and bin/oeb_bench_f32 (float) run each
  stencil with the reference's command line, print the reference's header
  and timer lines, verify every output and print a JSON result line whose
  checksums match the oracle on the same fillMath inputs.
- The reference's own drivers (src/<name>.cpp), compiled unchanged against
  this tree's util.h and kernels/<name>.h by oracle/Makefile into
  oracle/_ref/drivers/, run on the GPU with the reference's command line and
  print the reference's timer line."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from common import copy
from oeb.generators import generate
from oeb.layout import Layout
from oeb.spec import KERNELS, ORDER
from oracle import binding as B

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "open-earth-benchmarks_amd"
REF_DRIVERS = ROOT / "oracle" / "_ref" / "drivers"


def _run(args):
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout.strip().splitlines()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("name", ORDER)
def test_driver_runs_verifies_and_matches_oracle(name, precision):
    binary = PKG / "bin" / ("oeb_bench" if precision == "f64" else "oeb_bench_f32")
    if not binary.exists():
        pytest.fail(f"{binary} not built")
    n, nz = 24, 5
    print(f"[gpu] driver {name} {precision} {n}x{n}x{nz}", flush=True)
    lines = _run([binary, name, 0, n, nz, "--count", 3])
    assert lines[0] == "ALGO: 0" and lines[1] == f"domain_size: {n}" and lines[2] == f"domain_height: {nz}"
    assert any(ln.startswith(f"{name}: ") and "3 count, avg" in ln for ln in lines), lines
    for o in KERNELS[name].outputs:
        assert f"verification: {name}.{o} ok, region sum" in "\n".join(lines), o
    res = json.loads(lines[-1])
    assert res["kernel"] == name and res["count"] == 3 and res["gpu_ms"] > 0 and res["verified"] is True
    # The driver's inputs are the reference drivers' fillMath fields with halo 3.
    L = Layout(n, n, nz, h=3, hk=3, align=16)
    host = generate(name, L, "refdriver", 0)
    ref = copy(host)
    B.run_oracle(name, ref, L)
    for o in KERNELS[name].outputs:
        i0, i1, j0, j1 = KERNELS[name].region(o, n, n)
        reg = L.region(ref[o], i0, i1, j0, j1)
        want, want_abs = float(np.sum(reg)), float(np.sum(np.abs(reg)))
        got = res["checksums"][o]
        if precision == "f64":
            # Bit-exact outputs; only the summation order of the checksum differs.
            assert abs(got - want) <= 1e-11 * max(1.0, want_abs), (o, got, want)
        else:
            # Float fillMath inputs and FP32 arithmetic against the FP64 oracle.
            assert abs(got - want) <= 1e-4 * want_abs, (o, got, want)
            assert abs(res["abs_checksums"][o] - want_abs) <= 1e-4 * want_abs, o


@pytest.mark.gpu
@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("name", ORDER)
def test_reference_driver_runs_on_gpu(name, algo):
    """The reference's own driver, built against the drop-in headers, runs
    its timed loop on the GPU (1000 calls, src/<name>.cpp:7) and prints the
    reference's timer line.  Skipped where the reference tree was not there
    to build it from."""
    binary = REF_DRIVERS / name
    if not binary.exists():
        pytest.skip("reference drivers not built (oracle/Makefile ref needs the reference tree)")
    print(f"[gpu] reference driver {name} ALGO {algo}", flush=True)
    lines = _run([binary, algo, 24, 5])
    assert lines[0] == f"ALGO: {algo}" and lines[1] == "domain_size: 24" and lines[2] == "domain_height: 5"
    label = name if algo == 0 else f"{name} fullfusion"
    assert any(ln.startswith(f"{label}: ") and "ms / 1000 count, avg" in ln for ln in lines), lines
