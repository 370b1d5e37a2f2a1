"""The drop-in drivers (open-earth-benchmarks_amd/bin/<name>) on the GPU: they run
with the reference's command line and print the timer line and a JSON result.
Their output checksums match the oracle on the same fillMath inputs."""
import json
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from common import copy
from oeb.generators import generate
from oeb.layout import Layout
from oeb.spec import KERNELS, ORDER
from oracle import binding as B

PKG = Path(__file__).resolve().parent.parent / "open-earth-benchmarks_amd"


@pytest.mark.gpu
@pytest.mark.parametrize("name", ORDER)
def test_driver_runs_and_matches_oracle(name):
    binary = PKG / "bin" / name
    if not binary.exists():
        pytest.fail(f"{binary} not built")
    n, nz = 24, 5
    env = dict(os.environ, OEB_COUNT="3")
    print(f"[gpu] driver {name} {n}x{n}x{nz}", flush=True)
    r = subprocess.run([str(binary), "0", str(n), str(nz)], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0] == "ALGO: 0" and lines[1] == f"domain_size: {n}" and lines[2] == f"domain_height: {nz}"
    assert any(ln.startswith(f"{name}: ") and "count, avg" in ln for ln in lines), r.stdout
    res = json.loads(lines[-1])
    assert res["kernel"] == name and res["count"] == 3 and res["gpu_ms"] > 0
    # The drivers' inputs are the reference drivers' fillMath fields with halo 3.
    L = Layout(n, n, nz, h=3, hk=3, align=16)
    host = generate(name, L, "refdriver", 0)
    ref = copy(host)
    B.run_oracle(name, ref, L)
    for o in KERNELS[name].outputs:
        i0, i1, j0, j1 = KERNELS[name].region(o, n, n)
        want = float(np.sum(L.region(ref[o], i0, i1, j0, j1)))
        got = res["checksums"][o]
        assert abs(got - want) <= 1e-11 * max(1.0, abs(want)), (o, got, want)
