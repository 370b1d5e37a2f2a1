"""The oracle against the committed golden fixtures (outputs of the
reference's own kernels, written by tests/golden/make_golden.py).  These
pin the oracle where the reference tree, and so oracle/_ref, is absent."""
import json
from pathlib import Path

import numpy as np
import pytest

from common import bit_diffs, copy, out_region, rel_maxnorm
from golden.make_golden import cases, fixture_name, sha
from oeb.generators import generate
from oeb.layout import Layout
from oeb.spec import KERNELS, ORDER
from oracle import binding as B

GOLDEN = Path(__file__).resolve().parent / "golden"

CASES = [(k,) + c for k in ORDER for c in cases(k)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1]}-{c[2]}x{c[3]}x{c[4]}")
def test_oracle_matches_golden(case):
    kernel, family, nx, ny, nz, align, seed = case
    f = GOLDEN / fixture_name(kernel, family, nx, ny, nz)
    assert f.exists(), f"missing fixture {f.name}: run tests/golden/make_golden.py"
    data = np.load(f, allow_pickle=False)
    meta = json.loads(str(data["meta"]))
    L = Layout(meta["nx"], meta["ny"], meta["nz"], h=meta["h"], hk=meta["hk"], align=meta["align"])
    host = generate(kernel, L, family, meta["seed"])
    for name, digest in meta["input_sha256"].items():
        assert sha(host[name]) == digest, f"generator output for {kernel}.{name} changed; regenerate the fixtures"
    a = copy(host)
    B.run_oracle(kernel, a, L)
    for o in KERNELS[kernel].outputs:
        x = out_region(kernel, o, L, a[o])
        y = data[o]
        assert x.shape == y.shape
        assert bit_diffs(x, y) == 0, f"{kernel}.{o}: oracle differs from golden (rel {rel_maxnorm(x, y):.2e})"


def test_fixtures_are_small():
    total = sum(p.stat().st_size for p in GOLDEN.glob("*.npz"))
    assert 0 < total < 1_500_000
