"""Stencil reach of the CPU oracle under AddressSanitizer.  Every input point
outside the stencil's read box (tests/reach.py) and every output point
outside the output's region is poisoned (__asan_poison_memory_region) for
the duration of the oracle call, so a read or write of one of them aborts
with an ASan report; the outputs must equal those of an unpoisoned run.
Both precisions (FP32 arrays are poisoned by whole 8-byte granules only, so
a float that shares a granule with a point inside the box stays readable).

Started by tests/test_asan.py in a process with the gcc ASan runtime
preloaded.  argv: the instrumented liboracle.so; optionally
--shrink NAME:ARG, which narrows that array's box by one point on its low
i side, to show that a read there is caught."""
import argparse
import ctypes as C
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path[:0] = [str(ROOT / "open-earth-benchmarks_amd"), str(ROOT), str(ROOT / "tests")]

import numpy as np  # noqa: E402

from common import copy, inputs, out_region  # noqa: E402
from oeb.generators import cast  # noqa: E402
from oeb.spec import KERNELS, ORDER  # noqa: E402
from oracle import binding as B  # noqa: E402
from reach import BOXES, BOXES_1D, keep_mask, keep_mask_1d, outside_runs, output_boxes  # noqa: E402

GRANULE = 8
_rt = C.CDLL(None)
_poison = _rt.__asan_poison_memory_region
_unpoison = _rt.__asan_unpoison_memory_region
_is_poisoned = _rt.__asan_address_is_poisoned
for f in (_poison, _unpoison):
    f.argtypes, f.restype = [C.c_void_p, C.c_size_t], None
_is_poisoned.argtypes, _is_poisoned.restype = [C.c_void_p], C.c_int


def poison_outside(arr: np.ndarray, keep: np.ndarray) -> int:
    """Poisons the whole granules of `arr` outside `keep`; returns their bytes."""
    base, es, total = arr.ctypes.data, arr.itemsize, 0
    for a, b in outside_runs(keep):
        lo = -(-(base + a * es) // GRANULE) * GRANULE
        hi = (base + b * es) // GRANULE * GRANULE
        if hi > lo:
            _poison(lo, hi - lo)
            total += hi - lo
    return total


def run(name, shape, precision, shrink):
    L, host = inputs(name, *shape, "baseline", seed=7, align=1)
    host = cast(name, host, precision)
    want = copy(host)
    B.run_oracle(name, want, L, precision=precision)
    got = copy(host)
    boxes = dict(BOXES[name])
    if shrink and shrink[0] == name:
        i0, *rest = boxes[shrink[1]]
        boxes[shrink[1]] = (i0 + 1, *rest)
    masks = {a: keep_mask(L, box) for a, box in boxes.items()}
    masks.update({o: keep_mask(L, box) for o, box in output_boxes(name, L).items()})
    masks1 = {a: keep_mask_1d(L, box) for a, box in BOXES_1D.get(name, {}).items()}
    poisoned = 0
    for a, m in list(masks.items()) + list(masks1.items()):
        poisoned += poison_outside(got[a], m)
    assert poisoned > 0
    probe = got[KERNELS[name].outputs[0]]
    assert _is_poisoned(probe.ctypes.data), "the runtime does not poison"
    try:
        B.run_oracle(name, got, L, precision=precision)
    finally:
        for a in list(masks) + list(masks1):
            _unpoison(got[a].ctypes.data, got[a].nbytes)
    for o in KERNELS[name].outputs:
        x, y = out_region(name, o, L, got[o]), out_region(name, o, L, want[o])
        assert np.array_equal(x, y, equal_nan=True), f"{name}.{o}: poisoned run differs"
    print(f"{name} {shape} {precision}: {poisoned} bytes poisoned, outputs equal", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("lib")
    ap.add_argument("--shrink", default="")
    args = ap.parse_args()
    B.ORACLE_SO = Path(args.lib)
    assert Path(B.oracle_lib()._name) == B.ORACLE_SO
    shrink = tuple(args.shrink.split(":")) if args.shrink else None
    names = [shrink[0]] if shrink else ORDER
    for name in names:
        for shape in ((9, 7, 3), (16, 11, 2)):
            for precision in ("f64", "f32"):
                run(name, shape, precision, shrink)
    print("reach ok", flush=True)


if __name__ == "__main__":
    main()
