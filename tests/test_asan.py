"""AddressSanitizer runs on the CPU.

- The C oracle built with -fsanitize=address, run with every input point
  outside its read box and every output point outside its region poisoned
  (tests/asan_oracle_reach.py): no access outside them, and a box narrowed by
  one point is caught.
- The library's host code built with -fsanitize=address (make asan; device
  code as usual), running the whole C-ABI validation suite (tests/test_abi.py
  through tests/asan_abi_suite.py): no invalid access in the validation, the
  knob table, the alias and memory checks or the error reporting.

Each runs in a child process with its compiler's ASan runtime preloaded
(gcc's for the oracle, clang's for the library); leak detection is off (the
Python interpreter keeps its allocations until exit)."""
import glob
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
TESTS = ROOT / "tests"
ASAN_OPTIONS = "detect_leaks=0:halt_on_error=1:abort_on_error=0"


def _gcc_asan() -> str:
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not p or not os.path.isabs(p) or not os.path.exists(p):
        pytest.skip("gcc's ASan runtime is not installed")
    return p


def _clang_asan() -> str:
    hits = sorted(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not hits:
        pytest.skip("clang's ASan runtime is not installed")
    return hits[-1]


def _run(preload: str, args: list, timeout: int = 900) -> subprocess.CompletedProcess:
    env = dict(os.environ, LD_PRELOAD=preload, ASAN_OPTIONS=ASAN_OPTIONS)
    return subprocess.run([sys.executable, "-u", *args], env=env, capture_output=True, text=True, timeout=timeout,
                           cwd="/tmp")


@pytest.fixture(scope="module")
def asan_oracle() -> str:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "asan"], check=True)
    return str(ROOT / "oracle" / "_asan" / "liboracle.so")


def test_oracle_reach_under_asan(asan_oracle):
    r = _run(_gcc_asan(), [str(TESTS / "asan_oracle_reach.py"), asan_oracle])
    assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0 and "reach ok" in r.stdout, (r.stdout[-2000:], r.stderr[-2000:])


@pytest.mark.parametrize("shrink", ["laplace:in", "hdiffsa:in", "fastwavesuv:ppuv", "fvtp2d:crx"])
def test_asan_catches_a_read_outside_a_narrowed_box(asan_oracle, shrink):
    """The same run with one box narrowed by a point the oracle does read:
    ASan reports the read, in the oracle."""
    r = _run(_gcc_asan(), [str(TESTS / "asan_oracle_reach.py"), asan_oracle, "--shrink", shrink])
    assert r.returncode != 0
    assert "AddressSanitizer: use-after-poison" in r.stderr and "READ of size" in r.stderr, r.stderr[-3000:]
    first_frame = next(ln for ln in r.stderr.split("READ of size")[1].splitlines() if ln.strip().startswith("#0"))
    assert "/oracle/oracle" in first_frame, first_frame


def test_abi_validation_under_asan():
    subprocess.run(["make", "-s", "-C", str(ROOT / "open-earth-benchmarks_amd"), "-j8", "asan"], check=True)
    lib = ROOT / "open-earth-benchmarks_amd" / "_build" / "asan" / "liboeb.so"
    r = _run(_clang_asan(), [str(TESTS / "asan_abi_suite.py"), str(lib)])
    assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0 and " passed" in r.stdout, (r.stdout[-3000:], r.stderr[-2000:])
