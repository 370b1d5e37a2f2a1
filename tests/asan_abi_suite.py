"""Runs the C-ABI validation suite (tests/test_abi.py) against the library
built with its host code under AddressSanitizer (make -C
open-earth-benchmarks_amd asan).  Started by tests/test_asan.py in a process
with the clang ASan runtime preloaded; argv[1] is the instrumented library."""
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path[:0] = [str(ROOT / "open-earth-benchmarks_amd"), str(ROOT), str(ROOT / "tests")]

from oeb import abi  # noqa: E402

abi.LIB_PATH = Path(sys.argv[1])
import pytest  # noqa: E402

rc = pytest.main([str(ROOT / "tests" / "test_abi.py"), "-q", "-x", "-p", "no:cacheprovider"])
assert abi.load()._name == str(abi.LIB_PATH), "the suite did not run the instrumented library"
sys.exit(rc)
