"""Shared pytest setup: import paths, the `gpu` marker, and a skip for GPU
tests on machines without one."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (str(ROOT), str(ROOT / "open-earth-benchmarks_amd"), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built liboeb.so")


def _have_gpu() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False


def pytest_collection_modifyitems(config, items):
    if _have_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


class _Knobs:
    """Sets launch knobs of the library's tuning table for every domain and
    both precisions (oeb_tune_knob_set); the fixture clears the table after
    the test."""

    def set(self, kernel: str, knob: str, value) -> None:
        from oeb import tune

        tune.set_knob(kernel, None, None, knob, int(value))


@pytest.fixture
def knobs():
    from oeb import tune

    tune.clear()
    yield _Knobs()
    tune.clear()
