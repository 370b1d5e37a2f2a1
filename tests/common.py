"""Helpers shared by the tests: input sets, region comparison and error norms."""
from __future__ import annotations

import numpy as np

from oeb.generators import generate
from oeb.layout import Layout
from oeb.spec import KERNELS


def inputs(name: str, nx: int, ny: int, nz: int, family: str = "baseline", seed: int = 0, align: int = 16,
           h: int = 3) -> tuple[Layout, dict]:
    L = Layout(nx, ny, nz, h=h, hk=1, align=align)
    return L, generate(name, L, family, seed)


def copy(host: dict) -> dict:
    return {k: v.copy() for k, v in host.items()}


def out_region(name: str, out: str, L: Layout, arr: np.ndarray) -> np.ndarray:
    i0, i1, j0, j1 = KERNELS[name].region(out, L.nx, L.ny)
    return L.region(arr, i0, i1, j0, j1)


def outside_mask(name: str, out: str, L: Layout) -> np.ndarray:
    """Boolean (levels, rows, sj) mask of allocated points outside `out`'s region."""
    m = np.ones((L.levels, L.rows, L.sj), dtype=bool)
    i0, i1, j0, j1 = KERNELS[name].region(out, L.nx, L.ny)
    m[L.hk : L.hk + L.nz, j0 + L.h : j1 + L.h, i0 + L.lead : i1 + L.lead] = False
    return m


def rel_maxnorm(x: np.ndarray, ref: np.ndarray) -> float:
    """max|x - ref| / max|ref| (0 when both are identically zero)."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = float(np.max(np.abs(ref))) if ref.size else 0.0
    num = float(np.max(np.abs(x - ref))) if ref.size else 0.0
    if den == 0.0:
        return num
    return num / den


def bit_diffs(x: np.ndarray, ref: np.ndarray) -> int:
    """Number of elements whose bit patterns differ (same dtype required)."""
    assert x.dtype == ref.dtype, (x.dtype, ref.dtype)
    it = np.int64 if x.dtype == np.float64 else np.int32
    return int(np.count_nonzero(np.ascontiguousarray(x).view(it) != np.ascontiguousarray(ref).view(it)))


def rel_rms(x: np.ndarray, ref: np.ndarray) -> float:
    """RMS(x - ref) / RMS(ref)."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = float(np.sqrt(np.mean(ref * ref))) if ref.size else 0.0
    num = float(np.sqrt(np.mean((x - ref) ** 2))) if ref.size else 0.0
    return num if den == 0.0 else num / den
