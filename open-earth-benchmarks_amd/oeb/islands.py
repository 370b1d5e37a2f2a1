"""Deviation islands: an FP64 field of the FP32 forms stored as FP32
deviations from an FP64 per-level base profile (include/oeb.h, the
oeb_<name>_f32d entry points).

The base here is the mean of each level over every allocated point of its
plane; any per-level values are valid bases, and the entry points add the
base in FP64 whatever it is.  The stored field then differs from the FP64
field by the FP32 rounding of the deviation: at most 2^-24 of the deviation's
magnitude, for pressure about 1e-2 of the value, where FP32 storage of the
value itself would lose 2^-24 of the value."""
from __future__ import annotations

import numpy as np

from .layout import Layout
from .spec import DEVIATION_STENCILS  # noqa: F401  (re-exported)


def split_host(L: Layout, values: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(FP32 deviations of L.size, FP64 base of L.levels) of a float64 field."""
    v = L.view(np.asarray(values, dtype=np.float64))
    base = v.mean(axis=(1, 2))
    dev = (v - base[:, None, None]).astype(np.float32)
    return dev.reshape(-1), base


def split_device(L: Layout, t):
    """The same split of a flat float64 torch tensor, on its device."""
    import torch

    v = t.reshape(L.levels, L.rows, L.sj)
    base = v.mean(dim=(1, 2))
    dev = (v - base[:, None, None]).to(dtype=torch.float32)
    return dev.reshape(-1).contiguous(), base.contiguous()


def reconstruct(L: Layout, dev: np.ndarray, base: np.ndarray) -> np.ndarray:
    """The FP64 values an _f32d entry point reads: base(k) + dev, added in
    FP64 (the same IEEE sum the kernels form)."""
    d = L.view(np.asarray(dev, dtype=np.float32)).astype(np.float64)
    return (np.asarray(base, dtype=np.float64)[:, None, None] + d).reshape(-1)
