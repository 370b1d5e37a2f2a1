"""Array layout: the reference's i-fastest, halo-padded Storage, with optional
pitch padding so that every row's interior starts on a cache-line boundary.

Reference layout (include/util.h:99-141): i stride 1, j stride sx = nx + 2h,
k stride sx * sy, offset = h * (1 + sx + sx * sy), every axis padded by h.
Here k is padded by its own halo ``hk`` (0 or 1 is enough for every stencil).
With ``align`` > 1 the row pitch is rounded up to a multiple of ``align``
elements, and element (0, j, k) starts at a multiple of ``align`` from the
base.  With align = 16 (128 B) the interior rows line up with HBM/L2 cache
lines, so a wave's 512 B load of 64 doubles touches 4 lines instead of 5.
``align = 1`` reproduces the reference layout exactly.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def _round_up(x: int, m: int) -> int:
    return (x + m - 1) // m * m


@dataclass(frozen=True)
class Layout:
    nx: int
    ny: int
    nz: int
    h: int = 3
    hk: int = 1
    align: int = 16

    @property
    def lead(self) -> int:
        """Elements before i = 0 in each row (>= h)."""
        return self.h if self.align <= 1 else _round_up(self.h, self.align)

    @property
    def sj(self) -> int:
        return self.h + self.nx + self.h if self.align <= 1 else _round_up(self.lead + self.nx + self.h, self.align)

    @property
    def rows(self) -> int:
        return self.ny + 2 * self.h

    @property
    def sk(self) -> int:
        return self.sj * self.rows

    @property
    def levels(self) -> int:
        return self.nz + 2 * self.hk

    @property
    def size(self) -> int:
        """Elements in the allocation."""
        return self.sk * self.levels

    @property
    def offset(self) -> int:
        """Index of element (0, 0, 0)."""
        return self.lead + self.h * self.sj + self.hk * self.sk

    def sizes(self) -> tuple[int, int, int]:
        """Reversed extents [k, j, i]; the i extent is the full row pitch."""
        return (self.levels, self.rows, self.sj)

    def strides(self) -> tuple[int, int, int]:
        return (self.sk, self.sj, 1)

    # -- numpy views -------------------------------------------------------
    def view(self, flat: np.ndarray) -> np.ndarray:
        """(levels, rows, sj) view of a flat allocation."""
        return flat.reshape(self.levels, self.rows, self.sj)

    def index(self, i, j, k):
        """Flat index of (i, j, k) (numpy broadcasting allowed)."""
        return self.offset + np.asarray(i) + np.asarray(j) * self.sj + np.asarray(k) * self.sk

    def region(self, flat: np.ndarray, i0: int, i1: int, j0: int, j1: int, k0: int = 0, k1: int | None = None):
        """Sub-array [k0,k1) x [j0,j1) x [i0,i1) (domain-relative indices), as a view (k, j, i)."""
        k1 = self.nz if k1 is None else k1
        v = self.view(flat)
        a = self.lead
        return v[k0 + self.hk : k1 + self.hk, j0 + self.h : j1 + self.h, i0 + a : i1 + a]

    def interior(self, flat: np.ndarray) -> np.ndarray:
        return self.region(flat, 0, self.nx, 0, self.ny)

    def allocated_box(self) -> tuple[range, range, range]:
        """Domain-relative index ranges (i, j, k) covered by the allocation."""
        return (range(-self.lead, self.sj - self.lead), range(-self.h, self.ny + self.h), range(-self.hk, self.nz + self.hk))

    def ij_grid(self):
        """Domain-relative (i, j) index planes over the allocated rows, shape (rows, sj)."""
        i = np.arange(self.sj) - self.lead
        j = np.arange(self.rows) - self.h
        return np.meshgrid(i, j, indexing="xy")

    def k_levels(self) -> np.ndarray:
        return np.arange(self.levels) - self.hk


@dataclass(frozen=True)
class Layout1D:
    """1-D array over j with halo h (include/util.h:99-109: offset h, stride 1)."""

    n: int
    h: int = 3

    @property
    def size(self) -> int:
        return self.n + 2 * self.h

    @property
    def offset(self) -> int:
        return self.h

    def coords(self) -> np.ndarray:
        return np.arange(self.size) - self.h
