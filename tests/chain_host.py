"""Host execution of a chain program with the oracle and a numpy halo fill
(tests only): the reference a GPU chain is compared with."""
from __future__ import annotations

import numpy as np

from oeb import abi
from oeb.chain import Call, Halo
from oeb.layout import Layout
from oeb.spec import KERNELS
from oracle import binding as B


def halo_fill_host(a: np.ndarray, L: Layout, h: int, hk: int = 0, mode: int = abi.OEB_HALO_PERIODIC) -> None:
    """In place, the definition of oeb_halo_fill: every point of the box
    [-h, nx+h) x [-h, ny+h) on levels [0, nz) takes the value of the interior
    point at the wrapped (periodic) or clamped (zero-gradient) index; then
    the k-halo planes copy planes 0 and nz-1 over the box."""
    v = L.view(a)
    ii, jj = np.arange(-h, L.nx + h), np.arange(-h, L.ny + h)
    if mode == abi.OEB_HALO_PERIODIC:
        si, sj = ii % L.nx, jj % L.ny
    else:
        si, sj = np.clip(ii, 0, L.nx - 1), np.clip(jj, 0, L.ny - 1)
    box_j = slice(L.h - h, L.h + L.ny + h)
    box_i = slice(L.lead - h, L.lead + L.nx + h)
    for k in range(L.nz):
        plane = v[k + L.hk]
        plane[box_j, box_i] = plane[np.ix_(sj + L.h, si + L.lead)]
    for q in range(hk):
        v[L.hk - 1 - q, box_j, box_i] = v[L.hk, box_j, box_i]
        v[L.hk + L.nz + q, box_j, box_i] = v[L.hk + L.nz - 1, box_j, box_i]


def run_host(program: list, fields: dict, L: Layout, precision: str = "f64", steps: int = 1) -> None:
    """Runs `steps` steps of `program` in place on host arrays."""
    for _ in range(steps):
        for op in program:
            if isinstance(op, Halo):
                for name in op.fields:
                    halo_fill_host(fields[name], L, op.h, op.hk, op.mode)
            else:
                assert isinstance(op, Call)
                args = {a.name: fields[op.bind[a.name]] for a in KERNELS[op.stencil].args if a.name in op.bind}
                B.run_oracle(op.stencil, args, L, op.scalars or None, precision=precision)
