"""Runs a stencil through liboeb on cuda:0 with host inputs (tests only)."""
from __future__ import annotations

import numpy as np

from oeb import stencils
from oeb.layout import Layout, Layout1D
from oeb.spec import KERNELS
from oeb.storage import Storage1D, Storage3D


def to_device(name: str, host: dict, L: Layout) -> dict:
    L1 = Layout1D(L.ny, L.h)
    dev = {}
    for a in KERNELS[name].args:
        if a.kind == "scratch":
            continue
        if a.kind == "in1":
            dev[a.name] = Storage1D.to_device(L1, host[a.name])
        else:
            dev[a.name] = Storage3D.to_device(L, host[a.name])
    return dev


def run_gpu(name: str, host: dict, L: Layout, scalars: dict | None = None) -> dict[str, np.ndarray]:
    """Uploads every array (outputs included, so their NaN fill travels), runs
    the stencil on torch's current stream, and returns the output buffers."""
    import torch

    dev = to_device(name, host, L)
    stencils.run(name, dev, scalars, domain=(L.nx, L.ny, L.nz))
    torch.cuda.synchronize()
    return {o: dev[o].to_host() for o in KERNELS[name].outputs}
