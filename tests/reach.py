"""The read boxes of the ten stencils, and poisoning helpers for the checks
of stencil reach (tests/test_gpu_poison.py on the GPU with NaN,
tests/test_asan.py on the CPU oracle with AddressSanitizer).

The boxes are those of the reference semantics (kernels/<name>.h, the
_openmp variants and fvtp2d's DEFAULT), the same the C-ABI validates each
array against.  The reference leaves halos uninitialised (include/util.h:138)
and never initialises the k-halo planes although fastwavesuv reads planes -1
and nz (kernels/fastwavesuv.h:10-11), so what lies outside each box is
exactly what a kernel must not depend on."""
import numpy as np

from oeb.layout import Layout1D
from oeb.spec import KERNELS

# (i0, i1, j0, j1, k0, k1) offsets: box [i0, nx + i1) x [j0, ny + j1) x [k0, nz + k1).
IN0 = (0, 0, 0, 0, 0, 0)
BOXES = {
    "laplace": {"in": (-1, 1, -1, 1, 0, 0)},
    "hdiffsa": {"in": (-2, 2, -2, 2, 0, 0), "mask": IN0},
    "hdiffsmag": {"uin": (-1, 1, -1, 1, 0, 0), "vin": (-1, 1, -1, 1, 0, 0), "mask": IN0},
    "hadvuv": {"uin": (-2, 2, -2, 2, 0, 0), "vin": (-2, 2, -2, 2, 0, 0)},
    "hadvuv5th": {"uin": (-3, 3, -3, 3, 0, 0), "vin": (-3, 3, -3, 3, 0, 0)},
    "fastwavesuv": {"uin": IN0, "vin": IN0, "utens": IN0, "vtens": IN0, "wgtfac": (0, 1, 0, 1, 0, 1),
                    "ppuv": (0, 1, 0, 1, -1, 1), "hhl": (0, 1, 0, 1, 0, 1), "rho": (0, 1, 0, 1, 0, 0)},
    "uvbke": {"uc": (0, 0, -1, 0, 0, 0), "vc": (-1, 0, 0, 0, 0, 0), "cosa": IN0, "rsina": IN0},
    "p_grad_c": {"uin": IN0, "vin": IN0, "rdxc": IN0, "rdyc": IN0, "delpc": (-1, 0, -1, 0, 0, 0),
                 "gz": (-1, 0, -1, 0, 0, 1), "pkc": (-1, 0, -1, 0, 0, 1)},
    "nh_p_grad": {"uin": IN0, "vin": IN0, "rdx": IN0, "rdy": IN0, "gz": (0, 1, 0, 1, 0, 1),
                  "pp": (0, 1, 0, 1, 0, 1), "pk3": (0, 1, 0, 1, 0, 1), "wk1": (0, 1, 0, 1, 0, 0)},
    "fvtp2d": {"q": (-3, 3, -3, 3, 0, 0), "crx": (0, 1, -3, 2, 0, 0), "cry": (-3, 2, 0, 1, 0, 0),
               "ra_x": (0, 0, -3, 2, 0, 0), "ra_y": (-3, 2, 0, 0, 0, 0), "xfx": (0, 1, -3, 2, 0, 0),
               "yfx": (-3, 2, 0, 1, 0, 0), "area": (-3, 2, -3, 2, 0, 0)},
}
# 1-D arrays: index range [j0, ny + j1).
BOXES_1D = {
    "hdiffsa": {"crlato": (-1, 1), "crlatu": (-1, 1)},
    "hdiffsmag": {"crlavo": (0, 0), "crlavu": (0, 0), "crlato": (0, 0), "crlatu": (0, 0), "acrlat0": (-1, 1)},
    "hadvuv": {n: (0, 0) for n in ("acrlat0", "acrlat1", "tgrlatda0", "tgrlatda1")},
    "hadvuv5th": {n: (0, 0) for n in ("acrlat0", "acrlat1", "tgrlatda0", "tgrlatda1")},
    "fastwavesuv": {"fx": (0, 0)},
}


def keep_mask(L, box) -> np.ndarray:
    """Boolean (levels, rows, sj) mask of the points inside a 3-D box."""
    i0, i1, j0, j1, k0, k1 = box
    keep = np.zeros((L.levels, L.rows, L.sj), dtype=bool)
    keep[k0 + L.hk: L.nz + k1 + L.hk, j0 + L.h: L.ny + j1 + L.h, i0 + L.lead: L.nx + i1 + L.lead] = True
    return keep


def keep_mask_1d(L, box) -> np.ndarray:
    L1 = Layout1D(L.ny, L.h)
    keep = np.zeros(L1.size, dtype=bool)
    keep[box[0] + L1.offset: L.ny + box[1] + L1.offset] = True
    return keep


def output_boxes(name: str, L) -> dict:
    """Each output's region as a 3-D box over all levels."""
    out = {}
    for o in KERNELS[name].outputs:
        i0, i1, j0, j1 = KERNELS[name].region(o, L.nx, L.ny)
        out[o] = (i0, i1 - L.nx, j0, j1 - L.ny, 0, 0)
    return out


def outside_runs(keep: np.ndarray) -> list:
    """(start, stop) element ranges of the flattened array outside `keep`."""
    flat = ~keep.reshape(-1)
    edges = np.flatnonzero(np.diff(np.concatenate(([0], flat.view(np.int8), [0]))))
    return list(zip(edges[0::2].tolist(), edges[1::2].tolist()))
