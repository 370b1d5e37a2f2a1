"""Storage3D / Storage1D: arrays with the reference's descriptor, backed by a
torch tensor on the GPU (the product path) or a numpy array on the host.

Device memory comes from torch's allocator.  torch is the plumbing for
memory and streams only: every stencil runs in liboeb.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import OebStorage1D, OebStorage3D
from .layout import Layout, Layout1D


def desc3(ptr: int, L: Layout) -> OebStorage3D:
    """Descriptor of a flat allocation of L.size elements at address ptr."""
    d = OebStorage3D()
    d.allocatedPtr = ptr
    d.alignedPtr = ptr
    d.offset = L.offset
    d.sizes[:] = L.sizes()
    d.strides[:] = L.strides()
    return d


def desc1(ptr: int, L1: Layout1D) -> OebStorage1D:
    d = OebStorage1D()
    d.allocatedPtr = ptr
    d.alignedPtr = ptr
    d.offset = L1.offset
    d.sizes[0] = L1.size
    d.strides[0] = 1
    return d


def _ptr(buf) -> int:
    if isinstance(buf, np.ndarray):
        if buf.dtype not in (np.float64, np.float32) or not buf.flags.c_contiguous:
            raise TypeError("a storage buffer must be a contiguous float64 or float32 array")
        return buf.ctypes.data
    # torch tensor
    if not (buf.dtype.is_floating_point and buf.element_size() in (4, 8) and buf.is_contiguous()):
        raise TypeError("a storage buffer must be a contiguous float64 or float32 tensor")
    return int(buf.data_ptr())


def on_device(buf) -> bool:
    """True for a torch tensor in GPU memory (numpy arrays and CPU tensors
    are host memory, which a kernel cannot read)."""
    return not isinstance(buf, np.ndarray) and bool(getattr(buf, "is_cuda", False))


def require_device(what: str, arrays: dict) -> None:
    """Raises TypeError naming the first array of `arrays` (name -> Storage or
    None) whose buffer is not in GPU memory."""
    for name, a in arrays.items():
        if a is None:
            continue
        for buf in getattr(a, "bufs", (a.buf,)):
            if not on_device(buf):
                raise TypeError(f"{what}: array '{name}' is a host buffer; pass a GPU tensor (e.g. Storage3D.to_device)")


def _precision(buf) -> str:
    size = buf.itemsize if isinstance(buf, np.ndarray) else buf.element_size()
    return "f64" if size == 8 else "f32"


class Storage3D:
    """A 3-D field: flat float64 or float32 buffer plus its Layout.
    ``precision`` is "f64" or "f32", the element type of the buffer."""

    def __init__(self, layout: Layout, buf):
        if (buf.size if isinstance(buf, np.ndarray) else buf.numel()) != layout.size:
            raise ValueError("buffer size does not match layout")
        self.layout = layout
        self.buf = buf
        self.precision = _precision(buf)
        self._desc = desc3(_ptr(buf), layout)

    @property
    def desc(self) -> OebStorage3D:
        return self._desc

    def ref(self):
        return C.byref(self._desc)

    @classmethod
    def to_device(cls, layout: Layout, host: np.ndarray, device="cuda") -> "Storage3D":
        import torch

        return cls(layout, torch.from_numpy(np.ascontiguousarray(host)).to(device))

    def to_host(self) -> np.ndarray:
        if isinstance(self.buf, np.ndarray):
            return self.buf
        return self.buf.detach().to("cpu").numpy()


class Storage1D:
    def __init__(self, layout: Layout1D, buf):
        if (buf.size if isinstance(buf, np.ndarray) else buf.numel()) != layout.size:
            raise ValueError("buffer size does not match layout")
        self.layout = layout
        self.buf = buf
        self.precision = _precision(buf)
        self._desc = desc1(_ptr(buf), layout)

    @property
    def desc(self) -> OebStorage1D:
        return self._desc

    def ref(self):
        return C.byref(self._desc)

    @classmethod
    def to_device(cls, layout: Layout1D, host: np.ndarray, device="cuda") -> "Storage1D":
        import torch

        return cls(layout, torch.from_numpy(np.ascontiguousarray(host)).to(device))

    def to_host(self) -> np.ndarray:
        if isinstance(self.buf, np.ndarray):
            return self.buf
        return self.buf.detach().to("cpu").numpy()


class DeviationIsland:
    """An FP64 island of the FP32 forms (fastwavesuv's ppuv and hhl, p_grad_c's
    gz and pkc, nh_p_grad's gz, pp and pk3) stored as FP32 deviations from an
    FP64 per-level base profile: value(i, j, k) = base(k) + dev(i, j, k), the
    sum formed in FP64 by the oeb_<name>_f32d entry points.  `dev` is a
    float32 Storage3D on the island's layout, `base` a float64 Storage1D over
    the levels (Layout1D(nz, hk): level k at index k + hk)."""

    precision = "f32d"

    def __init__(self, dev: Storage3D, base: Storage1D):
        if dev.precision != "f32" or base.precision != "f64":
            raise TypeError("DeviationIsland: dev must be float32 and base float64")
        L = dev.layout
        if base.layout.size != L.levels or base.layout.offset != L.hk:
            raise ValueError(f"DeviationIsland: base must cover the {L.levels} levels of the layout "
                             f"(Layout1D({L.nz}, {L.hk}))")
        self.dev, self.base = dev, base
        self.layout = L

    @property
    def buf(self):
        return self.dev.buf

    @property
    def bufs(self):
        return (self.dev.buf, self.base.buf)

    def refs(self):
        return [self.dev.ref(), self.base.ref()]

    @classmethod
    def split(cls, layout: Layout, values: np.ndarray, device="cuda") -> "DeviationIsland":
        """Splits a host float64 island (layout.size values) into the mean of
        each level (over every allocated point of the plane) and the FP32
        deviations from it, on `device`."""
        from .islands import split_host

        dev, base = split_host(layout, values)
        return cls(Storage3D.to_device(layout, dev, device), Storage1D.to_device(level_layout(layout), base, device))


def level_layout(layout: Layout) -> Layout1D:
    """The 1-D layout of a per-level profile of `layout`: the levels
    [-hk, nz + hk), level k at index k + hk."""
    return Layout1D(layout.nz, layout.hk)
