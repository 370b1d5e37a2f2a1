"""oeb: host side of the MI355X implementation of the ten open-earth-benchmarks
climate stencils.

  abi         ctypes view of include/oeb.h (liboeb.so)
  layout      the i-fastest, halo-padded array layout
  storage     Storage3D / Storage1D descriptors over device or host buffers
  spec        the stencil registry (arguments, scalars, regions, bytes/point)
  generators  seeded synthetic inputs
  stencils    the reference's operator API over liboeb
"""
from . import abi, layout, spec, storage  # noqa: F401
from .abi import OebError, OebLibraryError  # noqa: F401
from .layout import Layout, Layout1D  # noqa: F401
from .spec import KERNELS, ORDER  # noqa: F401

__version__ = "0.1.0"
