"""The oracle: CPU checkers for the HIP stencils.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this package.  The product path never
does.

  liboracle.so     C restatement of the ten stencils (oracle.c)
  _ref/libref.so   the reference's own kernels compiled in place
                   (ref_harness.cpp), where the reference tree was present
                   at build time
"""
