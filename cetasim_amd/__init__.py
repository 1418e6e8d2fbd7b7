"""MI355X multi-particle tracking turn of CETASim (HIP, gfx950).

The product is the C-ABI library libcetasim_hip.so (include/cetasim_hip.h)
plus the C++ host layer; this package holds its sources (csrc/) and the
Python plumbing the tests and the benchmark use to drive it.
"""
