"""MI355X-native VSIPL multiple-FFT library (vsip_ccfftm*/rcfftmop/crfftmop).

The product is the C library jvsip_amd/lib/libvsip_amd.so (include/vsip.h);
this package holds its sources (csrc/) and a ctypes binding for Python callers.
"""
from .vsip import load, VsipLib, AmdLib  # noqa: F401
