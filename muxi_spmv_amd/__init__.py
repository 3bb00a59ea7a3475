"""MI355X (gfx950) SpMV/SpMM library: hand-written HIP kernels behind a C-ABI.

The Python modules are a host-side mirror of the reference's C++ interface;
all arithmetic runs in libmuxi_spmv_amd.so (built by muxi_spmv_amd.build).
"""
from ._lib import SpmvError, load, LIB_PATH  # noqa: F401

__all__ = ["SpmvError", "load", "LIB_PATH"]
__version__ = "0.1.0"
