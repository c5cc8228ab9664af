"""MI355X (gfx950) transient-elasticity topology-optimisation path.

The product is the C-ABI shared library libtopopt_gpu.so (HIP kernels + C++
host solver) declared in include/topopt_gpu.h, plus the C++ mirror of the
reference's LinearElasticity / PDEFilt / Filter / MMA classes.  This Python
package only locates and builds the library and binds its C-ABI.
"""
import os

from . import capi  # noqa: F401

LIB_PATH = capi.LIB_PATH


def library_loaded():
    return capi._lib is not None
