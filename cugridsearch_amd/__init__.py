"""cugridsearch_amd -- brute-force homography grid search for AMD MI355X (gfx950).

The compute lives in libcgs.so (hand-written HIP kernels behind the C-ABI of
include/cgs.h); this package holds the Python binding of that ABI, a host-side
mirror of the reference's CudaGrid / CudaTensor / CudaImage /
CudaGridSearcher API, and the seeded synthetic workloads.
"""
from . import _lib  # noqa: F401
from .api import (CudaGrid, CudaGridSearcher, CudaImage, CudaTensor, Functor, averageAbsoluteDifference,  # noqa: F401
                  averageAbsoluteDifferenceH, calcMI, calcMIstream, calcNCC, calcSQD, centredAbsoluteDifference,
                  eval_point, sumOfAbsoluteDifferences)

__version__ = "0.1.0"
