"""numrs_amd -- MI355X (gfx950) kernels layer behind NumRS2-style arrays.

The compute lives in libnumrs_amd.so (hand-written HIP for gfx950 behind the C ABI in
include/numrs_amd.h). This package is the host-side mirror of the reference's kernels interface
(`numrs_amd.kernels`), device-buffer helpers (`numrs_amd.device`) and the synthetic-input
generator (`numrs_amd.rng`). The shared library is loaded on first use; every entry point raises
if it is missing -- there is no CPU fallback.
"""
from ._lib import LIB_PATH, NrkError, load  # noqa: F401

__version__ = "0.1.0"
