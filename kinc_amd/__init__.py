"""kinc_amd — MI355X-native pairwise Similarity (KINC) for gfx950.

The compute path is the HIP library kinc_amd/libkinc.so behind the C-ABI of
include/kinc_sim.h; this package holds the Python binding of that ABI
(_native), the ACE-shaped analytic driver (similarity), the EMX/CCM/CMX
formats (formats) and the synthetic GEM generator (gemgen).
"""
from . import gemgen  # noqa: F401

__all__ = ["gemgen"]
