"""kinc_amd — MI355X-native pairwise Similarity (KINC) for gfx950.

The compute path is the HIP library kinc_amd/libkinc.so behind the C-ABI of
include/*.h; this package holds its Python bindings: _native (the
Similarity engine of include/kinc_sim.h: work blocks, pair lists, the
launch-configuration autotuner, HIP graphs), formats (the EMX/CCM/CMX files
and the CCM/CMX writer of include/kinc_formats.h), serving (the query server
of include/kinc_server.h) and rmt (the RMT threshold analytic of
include/kinc_rmt.h); gemgen is the seeded synthetic expression-matrix
generator of the benchmark configurations, and build compiles the library.
The command-line driver (kinc run similarity, import-emx, export-emx,
export-cmx, extract, rmt) is the executable kinc_amd/kinc.
"""
from . import gemgen  # noqa: F401

__all__ = ["gemgen"]
