"""graada3d_amd — MI355X-native engine for GraAda3D's prism-kernel / CG hot path.

The compute runs in libgraeng.so (hand-written HIP for gfx950 behind a C ABI,
include/graeng.h).  This package holds the Python host mirror of the
reference's interfaces and the input builders:

* ``engine``     - Fwd / GaussNewtonInversion / forward (ctypes over the C ABI)
* ``mesh``       - regular prism meshes, anomaly boxes, observation grids
* ``operators``  - regularisation operators (CSR) on regular meshes
* ``workloads``  - seeded benchmark problems
* ``build``      - compiles libgraeng.so with hipcc for gfx950
"""
from .engine import (GZ, GX, GY, TZZ, TXZ, TYZ, TXX, TXY, TYY, Context, Fwd,  # noqa: F401
                     GaussNewtonInversion, GraError, component_bits, flag_from_names, forward,
                     forward_rectilinear)

__all__ = ["Context", "Fwd", "GaussNewtonInversion", "forward", "forward_rectilinear",
           "GraError", "flag_from_names",
           "component_bits", "GZ", "GX", "GY", "TZZ", "TXZ", "TYZ", "TXX", "TXY", "TYY"]
