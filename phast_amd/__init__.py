"""MI355X-native (gfx950 HIP) matrix-free phase-field hot path of PhAST.

Submodules: ``q4`` (2D native-Q4 explicit path), ``solid3d`` (periodic hex8 voxel
operator and solves), ``batch`` (many plates per launch), ``geometry`` (voxel inputs),
``build`` (hipcc build of ``libphast_mi.so``).  Compute entry points need the built
library and a ROCm device; there is no CPU fallback.
"""
__version__ = "0.1.0"
