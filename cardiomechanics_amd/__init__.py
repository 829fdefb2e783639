"""MI355X-native mechanics hot path of the CardioMechanics cardiac solver:
T10/T4 hyperelastic residual + analytic tangent (Guccione, Holzapfel-Ogden, active
fibre tension, volumetric penalty), deterministic 3x3-block assembly, and the
GMRES/block-Jacobi Newton-Newmark solver, all in libcmg.so (include/cmg.h)."""
from . import materials, mesh  # noqa: F401
