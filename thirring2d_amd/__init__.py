"""thirring2d_amd -- MI355X (gfx950) implementation of the Thirring2D
fermion hot path: staggered Dslash, CG on M^dag M, HMC force and
molecular dynamics, fermion-bag determinants.

The numerics live in the HIP library libth2d.so (C ABI: include/th2d.h);
this package is the host-side mirror of the reference's interface.
"""
from . import _native  # noqa: F401
from ._native import PREC_F32, PREC_F64, PREC_MIXED, Th2dError  # noqa: F401

__all__ = ["PREC_F32", "PREC_F64", "PREC_MIXED", "Th2dError"]
