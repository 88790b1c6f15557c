"""MI355X engine for the quasi-dynamic boundary-integral hot path of
jdpmoore/unicycle (fortran/3d): Green's-function assembly, the traction
matvec and the rate-and-state RHS under Cash-Karp RK45, on HIP for gfx950.
"""
from .engine import (Engine, UcyError, PREC_F64, PREC_SPLIT_BF16, PREC_MIXED,  # noqa: F401
                     PREC_BF16_I8,
                     computeReferenceSystemOkada92, computeReferenceSystemNikkhoo15,
                     computeReferenceSystemVerticalStrainVolume)
from . import configs  # noqa: F401

__all__ = ["Engine", "UcyError", "PREC_F64", "PREC_SPLIT_BF16", "PREC_MIXED", "PREC_BF16_I8",
           "configs",
           "computeReferenceSystemOkada92", "computeReferenceSystemNikkhoo15",
           "computeReferenceSystemVerticalStrainVolume"]
