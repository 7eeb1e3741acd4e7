"""suetes_amd -- MI355X-native semi-implicit semi-Lagrangian (SISL) step.

Drop-in for the reference's ``build_dynamical_core('sisl', ...)`` /
``SISLStepper3D.step`` (suetes/regional3d/steppers.py): hand-written HIP
kernels for gfx950 behind a C ABI (include/suetes_sisl.h), torch for device
memory, streams and the per-step glue.
"""
from .geometry import (GalChenSigma, HybridSigma, ObliqueStereographic, RegionalGrid3D, SleveSimple,
                       StretchedSleveSimple, ref_view, to_internal)
from .operators import CGridOperator3D, Coords, tensor_product_interp_3d
from .euler import Euler3D, HyperFilter
from .steppers import (FluxFormAdvector, SemiImplicitSolver3D, SemiLagrangianAdvector3D, SISLStepper3D,
                       VerticalPreconditioner, build_dynamical_core, state_to_internal, state_to_reference)
from .boundaries import BenchmarkSponge, DaviesSponge, XRelaxation

__all__ = [
    "GalChenSigma", "HybridSigma", "ObliqueStereographic", "RegionalGrid3D", "SleveSimple",
    "StretchedSleveSimple", "CGridOperator3D", "Coords", "tensor_product_interp_3d", "Euler3D",
    "HyperFilter", "FluxFormAdvector", "SemiImplicitSolver3D", "SemiLagrangianAdvector3D", "SISLStepper3D",
    "VerticalPreconditioner", "build_dynamical_core", "BenchmarkSponge", "DaviesSponge", "XRelaxation", "ref_view",
    "to_internal", "state_to_internal", "state_to_reference",
]
