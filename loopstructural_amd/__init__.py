"""MI355X-native discrete implicit interpolation (LoopStructural's
FiniteDifferenceInterpolator / P1Interpolator path) on hand-written HIP kernels.
"""
from .constraints import (DtypeError, FiniteValueError, ShapeError, UnsupportedCombinationError,
                          ValidationError, VectorError, WeightError)
from .diagnostics import (ConstraintDiagnosticsReport, ConstraintFamilyDiagnostics,
                          RegionCoverageDiagnostics)
from .interpolators import (DiscreteInterpolator, FiniteDifferenceInterpolator,
                            GeologicalInterpolator, InterpolatorType, P1Interpolator)
from .supports import Operator, StructuredGrid, SupportType, TetMesh

__all__ = [
    "ConstraintDiagnosticsReport", "ConstraintFamilyDiagnostics", "DiscreteInterpolator",
    "DtypeError", "FiniteDifferenceInterpolator", "FiniteValueError", "GeologicalInterpolator",
    "InterpolatorType", "Operator", "P1Interpolator", "RegionCoverageDiagnostics", "ShapeError", "StructuredGrid",
    "SupportType", "TetMesh", "UnsupportedCombinationError", "ValidationError", "VectorError",
    "WeightError",
]
__version__ = "0.1.0"
