"""MI355X-native discrete implicit interpolation (LoopStructural's
FiniteDifferenceInterpolator / P1Interpolator path) on hand-written HIP kernels.
"""
from .constraints import (DtypeError, FiniteValueError, ShapeError, UnsupportedCombinationError,
                          ValidationError, VectorError, WeightError)
from .diagnostics import (ConstraintDiagnosticsReport, ConstraintFamilyDiagnostics,
                          RegionCoverageDiagnostics)
from .interpolators import (DiscreteInterpolator, FiniteDifferenceInterpolator,
                            GeologicalInterpolator, InterpolatorType, P1Interpolator)
from .supports import Operator, RectilinearGrid, StructuredGrid, SupportType, TetMesh
from .factory import (InterpolatorFactory, SupportFactory, interpolator_map,
                      interpolator_string_map, support_interpolator_map, support_map)

__all__ = [
    "ConstraintDiagnosticsReport", "ConstraintFamilyDiagnostics", "DiscreteInterpolator",
    "DtypeError", "FiniteDifferenceInterpolator", "FiniteValueError", "GeologicalInterpolator",
    "InterpolatorFactory", "InterpolatorType", "Operator", "P1Interpolator", "RegionCoverageDiagnostics", "ShapeError", "StructuredGrid", "RectilinearGrid",
    "SupportFactory", "SupportType", "TetMesh", "UnsupportedCombinationError", "ValidationError", "VectorError",
    "WeightError", "interpolator_map", "interpolator_string_map", "support_interpolator_map",
    "support_map",
]
__version__ = "0.1.0"
