"""Registration and reconstruction of this package's interpolators.

The maps and the factory follow the reference's registration point
(LI/__init__.py:94-117: interpolator_string_map, interpolator_map;
LI/_interpolator_factory.py:24-140: InterpolatorFactory;
LC/supports/_support_factory.py:10-25: SupportFactory), restricted to the
two interpolators built here: 'FDI' -> FiniteDifferenceInterpolator on a
StructuredGrid and 'PLI' / 'P1' -> P1Interpolator on a TetMesh.  Other types
are recognised and refused with NotImplementedError.  Supports are passed
in; building one from a bounding box (SupportFactory.create_support_from_bbox)
belongs to the model layer and is refused the same way.
"""
from __future__ import annotations

import numpy as np

from .interpolators import FiniteDifferenceInterpolator, InterpolatorType, P1Interpolator
from .supports import StructuredGrid, SupportType, TetMesh

interpolator_string_map = {
    "FDI": InterpolatorType.FINITE_DIFFERENCE,
    "PLI": InterpolatorType.PIECEWISE_LINEAR,
    "P2": InterpolatorType.PIECEWISE_QUADRATIC,
    "P1": InterpolatorType.PIECEWISE_LINEAR,
    "DFI": InterpolatorType.DISCRETE_FOLD,
    "surfe": InterpolatorType.SURFE,
    "FDI_CN": InterpolatorType.FINITE_DIFFERENCE_CONSTANT_NORM,
    "P1_CN": InterpolatorType.PIECEWISE_LINEAR_CONSTANT_NORM,
}

interpolator_map = {
    InterpolatorType.FINITE_DIFFERENCE: FiniteDifferenceInterpolator,
    InterpolatorType.PIECEWISE_LINEAR: P1Interpolator,
}

support_map = {
    SupportType.StructuredGrid: StructuredGrid,
    SupportType.TetMesh: TetMesh,
}


class SupportFactory:
    """LC/supports/_support_factory.py:10-25 for StructuredGrid and TetMesh."""

    @staticmethod
    def create_support(support_type, **kwargs):
        if support_type is None:
            raise ValueError("No support type specified")
        if isinstance(support_type, str):
            support_type = SupportType[support_type]
        support_type = SupportType(int(support_type))
        if support_type not in support_map:
            raise NotImplementedError(f"support {support_type.name} is not part of this package")
        return support_map[support_type](**kwargs)

    @staticmethod
    def from_dict(d):
        d = dict(d)
        support_type = d.pop("type", None)
        if support_type is None:
            raise ValueError("No support type specified")
        return SupportFactory.create_support(support_type, **d)


class InterpolatorFactory:
    """LI/_interpolator_factory.py:24-140 for the GPU interpolators."""

    @staticmethod
    def _normalise_interpolator_type(interpolatortype) -> InterpolatorType:
        if isinstance(interpolatortype, str):
            if interpolatortype in interpolator_string_map:
                return interpolator_string_map[interpolatortype]
            if interpolatortype in InterpolatorType.__members__:
                return InterpolatorType[interpolatortype]
            return InterpolatorType(interpolatortype)
        return interpolatortype

    @staticmethod
    def create_interpolator(interpolatortype=None, boundingbox=None, nelements=None,
                            element_volume=None, support=None, buffer=None, solver=None):
        if interpolatortype is None:
            raise ValueError("No interpolator type specified")
        if support is None and boundingbox is None:
            raise ValueError("No bounding box specified")
        interpolatortype = InterpolatorFactory._normalise_interpolator_type(interpolatortype)
        if interpolatortype not in interpolator_map:
            raise NotImplementedError(
                f"{interpolatortype.name} has no GPU implementation in this package "
                f"(available: {', '.join(t.name for t in interpolator_map)})")
        if support is None:
            raise NotImplementedError(
                "building a support from a bounding box is outside this package; pass support=")
        interpolator = interpolator_map[interpolatortype](support)
        if boundingbox is not None:
            interpolator.bounding_box = boundingbox
        if solver is not None:
            interpolator.solver = solver
        return interpolator

    @staticmethod
    def from_dict(d):
        """Rebuild an interpolator from its to_dict() payload: support,
        constraint arrays, c and the up_to_date / valid flags."""
        d = dict(d)
        interpolator_type = d.pop("type", None)
        if interpolator_type is None:
            raise ValueError("No interpolator type specified")
        interpolator_type = InterpolatorFactory._normalise_interpolator_type(interpolator_type)
        support = d.pop("support", None)
        if isinstance(support, dict):
            payload = dict(support)
            try:
                support = SupportFactory.from_dict(payload)
            except TypeError as exc:
                # TetMesh takes no rotation_xy
                if "rotation_xy" in payload and "rotation_xy" in str(exc):
                    payload.pop("rotation_xy", None)
                    support = SupportFactory.from_dict(payload)
                else:
                    raise
        data = d.pop("data", None)
        c = d.pop("c", None)
        up_to_date = bool(d.pop("up_to_date", False))
        valid = bool(d.pop("valid", True))
        interpolator = InterpolatorFactory.create_interpolator(interpolator_type, support=support,
                                                               **d)
        if data is not None:
            for key, setter in (("value", "set_value_constraints"),
                                ("gradient", "set_gradient_constraints"),
                                ("normal", "set_normal_constraints"),
                                ("tangent", "set_tangent_constraints"),
                                ("interface", "set_interface_constraints"),
                                ("inequality", "set_value_inequality_constraints"),
                                ("inequality_pairs", "set_inequality_pairs_constraints")):
                if data.get(key) is not None:
                    arr = np.asarray(data[key], dtype=float)
                    if arr.size > 0:
                        getattr(interpolator, setter)(arr)
        if c is not None and hasattr(interpolator, "c"):
            interpolator.c = np.asarray(c, dtype=float)
        interpolator.up_to_date = up_to_date
        interpolator.valid = valid
        return interpolator

    @staticmethod
    def get_supported_interpolators():
        return interpolator_map.keys()
