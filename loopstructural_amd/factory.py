"""Registration and reconstruction of this package's interpolators.

The maps and the factory follow the reference's registration point
(LI/__init__.py:94-117: interpolator_string_map, interpolator_map;
LI/_interpolator_factory.py:24-140: InterpolatorFactory;
LC/supports/_support_factory.py:10-25: SupportFactory), restricted to the
two interpolators built here: 'FDI' -> FiniteDifferenceInterpolator on a
StructuredGrid and 'PLI' / 'P1' -> P1Interpolator on a TetMesh.  Other types
are recognised and refused with NotImplementedError.  A support is passed
in, or built from the caller's bounding box object (the reference's
BoundingBox or anything with its origin / maximum / nsteps / nelements /
dimensions / project / with_buffer members), as
SupportFactory.create_support_from_bbox does.
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

# LI/__init__.py:119-126, 3-D entries of the two interpolators built here
support_interpolator_map = {
    InterpolatorType.FINITE_DIFFERENCE: {3: SupportType.StructuredGrid},
    InterpolatorType.PIECEWISE_LINEAR: {3: SupportType.TetMesh},
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

    @staticmethod
    def create_support_from_bbox(support_type, bounding_box, nelements, element_volume=None,
                                 buffer=None, local_coordinates=True):
        """LC/supports/_support_factory.py:37-78: the box (buffered when asked)
        takes the element count, its own nelements setter picks near-cubic
        cells, and the support spans the box in its local frame (or in world
        coordinates with the box's step vector)."""
        if isinstance(support_type, str):
            support_type = SupportType[support_type]
        support_type = SupportType(int(support_type))
        if support_type not in support_map:
            raise NotImplementedError(f"support {support_type.name} is not part of this package")
        if buffer is not None:
            bounding_box = bounding_box.with_buffer(buffer=buffer)
        if element_volume is not None:
            length = np.asarray(bounding_box.maximum, float) - np.asarray(bounding_box.origin, float)
            nelements = int(np.prod(length) / element_volume)
        if nelements is not None:
            bounding_box.nelements = nelements
        nsteps = np.asarray(bounding_box.nsteps).astype(int)
        if local_coordinates:
            local = np.asarray(bounding_box.project(
                np.array([bounding_box.origin, bounding_box.maximum], dtype=float)), dtype=float)
            origin = local.min(axis=0)
            step_vector = (local.max(axis=0) - origin) / nsteps
        else:
            origin = np.asarray(bounding_box.origin, dtype=float)
            step_vector = np.asarray(bounding_box.step_vector, dtype=float)
        return support_map[support_type](origin=origin, step_vector=step_vector,
                                         nsteps_cells=nsteps)


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
            supporttype = support_interpolator_map[interpolatortype].get(
                int(getattr(boundingbox, "dimensions", 3)))
            if supporttype is None:
                raise NotImplementedError("only 3-D supports are part of this package")
            support = SupportFactory.create_support_from_bbox(
                supporttype, bounding_box=boundingbox, nelements=nelements,
                element_volume=element_volume, buffer=buffer)
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
