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


# data keys of a to_dict() payload and the setter that takes each array
# (LI/_geological_interpolator.py:799-862 writes them under "data")
_DATA_SETTERS = {
    "value": "set_value_constraints",
    "gradient": "set_gradient_constraints",
    "normal": "set_normal_constraints",
    "tangent": "set_tangent_constraints",
    "interface": "set_interface_constraints",
    "inequality": "set_value_inequality_constraints",
    "inequality_pairs": "set_inequality_pairs_constraints",
}


def _type_aliases():
    """Every spelling a caller may use for an interpolator type: the short
    registry keys ('FDI', 'PLI', ...) and the enum member names."""
    table = {m.name: m for m in InterpolatorType}
    table.update(interpolator_string_map)
    return table


def _support_from_payload(payload: dict):
    """A support from its to_dict() payload.  Keys the support class does not
    take (a TetMesh has no rotation_xy) are left out, found from the
    constructor's signature rather than by retrying on TypeError."""
    import inspect

    payload = dict(payload)
    kind = payload.pop("type", None)
    if kind is None:
        raise ValueError("No support type specified")
    if isinstance(kind, str):
        kind = SupportType[kind]
    cls = support_map.get(SupportType(int(kind)))
    if cls is None:
        raise NotImplementedError(f"support {SupportType(int(kind)).name} is not part of this "
                                  f"package")
    params = inspect.signature(cls.__init__).parameters
    takes_any = any(p.kind is inspect.Parameter.VAR_KEYWORD for p in params.values())
    kwargs = {k: v for k, v in payload.items() if takes_any or k in params}
    return cls(**kwargs)


class InterpolatorFactory:
    """LI/_interpolator_factory.py:24-140 for the GPU interpolators: the same
    entry points, arguments and errors; types without a GPU class here are
    refused with NotImplementedError."""

    _aliases = _type_aliases()

    @classmethod
    def _normalise_interpolator_type(cls, interpolatortype) -> InterpolatorType:
        if isinstance(interpolatortype, InterpolatorType):
            return interpolatortype
        if isinstance(interpolatortype, str) and interpolatortype in cls._aliases:
            return cls._aliases[interpolatortype]
        return InterpolatorType(interpolatortype)

    @classmethod
    def _class_of(cls, kind: InterpolatorType):
        impl = interpolator_map.get(kind)
        if impl is None:
            raise NotImplementedError(
                f"{kind.name} has no GPU implementation in this package "
                f"(available: {', '.join(t.name for t in interpolator_map)})")
        return impl

    @classmethod
    def create_interpolator(cls, interpolatortype=None, boundingbox=None, nelements=None,
                            element_volume=None, support=None, buffer=None, solver=None):
        if interpolatortype is None:
            raise ValueError("No interpolator type specified")
        if support is None and boundingbox is None:
            raise ValueError("No bounding box specified")
        kind = cls._normalise_interpolator_type(interpolatortype)
        impl = cls._class_of(kind)
        if support is None:
            dim = int(getattr(boundingbox, "dimensions", 3))
            stype = support_interpolator_map[kind].get(dim)
            if stype is None:
                raise NotImplementedError("only 3-D supports are part of this package")
            support = SupportFactory.create_support_from_bbox(
                stype, bounding_box=boundingbox, nelements=nelements,
                element_volume=element_volume, buffer=buffer)
        interp = impl(support)
        for attr, val in (("bounding_box", boundingbox), ("solver", solver)):
            if val is not None:
                setattr(interp, attr, val)
        return interp

    @classmethod
    def from_dict(cls, d):
        """Rebuild an interpolator from its to_dict() payload: support,
        constraint arrays (each non-empty one through its setter), c and the
        up_to_date / valid flags; other keys go to create_interpolator."""
        rest = dict(d)
        kind = rest.pop("type", None)
        if kind is None:
            raise ValueError("No interpolator type specified")
        support = rest.pop("support", None)
        if isinstance(support, dict):
            support = _support_from_payload(support)
        data = rest.pop("data", None) or {}
        c = rest.pop("c", None)
        flags = {"up_to_date": bool(rest.pop("up_to_date", False)),
                 "valid": bool(rest.pop("valid", True))}
        interp = cls.create_interpolator(cls._normalise_interpolator_type(kind), support=support,
                                         **rest)
        arrays = {k: np.asarray(data[k], dtype=float) for k in _DATA_SETTERS
                  if data.get(k) is not None}
        for key, arr in arrays.items():
            if arr.size:
                getattr(interp, _DATA_SETTERS[key])(arr)
        if c is not None and hasattr(interp, "c"):
            interp.c = np.asarray(c, dtype=float)
        for name, val in flags.items():
            setattr(interp, name, val)
        return interp

    @staticmethod
    def get_supported_interpolators():
        return interpolator_map.keys()
