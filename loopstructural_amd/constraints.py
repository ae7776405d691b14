"""Constraint array validation with the reference's observable semantics.

Mirrors ``ValueConstraint`` / ``GradientConstraint`` / ``InterfaceConstraint``
(LI/constraints.py:128-287) and the helpers of LI/_validation.py:25-121:
  * arrays are converted to float64 (DtypeError otherwise), must be 2-D with
    the right column count (ShapeError);
  * non-finite weights are replaced by 1.0 (LI/constraints.py:61-85);
  * rows with a non-finite position / data column are dropped
    (LI/_validation.py:54-85);
  * gradient/normal/tangent rows whose vector magnitude is < 1e-14 are dropped
    (LI/constraints.py:233-254);
  * a weight column of 1.0 is appended when absent (LI/constraints.py:167-170,
    257-260).
"""
from __future__ import annotations

import logging

import numpy as np

logger = logging.getLogger(__name__)


class ValidationError(ValueError):
    """Base exception for constraint validation errors (LI/_validation.py:88)."""


class ShapeError(ValidationError):
    """Shape mismatch in a constraint array."""


class DtypeError(ValidationError):
    """Constraint array not convertible to float64."""


class FiniteValueError(ValidationError):
    """Non-finite values in a constraint array."""


class VectorError(ValidationError):
    """Invalid direction vector."""


class WeightError(ValidationError):
    """Invalid weight values."""


class UnsupportedCombinationError(ValidationError):
    """Unsupported combination of constraint families."""


def _to_float(points, name):
    try:
        return np.asarray(points, dtype=float)
    except (TypeError, ValueError) as e:
        raise DtypeError(f"{name} could not be converted to float64. Error: {e}") from e


def _fix_weights(w: np.ndarray) -> np.ndarray:
    bad = ~np.isfinite(w)
    if np.any(bad):
        logger.warning("%d NaN weight value(s) replaced with 1.0 (default weight).",
                       int(np.sum(bad)))
        w = w.copy()
        w[bad] = 1.0
    return w


def _drop_bad(cols: np.ndarray, name: str) -> np.ndarray:
    good = np.isfinite(cols).all(axis=1)
    if not np.all(good):
        logger.warning("%d %s row(s) dropped: position or data columns contain NaN or inf.",
                       int(np.sum(~good)), name)
    return good


def value_array(points, dimensions: int = 3) -> np.ndarray:
    """(N, d+1|d+2) -> validated (M, d+2) [xyz, value, weight]."""
    pts = _to_float(points, "Value constraint array")
    if pts.ndim != 2:
        raise ShapeError("Value constraint array must be 2D")
    if pts.shape[1] not in (dimensions + 1, dimensions + 2):
        raise ShapeError(
            f"Value constraint array must have {dimensions + 1} or {dimensions + 2} columns")
    n = pts.shape[0]
    w = pts[:, dimensions + 1] if pts.shape[1] == dimensions + 2 else np.ones(n)
    w = _fix_weights(np.array(w, dtype=float))
    data = pts[:, : dimensions + 1]
    keep = _drop_bad(data, "value constraint")
    return np.hstack([data[keep], w[keep, None]])


def gradient_array(points, dimensions: int = 3, is_normal: bool = False) -> np.ndarray:
    """(N, 2d|2d+1) -> validated (M, 2d+1) [xyz, vector, weight]."""
    pts = _to_float(points, "Gradient constraint array")
    if pts.ndim != 2:
        raise ShapeError("Gradient constraint array must be 2D")
    if pts.shape[1] not in (2 * dimensions, 2 * dimensions + 1):
        raise ShapeError(
            f"Gradient constraint array must have {2 * dimensions} or {2 * dimensions + 1} columns")
    n = pts.shape[0]
    w = pts[:, 2 * dimensions] if pts.shape[1] == 2 * dimensions + 1 else np.ones(n)
    w = _fix_weights(np.array(w, dtype=float))
    data = pts[:, : 2 * dimensions]
    keep = _drop_bad(data, ("normal" if is_normal else "gradient") + " constraint")
    data, w = data[keep], w[keep]
    mag = np.linalg.norm(data[:, dimensions:], axis=1)
    zero = mag < 1e-14
    if np.any(zero):
        logger.warning("Dropping %d %s constraints with zero or near-zero magnitude vectors.",
                       int(np.sum(zero)), "normal" if is_normal else "gradient")
        data, w = data[~zero], w[~zero]
    return np.hstack([data, w[:, None]])


def interface_array(points, dimensions: int = 3) -> np.ndarray:
    """(N, d+1|d+2) -> validated (M, d+2) [xyz, interface id, weight]."""
    return value_array(points, dimensions)


def _require_finite(arr: np.ndarray, name: str) -> None:
    bad = ~np.isfinite(arr)
    if np.any(bad):
        raise FiniteValueError(f"{name} contains {int(np.sum(bad))} non-finite value(s)")


def _weight_column(pts: np.ndarray, col: int, name: str) -> np.ndarray:
    if pts.shape[1] <= col:
        return np.ones(pts.shape[0])
    return _fix_weights(np.array(pts[:, col], dtype=float))


def inequality_array(points, dimensions: int = 3) -> np.ndarray:
    """(N, d+2|d+3) -> (N, d+3) [xyz, lower, upper, weight]; positions and
    bounds must be finite and lower < upper (LI/constraints.py:290-348)."""
    pts = _to_float(points, "Inequality constraint array")
    if pts.ndim != 2:
        raise ShapeError("Inequality constraint array must be 2D")
    if pts.shape[1] not in (dimensions + 2, dimensions + 3):
        raise ShapeError(f"Inequality constraint array must have {dimensions + 2} or "
                         f"{dimensions + 3} columns")
    w = _weight_column(pts, dimensions + 2, "Inequality constraint")
    _require_finite(pts[:, :dimensions], "Position (X, Y, Z)")
    bounds = pts[:, dimensions:dimensions + 2]
    _require_finite(bounds, "Bound values")
    bad = bounds[:, 0] >= bounds[:, 1]
    if np.any(bad):
        raise ValidationError(
            f"Found {int(np.sum(bad))} inequality constraints with lower_bound >= upper_bound. "
            "For each constraint, lower_bound must be strictly less than upper_bound. "
            f"Invalid constraints at indices: {np.nonzero(bad)[0]}")
    return np.hstack([pts[:, :dimensions + 2], w[:, None]])


def inequality_pair_array(points, dimensions: int = 3) -> np.ndarray:
    """(N, d+1|d+2) -> (N, d+2) [xyz, pair id, weight]; positions and ids must
    be finite (LI/constraints.py:351-400)."""
    pts = _to_float(points, "Inequality pair constraint array")
    if pts.ndim != 2:
        raise ShapeError("Inequality pair constraint array must be 2D")
    if pts.shape[1] not in (dimensions + 1, dimensions + 2):
        raise ShapeError(f"Inequality pair constraint array must have {dimensions + 1} or "
                         f"{dimensions + 2} columns")
    w = _weight_column(pts, dimensions + 1, "Inequality pairs constraint")
    _require_finite(pts[:, :dimensions], "Position (X, Y, Z)")
    _require_finite(pts[:, dimensions], "Pair id column")
    return np.hstack([pts[:, :dimensions + 1], w[:, None]])
