"""Structured-grid support with the reference's ``StructuredGrid`` interface.

Geometry bookkeeping (origin, steps, node counts, setters) follows
LC/supports/_3d_base_structured.py:19-245; point location, shape functions,
their gradients and evaluation run in the HIP kernels of
libloopstructural_amd (LC/supports/_3d_structured_grid.py:90-454), so every
numerical result is produced on the GPU and copied back as numpy where the
reference returns numpy.
"""
from __future__ import annotations

import logging
from enum import IntEnum

import numpy as np
import torch

from . import _lib
from . import fd_system as fs

logger = logging.getLogger(__name__)


class SupportType(IntEnum):
    """LC/supports/__init__.py support type ids used in ``to_dict``."""

    StructuredGrid2D = 0
    StructuredGrid = 1
    UnStructuredTetMesh = 2
    P1Unstructured2d = 3
    P2Unstructured2d = 4
    BaseUnstructured2d = 5
    BaseStructured = 6
    TetMesh = 10
    P2UnstructuredTetMesh = 11
    P2StructuredTetMesh = 12
    DataSupported = 13
    RectilinearGrid = 14


class LoopException(Exception):
    """LC/supports/_3d_base_structured.py:14."""


class Operator:
    """Finite-difference masks, (3,3,3) indexed [dz+1][dy+1][dx+1]
    (LC/math/finite_difference_stencil.py:12-38)."""

    z = np.zeros((3, 3))
    Dx_mask = np.array([z, [[0.0, 0.0, 0.0], [-0.5, 0.0, 0.5], [0.0, 0.0, 0.0]], z])
    Dy_mask = Dx_mask.swapaxes(1, 2)
    Dz_mask = Dx_mask.swapaxes(0, 2)
    Dxx_mask = np.array([z, [[0, 0, 0], [1, -2, 1], [0, 0, 0]], z])
    Dyy_mask = Dxx_mask.swapaxes(1, 2)
    Dzz_mask = Dxx_mask.swapaxes(0, 2)
    Dxy_mask = np.array([z, [[-0.25, 0, 0.25], [0, 0, 0], [0.25, 0, -0.25]], z]) / np.sqrt(2)
    Dxz_mask = Dxy_mask.swapaxes(0, 1)
    Dyz_mask = Dxy_mask.swapaxes(0, 2)
    Lapacian = np.array([[[0, 0, 0], [0, 1, 0], [0, 0, 0]],
                         [[0, 1, 0], [1, -6, 1], [0, 1, 0]],
                         [[0, 0, 0], [0, 1, 0], [0, 0, 0]]])


def _as_points(pos) -> np.ndarray:
    """check_position, LC/supports/_3d_base_structured.py:295-326."""
    if not isinstance(pos, (np.ndarray, torch.Tensor)):
        try:
            pos = np.array(pos, dtype=float)
        except Exception as e:
            raise ValueError(f"Position array should be a numpy array or list of points, "
                             f"not {type(pos)}") from e
    if isinstance(pos, torch.Tensor):
        pos = pos.detach().cpu().numpy()
    if pos.ndim == 1:
        pos = np.array([pos])
    if pos.ndim != 2:
        raise ValueError("Position array needs to be a list of points or a point")
    return np.ascontiguousarray(pos[:, :3], dtype=float)


class BaseStructuredSupport:
    """Geometry of a regular grid: ``nsteps`` cells per axis -> ``nsteps+1`` nodes."""

    dimension = 3

    def __init__(self, origin=None, nsteps=None, step_vector=None, rotation_xy=None):
        origin = np.zeros(3) if origin is None else np.array(origin, dtype=float)
        nsteps = np.array([10, 10, 10]) if nsteps is None else np.array(nsteps)
        step_vector = np.ones(3) if step_vector is None else np.array(step_vector, dtype=float)
        if np.any(step_vector == 0):
            logger.warning(f"Step vector {step_vector} has zero values")
        if np.any(nsteps == 0):
            raise LoopException("nsteps cannot be zero")
        if np.any(nsteps < 0):
            raise LoopException("nsteps cannot be negative")
        self._nsteps = np.array(nsteps, dtype=int) + 1
        self._step_vector = np.array(step_vector, dtype=float)
        self._origin = np.array(origin, dtype=float)
        self.supporttype = "Base"
        self.type = SupportType.BaseStructured
        self._rotation_xy = np.eye(3)
        self.rotation_xy = rotation_xy
        self.interpolator = None

    # ---------------------------------------------------------- geometry --
    @property
    def nsteps(self):
        return self._nsteps

    @nsteps.setter
    def nsteps(self, nsteps):
        change_factor = nsteps / self.nsteps
        self._step_vector /= change_factor
        self._nsteps = np.array(nsteps, dtype=int)
        self.onGeometryChange()

    @property
    def nsteps_cells(self):
        return self.nsteps - 1

    @property
    def step_vector(self):
        return self._step_vector

    @step_vector.setter
    def step_vector(self, step_vector):
        change_factor = step_vector / self._step_vector
        self._nsteps = np.ceil(self._nsteps / change_factor).astype(int)
        self._step_vector = np.array(step_vector, dtype=float)
        self.onGeometryChange()

    @property
    def origin(self):
        return self._origin

    @origin.setter
    def origin(self, origin):
        origin = np.array(origin, dtype=float)
        length = (self.maximum - origin) / self.step_vector
        self._nsteps = np.ceil(length).astype(np.int64)
        self._nsteps[self._nsteps == 0] = 3
        if np.any(~(self._nsteps > 0)):
            raise ValueError("Cannot resize the interpolation support.")
        self._origin = origin
        self.onGeometryChange()

    @property
    def maximum(self):
        return self.origin + self.nsteps_cells * self.step_vector

    @maximum.setter
    def maximum(self, maximum):
        length = (np.array(maximum, dtype=float) - self.origin) / self.step_vector
        self._nsteps = np.ceil(length).astype(np.int64)
        self._nsteps[self._nsteps == 0] = 3
        if np.any(~(self._nsteps > 0)):
            raise ValueError("Cannot resize the interpolation support.")
        self.onGeometryChange()

    @property
    def rotation_xy(self):
        return self._rotation_xy

    @rotation_xy.setter
    def rotation_xy(self, rotation_xy):
        if rotation_xy is None:
            return
        if isinstance(rotation_xy, (float, int)):
            a = np.deg2rad(rotation_xy)
            rotation_xy = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0],
                                    [0, 0, 1]])
        rotation_xy = np.array(rotation_xy)
        if rotation_xy.shape != (3, 3):
            raise ValueError(f"Rotation matrix should be 3x3, not {rotation_xy.shape}")
        self._rotation_xy = rotation_xy

    @property
    def volume(self):
        return np.prod(self.maximum - self.origin)

    def set_nelements(self, nelements) -> int:
        """LC/supports/_3d_base_structured.py:81-92."""
        ele_vol = self.volume / nelements
        step_vector = np.zeros(3)
        step_vector[:] = ele_vol ** (1.0 / 3.0)
        self.nsteps = np.ceil((self.maximum - self.origin) / step_vector).astype(int)
        return self.n_elements

    @property
    def n_nodes(self):
        return int(np.prod(self.nsteps))

    @property
    def n_elements(self):
        return int(np.prod(self.nsteps_cells))

    @property
    def element_size(self):
        return np.prod(self.step_vector)

    @property
    def element_scale(self):
        return 1.0

    def onGeometryChange(self):
        if self.interpolator is not None:
            self.interpolator.reset()

    def associateInterpolator(self, interpolator):
        self.interpolator = interpolator

    @property
    def nodes(self):
        mx = self.origin + self.nsteps_cells * self.step_vector
        x = np.linspace(self.origin[0], mx[0], self.nsteps[0])
        y = np.linspace(self.origin[1], mx[1], self.nsteps[1])
        z = np.linspace(self.origin[2], mx[2], self.nsteps[2])
        xx, yy, zz = np.meshgrid(x, y, z, indexing="ij")
        return np.array([xx.flatten(order="F"), yy.flatten(order="F"),
                         zz.flatten(order="F")]).T

    def to_dict(self):
        return {"origin": self.origin, "nsteps": self.nsteps, "step_vector": self.step_vector,
                "rotation_xy": self.rotation_xy}

    def __str__(self):
        return (f"LoopStructural interpolation support:  {self.supporttype} \n"
                f"Origin: {self.origin[0]} {self.origin[1]} {self.origin[2]} \n"
                f"Maximum: {self.maximum[0]} {self.maximum[1]} {self.maximum[2]} \n"
                f"Step Vector: {self.step_vector[0]} {self.step_vector[1]} "
                f"{self.step_vector[2]} \n"
                f"Number of Steps: {self.nsteps[0]} {self.nsteps[1]} {self.nsteps[2]} \n"
                f"Degrees of freedon {self.n_nodes}")

    # ------------------------------------------------- index arithmetic --
    def global_index_to_node_index(self, global_index):
        gi = np.asarray(global_index)
        shape = gi.shape
        gi = gi.reshape(-1)
        out = np.zeros((gi.shape[0], 3), dtype=int)
        out[:, 0] = gi % self.nsteps[0]
        out[:, 1] = gi // self.nsteps[0] % self.nsteps[1]
        out[:, 2] = gi // self.nsteps[0] // self.nsteps[1]
        return out.reshape(*shape, 3)

    def global_index_to_cell_index(self, global_index):
        gi = np.asarray(global_index)
        out = np.zeros((gi.shape[0], 3), dtype=int)
        nc = self.nsteps_cells
        out[:, 0] = gi % nc[0]
        out[:, 1] = gi // nc[0] % nc[1]
        out[:, 2] = gi // nc[0] // nc[1]
        return out

    def _global_indicies(self, indexes, nsteps):
        indexes = np.asarray(indexes)
        if indexes.ndim == 1:
            raise ValueError("Cell indexes needs to be Nx3")
        if indexes.shape[-1] != 3:
            raise ValueError("Last dimensions should be ijk indexing")
        shape = indexes.shape
        ix = indexes.reshape(-1, 3)
        gi = ix[:, 0] + nsteps[0] * ix[:, 1] + nsteps[0] * nsteps[1] * ix[:, 2]
        return gi.reshape(shape[:-1])

    def global_node_indices(self, indexes):
        return self._global_indicies(indexes, self.nsteps)

    def global_cell_indices(self, indexes):
        return self._global_indicies(indexes, self.nsteps_cells)

    def cell_corner_indexes(self, cell_indexes):
        """(N,3) cell ijk -> (N,8,3) corner ijk, corner order 000,100,010,110,001,..."""
        ci = np.asarray(cell_indexes)
        off = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1],
                        [0, 1, 1], [1, 1, 1]])
        return ci[:, None, :] + off[None, :, :]

    @property
    def elements(self):
        """(n_cells, 8) corner node ids of every cell, cells x fastest
        (LC/supports/_3d_base_structured.py:216-221)."""
        cells = self.global_index_to_cell_index(np.arange(self.n_elements))
        return self.global_node_indices(self.cell_corner_indexes(cells))

    def node_indexes_to_position(self, node_indexes):
        ni = np.asarray(node_indexes)
        shape = ni.shape
        ni = ni.reshape(-1, 3)
        xyz = self.origin[None, :] + self.step_vector[None, :] * ni
        return xyz.reshape(shape)

    # --------------------------------------------------- GPU-backed ------
    def _lsgrid(self) -> _lib.Grid:
        return _lib.make_grid(self.origin, self.step_vector, self.nsteps)

    def _locate(self, pos, want_w=False, want_T=False):
        pts = _as_points(pos)
        idc, ins, w, T = fs.locate(self._lsgrid(), torch.from_numpy(pts), want_w, want_T)
        return pts, idc, ins, w, T

    def inside(self, pos):
        """LC/supports/_3d_base_structured.py:287-293 (on the GPU)."""
        _, _, ins, _, _ = self._locate(pos)
        return ins.cpu().numpy()

    def position_to_cell_corners(self, pos):
        """(N,8) corner node ids (-1 outside) and (N,) inside flags
        (LC/supports/_3d_base_structured.py:386-412)."""
        _, idc, ins, _, _ = self._locate(pos)
        return idc.cpu().numpy(), ins.cpu().numpy()

    def position_to_cell_index(self, pos):
        """(N,3) cell ijk (0 outside) and inside flags (:251-282)."""
        idc, ins = self.position_to_cell_corners(pos)
        g0 = np.where(ins, idc[:, 0], 0)
        ci = self.global_index_to_node_index(g0)
        ci[~ins] = 0
        return ci, ins

    def position_to_cell_vertices(self, pos):
        gi, inside = self.position_to_cell_corners(pos)
        gi = np.where(gi < 0, 0, gi)
        return self.node_indexes_to_position(self.global_index_to_node_index(gi)), inside


class StructuredGrid(BaseStructuredSupport):
    """Regular cartesian grid (LC/supports/_3d_structured_grid.py:18-510)."""

    def __init__(self, origin=None, nsteps_cells=None, step_vector=None, nsteps=None,
                 rotation_xy=None, properties=None, cell_properties=None,
                 name="StructuredGrid"):
        if nsteps_cells is None:
            nsteps_cells = np.array([10, 10, 10]) if nsteps is None else np.array(nsteps)
        super().__init__(origin, nsteps_cells, step_vector, rotation_xy=rotation_xy)
        self.type = SupportType.StructuredGrid
        self.name = name
        self.regions = {"everywhere": np.ones(self.n_nodes).astype(bool)}

    @property
    def barycentre(self):
        return self.cell_centres(np.arange(self.n_elements))

    def cell_centres(self, global_index):
        ci = self.global_index_to_cell_index(np.asarray(global_index))
        return (self.origin[None, :] + self.step_vector[None, :] * ci
                + self.step_vector[None, :] * 0.5)

    def trilinear(self, local_coords):
        """Shape functions at local coordinates (closed form, corner order
        000,100,010,110,001,101,011,111; :90-124).  Host formula helper: the
        assembly and evaluation paths compute these in the HIP kernels."""
        lc = np.asarray(local_coords, dtype=float)
        x, y, z = lc[:, 0], lc[:, 1], lc[:, 2]
        w = np.zeros((lc.shape[0], 8))
        w[:, 0] = (1 - x) * (1 - y) * (1 - z)
        w[:, 1] = x * (1 - y) * (1 - z)
        w[:, 2] = (1 - x) * y * (1 - z)
        w[:, 4] = (1 - x) * (1 - y) * z
        w[:, 5] = x * (1 - y) * z
        w[:, 6] = (1 - x) * y * z
        w[:, 3] = x * y * (1 - z)
        w[:, 7] = x * y * z
        return w

    def position_to_local_coordinates(self, pos):
        """clip((p - cell origin)/h, 0, 1) (:126-146), from the GPU cell lookup."""
        pts = _as_points(pos)
        ci, _ = self.position_to_cell_index(pts)
        cell_origin = self.origin[None, :] + self.step_vector[None, :] * ci
        return np.clip((pts - cell_origin) / self.step_vector[None, :], 0.0, 1.0)

    def position_to_dof_coefs(self, pos):
        """Trilinear weights (:148-162), computed by the locate kernel."""
        _, _, _, w, _ = self._locate(pos, want_w=True)
        return w.cpu().numpy()

    def get_element_gradient_for_location(self, pos):
        """(vertices, T (N,3,8), elements, inside) (:393-454)."""
        pts, idc, ins, _, T = self._locate(pos, want_T=True)
        vertices, inside = self.position_to_cell_vertices(pts)
        ci, _ = self.position_to_cell_index(pts)
        return vertices, T.cpu().numpy(), self.global_cell_indices(ci), inside

    def get_element_for_location(self, pos):
        pts, idc, ins, w, _ = self._locate(pos, want_w=True)
        vertices, inside = self.position_to_cell_vertices(pts)
        ci, _ = self.position_to_cell_index(pts)
        return vertices, w.cpu().numpy(), self.global_cell_indices(ci), inside

    def neighbour_global_indexes(self, mask=None, **kwargs):
        """Index helper for custom stencils (:164-293)."""
        if "indexes" in kwargs:
            indexes = kwargs["indexes"]
        else:
            gi = np.arange(self.n_nodes)
            ix = self.global_index_to_node_index(gi)
            m = np.all((ix > 0) & (ix < self.nsteps[None, :] - 1), axis=1)
            indexes = ix[m].T
        if indexes.ndim != 2:
            return None
        if mask is None:
            mask = np.array([[c - 1 for a in range(3) for b in range(3) for c in range(3)],
                             [b - 1 for a in range(3) for b in range(3) for c in range(3)],
                             [a - 1 for a in range(3) for b in range(3) for c in range(3)]])
        nb = indexes[:, None, :] + mask[:, :, None]
        return (nb[0] + self.nsteps[0] * nb[1]
                + self.nsteps[0] * self.nsteps[1] * nb[2]).astype(np.int64)

    def _check_property(self, property_array):
        if np.asarray(property_array).shape[0] != self.n_nodes:
            raise ValueError(f"cannot assign {np.asarray(property_array).shape[0]} vlaues "
                             f"to array of shape {self.n_nodes}")

    def evaluate_value(self, evaluation_points, property_array):
        """Trilinear interpolation of node values (:295-322); outside -> NaN."""
        self._check_property(property_array)
        c = property_array if isinstance(property_array, torch.Tensor) else \
            torch.from_numpy(np.ascontiguousarray(property_array, dtype=float))
        pts = _as_points(evaluation_points)
        v, _ = fs.evaluate(self._lsgrid(), c, torch.from_numpy(pts), True, False)
        v = v.cpu().numpy()
        # the support-level call has no NaN-coordinate special case: NaN -> outside
        v[np.any(np.isnan(pts), axis=1)] = np.nan
        return v

    def evaluate_gradient(self, evaluation_points, property_array):
        """Gradient of the trilinear interpolant (:324-391); outside -> 0."""
        self._check_property(property_array)
        c = property_array if isinstance(property_array, torch.Tensor) else \
            torch.from_numpy(np.ascontiguousarray(property_array, dtype=float))
        pts = _as_points(evaluation_points)
        _, g = fs.evaluate(self._lsgrid(), c, torch.from_numpy(pts), False, True)
        return g.cpu().numpy()

    def get_operators(self, weights):
        """(mask, weight) per interior family, cross weights / 4 (:489-510)."""
        return {"dxy": (Operator.Dxy_mask, weights["dxy"] / 4),
                "dyz": (Operator.Dyz_mask, weights["dyz"] / 4),
                "dxz": (Operator.Dxz_mask, weights["dxz"] / 4),
                "dxx": (Operator.Dxx_mask, weights["dxx"] / 1),
                "dyy": (Operator.Dyy_mask, weights["dyy"] / 1),
                "dzz": (Operator.Dzz_mask, weights["dzz"] / 1)}

    def to_dict(self):
        return {"type": int(self.type), **super().to_dict()}

    def vtk(self, node_properties=None, cell_properties=None):
        try:
            import pyvista  # noqa: F401
        except ImportError:
            raise ImportError("pyvista is required for vtk support")
        raise NotImplementedError("vtk export is outside this package's scope")


class RectilinearGrid(StructuredGrid):
    """Structured grid with its own node coordinates per axis
    (LC/supports/_3d_rectilinear_grid.py:20-315).

    Point location (binary search per axis, strict inside test, unclipped
    local coordinates), shape functions and their gradients scaled by each
    cell's size, and evaluation run in the same HIP kernels as the uniform
    grid, reading the node arrays from device memory.  The FD regulariser on
    this support is built as explicit rows scaled by the local spacing
    (``build_scaled_operator_rows``)."""

    def __init__(self, xnodes, ynodes, znodes):
        xnodes = np.asarray(xnodes, dtype=float)
        ynodes = np.asarray(ynodes, dtype=float)
        znodes = np.asarray(znodes, dtype=float)
        if xnodes.ndim != 1 or ynodes.ndim != 1 or znodes.ndim != 1:
            raise ValueError("Node arrays must be 1-D.")
        self._axis = (xnodes, ynodes, znodes)
        self._axis_dev = None
        nsteps = np.array([len(a) - 1 for a in self._axis], dtype=int)
        origin = np.array([a[0] for a in self._axis])
        with np.errstate(divide="ignore", invalid="ignore"):
            extent = np.array([a[-1] - a[0] for a in self._axis])
            step_vector = np.where(nsteps > 0, extent / nsteps, 1.0)
        BaseStructuredSupport.__init__(self, origin=origin, nsteps=nsteps,
                                       step_vector=step_vector)
        self.type = SupportType.RectilinearGrid
        self.name = "RectilinearGrid"
        self.regions = {"everywhere": np.ones(self.n_nodes).astype(bool)}

    @property
    def xnodes(self):
        return self._axis[0]

    @property
    def ynodes(self):
        return self._axis[1]

    @property
    def znodes(self):
        return self._axis[2]

    @property
    def nodes(self):
        xx, yy, zz = np.meshgrid(*self._axis, indexing="ij")
        return np.column_stack([xx.ravel(order="F"), yy.ravel(order="F"), zz.ravel(order="F")])

    @property
    def maximum(self):
        return np.array([a[-1] for a in self._axis])

    def cell_centres(self, global_index):
        ci = self.global_index_to_cell_index(np.asarray(global_index))
        return np.column_stack([0.5 * (a[ci[:, d]] + a[ci[:, d] + 1])
                                for d, a in enumerate(self._axis)])

    def node_indexes_to_position(self, node_indexes):
        ni = np.asarray(node_indexes)
        shape = ni.shape
        ni = ni.reshape(-1, 3)
        xyz = np.column_stack([a[np.clip(ni[:, d], 0, len(a) - 1)]
                               for d, a in enumerate(self._axis)])
        return xyz.reshape(shape)

    def position_to_cell_index(self, pos):
        """(N,3) cell ijk by binary search on each axis, and the strict inside
        flags (:115-130); cells are returned for outside points too."""
        pts = _as_points(pos)
        inside = self.inside(pts)
        ci = np.column_stack([np.clip(np.searchsorted(a, pts[:, d], side="right") - 1, 0,
                                      len(a) - 2) for d, a in enumerate(self._axis)])
        return ci, inside

    def position_to_local_coordinates(self, pos):
        """(p - x0) / h in the cell, not clipped (:145-159)."""
        pts = _as_points(pos)
        ci, _ = self.position_to_cell_index(pts)
        out = np.zeros(pts.shape)
        for d, a in enumerate(self._axis):
            x0 = a[ci[:, d]]
            h = a[np.minimum(ci[:, d] + 1, len(a) - 1)] - x0
            out[:, d] = (pts[:, d] - x0) / np.where(h == 0, 1.0, h)
        return out

    def _lsgrid(self) -> _lib.Grid:
        g = _lib.make_grid(self.origin, self.step_vector, self.nsteps)
        if self._axis_dev is None:
            self._axis_dev = [torch.as_tensor(np.ascontiguousarray(a), device=_lib.device())
                              for a in self._axis]
        for d in range(3):
            g.axis_nodes[d] = self._axis_dev[d].data_ptr()
        return g

    def get_operators(self, weights):
        """No fixed masks: the interpolator builds spacing-scaled rows
        (:217-230)."""
        return {"dxx": (None, weights["dxx"]), "dyy": (None, weights["dyy"]),
                "dzz": (None, weights["dzz"]), "dxy": (None, weights["dxy"] / 4),
                "dyz": (None, weights["dyz"] / 4), "dxz": (None, weights["dxz"] / 4)}

    def build_scaled_operator_rows(self, axis, cross_axis=-1):
        """(coefficients (n,3|4), node ids (n,3|4), centre ids (n,)) of the
        second derivative along ``axis`` (or the mixed one with
        ``cross_axis``) at every node interior along the axes involved
        (:232-315)."""
        A, cols, rows = scaled_operator_rows(self._axis, axis, cross_axis)
        return A.cpu().numpy(), cols.cpu().numpy(), rows.cpu().numpy()

    def to_dict(self):
        return {"type": int(self.type), "xnodes": self._axis[0], "ynodes": self._axis[1],
                "znodes": self._axis[2]}


def scaled_operator_rows(axis_nodes, axis, cross_axis=-1):
    """Spacing-scaled second-derivative rows on a rectilinear grid, built on
    the device: the three-point second derivative
    2[f_(i-1)/(hL (hL+hR)) - f_i/(hL hR) + f_(i+1)/(hR (hL+hR))] at nodes
    interior along ``axis``, or the four-point mixed derivative
    [f(+,+) - f(+,-) - f(-,+) + f(-,-)] / ((h_a) (h_c)) (h over two cells) at
    nodes interior along both axes; centres in ascending node order
    (LC/supports/_3d_rectilinear_grid.py:232-315)."""
    dev = _lib.device()
    ax = [torch.as_tensor(np.ascontiguousarray(a, dtype=float), device=dev) for a in axis_nodes]
    n = [int(a.shape[0]) for a in ax]
    ijk = torch.stack(torch.meshgrid(*[torch.arange(v, device=dev) for v in n], indexing="ij"),
                      dim=-1).permute(2, 1, 0, 3).reshape(-1, 3)
    pure = cross_axis < 0 or cross_axis == axis
    involved = (axis,) if pure else (axis, cross_axis)
    keep = torch.ones(ijk.shape[0], dtype=torch.bool, device=dev)
    for a in involved:
        keep &= (ijk[:, a] > 0) & (ijk[:, a] < n[a] - 1)
    c = ijk[keep]
    stride = (1, n[0], n[0] * n[1])
    rows = c[:, 0] + stride[1] * c[:, 1] + stride[2] * c[:, 2]
    if pure:
        i = c[:, axis]
        hL = ax[axis][i] - ax[axis][i - 1]
        hR = ax[axis][i + 1] - ax[axis][i]
        cm = 2.0 / (hL * (hL + hR))
        cp = 2.0 / (hR * (hL + hR))
        A = torch.stack([cm, -(cm + cp), cp], dim=1)
        cols = torch.stack([rows - stride[axis], rows, rows + stride[axis]], dim=1)
    else:
        ia, ic = c[:, axis], c[:, cross_axis]
        hx = ax[axis][ia + 1] - ax[axis][ia - 1]
        hy = ax[cross_axis][ic + 1] - ax[cross_axis][ic - 1]
        den = hx * hy
        signs = torch.tensor([1.0, -1.0, -1.0, 1.0], dtype=torch.float64, device=dev)
        A = signs[None, :] / den[:, None]
        sa, sc = stride[axis], stride[cross_axis]
        cols = torch.stack([rows + sa + sc, rows + sa - sc, rows - sa + sc, rows - sa - sc],
                           dim=1)
    return A.contiguous(), cols.contiguous(), rows


def _dev_points(pos) -> torch.Tensor:
    return torch.from_numpy(_as_points(pos)).to(_lib.device())


class TetMesh(BaseStructuredSupport):
    """Structured tetrahedral mesh, 5 tets per cell
    (LC/supports/_3d_structured_tetra.py:16-740).

    Topology (elements, neighbours, face table), location, barycentric
    coordinates, element gradients and evaluation come from the P1 HIP
    kernels; arrays the reference returns as numpy are returned as numpy.
    """

    tetra_mask_even = np.array([[7, 1, 2, 4], [6, 2, 4, 7], [5, 1, 4, 7], [0, 1, 2, 4],
                                [3, 1, 2, 7]])
    tetra_mask = np.array([[0, 6, 5, 3], [7, 3, 5, 6], [4, 0, 5, 6], [2, 0, 3, 6], [1, 0, 3, 5]])

    def __init__(self, origin=None, nsteps_cells=None, step_vector=None, nsteps=None):
        if nsteps_cells is None:
            nsteps_cells = np.ones(3) * 10 if nsteps is None else np.array(nsteps)
        super().__init__(origin, np.asarray(nsteps_cells).astype(int), step_vector)
        self.type = SupportType.TetMesh
        self.cg = None
        self._elements = None
        self._faces = None

    def onGeometryChange(self):
        self._elements = None
        self._faces = None
        if self.interpolator is not None:
            self.interpolator.reset()

    # ---------------------------------------------------------- topology --
    @property
    def ntetra(self) -> int:
        return int(np.prod(self.nsteps_cells)) * 5

    @property
    def n_elements(self) -> int:
        return self.ntetra

    @property
    def n_cells(self) -> int:
        return int(np.prod(self.nsteps_cells))

    @property
    def elements(self):
        """(ntetra, 4) node ids (:376-399)."""
        if self._elements is None:
            el, _ = fs.p1_elements(self._lsgrid())
            self._elements = el.cpu().numpy()
        return self._elements

    def get_elements(self):
        return self.elements

    def get_neighbours(self) -> np.ndarray:
        """(ntetra, 4) face neighbours, -1 = none (:626-740)."""
        _, nb = fs.p1_elements(self._lsgrid(), want_neighbours=True)
        return nb.cpu().numpy()

    @property
    def neighbours(self):
        return self.get_neighbours()

    def _face_table(self):
        if self._faces is None:
            pairs, faces = fs.p1_pairs(self._lsgrid(), want_faces=True)
            self._faces = (pairs, faces)
        return self._faces

    @property
    def shared_element_relationships(self):
        """(P,2) face-sharing tet pairs, larger id first (:120-184)."""
        return self._face_table()[0].cpu().numpy()

    @property
    def shared_elements(self):
        """(P,3) node ids of each shared face (ascending)."""
        return self._face_table()[1].cpu().numpy()

    def _node_pos(self, ids: torch.Tensor) -> torch.Tensor:
        """Support node coordinates (np.linspace per axis) of node ids."""
        n = torch.as_tensor(self.nsteps, device=ids.device)
        ijk = torch.stack([ids % n[0], ids // n[0] % n[1], ids // n[0] // n[1]], dim=-1)
        out = []
        for a in range(3):
            o = float(self.origin[a])
            stop = o + float(self.nsteps_cells[a]) * float(self.step_vector[a])
            step = (stop - o) / float(self.nsteps[a] - 1)
            v = ijk[..., a].to(torch.float64) * step + o
            out.append(torch.where(ijk[..., a] == self.nsteps[a] - 1,
                                   torch.full_like(v, stop), v))
        return torch.stack(out, dim=-1)

    @property
    def shared_element_norm(self):
        """(v1-v0) x (v2-v0) of each shared face (:186-194)."""
        f = self._face_table()[1]
        p = self._node_pos(f)
        return torch.linalg.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]).cpu().numpy()

    @property
    def shared_element_size(self):
        return 0.5 * np.linalg.norm(self.shared_element_norm, axis=1)

    @property
    def shared_element_scale(self):
        s = self.shared_element_size
        return s / np.mean(s)

    @property
    def element_size(self):
        """|det| / 6 of every tet (:80-94)."""
        t = torch.arange(self.ntetra, dtype=torch.int64, device=_lib.device())
        _, s = fs.p1_tet_gradients(self._lsgrid(), t, want_size=True)
        return s.cpu().numpy()

    @property
    def element_scale(self):
        size = self.element_size
        size -= np.min(size)
        size /= np.max(size)
        size += 1.0
        return size

    @property
    def barycentre(self):
        el, _ = fs.p1_elements(self._lsgrid())
        return (self._node_pos(el).sum(dim=1) / 4.0).cpu().numpy()

    def tetra_global_index(self, indices, tetra_index):
        """:401-419."""
        indices = np.asarray(indices)
        nc = self.nsteps_cells
        return (tetra_index + indices[:, 0] * 5 + nc[0] * indices[:, 1] * 5
                + nc[0] * nc[1] * indices[:, 2] * 5)

    # ---------------------------------------------------------- location --
    def inside(self, pos):
        """Strict box test (:264-272)."""
        p = _dev_points(pos)
        o = torch.as_tensor(self.origin, device=p.device)
        mx = o + torch.as_tensor(self.step_vector * self.nsteps_cells, device=p.device)
        return torch.all((p > o) & (p < mx), dim=1).cpu().numpy()

    def position_to_cell_index(self, pos):
        """Cell ijk of strictly inside points (0 elsewhere) and the flags."""
        p = _dev_points(pos)
        o = torch.as_tensor(self.origin, device=p.device)
        h = torch.as_tensor(self.step_vector, device=p.device)
        nc = torch.as_tensor(self.nsteps_cells, device=p.device)
        mx = o + h * nc
        ins = torch.all((p > o) & (p < mx), dim=1)
        ci = torch.floor((p - o) / h)
        ci = torch.where(ins[:, None], ci, torch.zeros_like(ci)).to(torch.int64)
        ci = torch.minimum(torch.clamp(ci, min=0), nc[None, :] - 1)
        return ci.cpu().numpy(), ins.cpu().numpy()

    def _p1_locate(self, pos, want_grad=False):
        return fs.p1_locate(self._lsgrid(), _dev_points(pos), True, want_grad)

    def get_element_for_location(self, pos):
        """(vertices (N,4,3), bc (N,4), tet (N,), inside (N,)); NaN / -1
        outside (:274-365)."""
        tet, nodes, ins, bc, _ = self._p1_locate(pos)
        nd = nodes.clamp(min=0)
        ijk = torch.stack([nd % int(self.nsteps[0]), nd // int(self.nsteps[0]) % int(self.nsteps[1]),
                           nd // int(self.nsteps[0]) // int(self.nsteps[1])], dim=-1)
        o = torch.as_tensor(self.origin, device=nd.device)
        h = torch.as_tensor(self.step_vector, device=nd.device)
        vert = o + h * ijk.to(torch.float64)
        vert[~ins] = float("nan")
        return vert.cpu().numpy(), bc.cpu().numpy(), tet.cpu().numpy(), ins.cpu().numpy()

    def evaluate_shape(self, locations):
        """(bc, tet, inside) (:367-374)."""
        tet, _, ins, bc, _ = self._p1_locate(locations)
        return bc.cpu().numpy(), tet.cpu().numpy(), ins.cpu().numpy()

    def get_element_gradients(self, elements=None):
        """(n,3,4) gradients of the barycentric coordinates (:421-486)."""
        if elements is None:
            elements = np.arange(0, self.ntetra)
        t = torch.as_tensor(np.asarray(elements, dtype=np.int64), device=_lib.device())
        g, _ = fs.p1_tet_gradients(self._lsgrid(), t.reshape(-1))
        return g.reshape(*t.shape, 3, 4).cpu().numpy()

    def evaluate_shape_derivatives(self, pos, elements=None):
        """(gradients, elements, inside) (:488-500)."""
        if elements is not None:
            inside = np.ones(np.asarray(elements).shape[0], dtype=bool)
        else:
            _, _, elements, inside = self.get_element_for_location(pos)
        return self.get_element_gradients(elements), elements, inside

    def get_element_gradient_for_location(self, pos):
        """(vertices, gradients (N,3,4) NaN outside, tet, inside) (:502-545)."""
        vert, _, tet, ins = self.get_element_for_location(pos)
        _, _, _, _, g = self._p1_locate(pos, want_grad=True)
        return vert, g.cpu().numpy(), tet, ins

    # -------------------------------------------------------- evaluation --
    def _check_property(self, property_array):
        if np.asarray(property_array).shape[0] != self.n_nodes:
            raise ValueError(f"cannot assign {np.asarray(property_array).shape[0]} vlaues "
                             f"to array of shape {self.n_nodes}")

    def _prop(self, property_array):
        if isinstance(property_array, torch.Tensor):
            return property_array
        return torch.from_numpy(np.ascontiguousarray(property_array, dtype=float))

    def evaluate_value(self, pos, property_array):
        """Barycentric interpolation (:208-229); NaN outside (strict)."""
        v, _ = fs.p1_evaluate(self._lsgrid(), self._prop(property_array), _dev_points(pos),
                              True, False)
        return v.cpu().numpy()

    def evaluate_gradient(self, pos, property_array):
        """Element gradient (:231-262); NaN outside."""
        _, g = fs.p1_evaluate(self._lsgrid(), self._prop(property_array), _dev_points(pos),
                              False, True)
        return g.cpu().numpy()

    def to_dict(self):
        return {"type": int(self.type), **super().to_dict()}

    def vtk(self, node_properties=None, cell_properties=None):
        try:
            import pyvista  # noqa: F401
        except ImportError:
            raise ImportError("pyvista is required for vtk support")
        raise NotImplementedError("vtk export is outside this package's scope")
