"""Drop-in interpolators: ``FiniteDifferenceInterpolator`` on the MI355X.

Same class names, method names, keyword arguments, attributes and error
behaviour as the reference (LI/_geological_interpolator.py,
LI/_discrete_interpolator.py, LI/_finite_difference_interpolator.py); the
numerical work (point location, row assembly and normalisation, the
matrix-free regularised normal operator, the PCG solve and evaluation) runs in
the HIP kernels of libloopstructural_amd.  Constraint rows live in HBM; the
regulariser is never materialised by the solvers (``build_matrix``,
``constraints[name]['matrix']`` and callable solvers get an explicit copy on
request).
"""
from __future__ import annotations

import ctypes
import itertools
import json
import logging
from collections import OrderedDict, defaultdict
from enum import Enum
from time import perf_counter

import numpy as np
import torch
from scipy import sparse

from . import _lib
from . import constraints as cons
from . import explicit
from . import fd_system as fs
from .diagnostics import (ConstraintDiagnosticsReport, ConstraintFamilyDiagnostics,
                          RegionCoverageDiagnostics)
from .supports import Operator

logger = logging.getLogger(__name__)

ValidationError = cons.ValidationError



# pair generators of the interface constraints, by mode: (n_pairs, 2) indices
_PAIR_MODES = {
    "all": lambda n: np.column_stack(np.triu_indices(n, k=1)),
    "star": lambda n: np.column_stack([np.zeros(n - 1, dtype=int), np.arange(1, n)]),
    "chain": lambda n: np.column_stack([np.arange(n - 1), np.arange(1, n)]),
}

class InterpolatorType(Enum):
    """LI/_interpolatortype.py."""

    BASE = "BASE"
    BASE_DISCRETE = "BASE_DISCRETE"
    FINITE_DIFFERENCE = "FINITE_DIFFERENCE"
    DISCRETE_FOLD = "DISCRETE_FOLD"
    PIECEWISE_LINEAR = "PIECEWISE_LINEAR"
    PIECEWISE_QUADRATIC = "PIECEWISE_QUADRATIC"
    BASE_DATA_SUPPORTED = "BASE_DATA_SUPPORTED"
    SURFE = "SURFE"
    PIECEWISE_LINEAR_CONSTANT_NORM = "PIECEWISE_LINEAR_CONSTANT_NORM"
    FINITE_DIFFERENCE_CONSTANT_NORM = "FINITE_DIFFERENCE_CONSTANT_NORM"


class _Lazy:
    """A ``constraints`` entry computed when it is first read."""

    def __init__(self, fn):
        self.fn = fn


class _Constraint(dict):
    """``constraints[name]`` entry: {"matrix", "b", "w"} as in
    LI/_discrete_interpolator.py:383-389.  Device tensors are returned as
    numpy arrays.  Row blocks held on the device and matrix-free families
    return ``matrix`` as a scipy CSC matrix of their normalised rows, built on
    first access and kept."""

    def __getitem__(self, key):
        v = dict.__getitem__(self, key)
        if key == "matrix" and hasattr(v, "to_scipy"):
            m = getattr(self, "_csc", None)
            if m is None:
                m = self._csc = v.to_scipy().tocsc()
            return m
        if isinstance(v, _Lazy):
            v = v.fn()
        if isinstance(v, torch.Tensor):
            return v.detach().cpu().numpy()
        return v

    def raw(self, key):
        """The stored object (device block, stencil family) behind ``key``."""
        return dict.__getitem__(self, key)

    def get(self, key, default=None):
        return self[key] if key in self else default


class StencilFamily:
    """A regulariser family kept matrix-free: rows of one normalised mask at a
    set of centre nodes, the interior nodes or one border layer
    (``layer`` = (axis, upper))."""

    def __init__(self, name, kind, weight, nrows, dof, mask=None, nnodes=None, layer=None):
        self.name, self.kind, self.weight, self.nrows, self.dof = name, kind, weight, nrows, dof
        self.mask, self.nnodes, self.layer = mask, nnodes, layer

    @property
    def shape(self):
        return (self.nrows, self.dof)

    def centres(self):
        if self.layer is None:
            return explicit.interior_centres(self.nnodes)
        return explicit.layer_centres(self.nnodes, *self.layer)

    def to_scipy(self):
        return explicit.stencil_matrix(self.nnodes, self.mask, self.centres())


def _mask_kind(mask):
    """Map a (3,3,3) mask to the canonical family it equals after row
    normalisation (LI/_discrete_interpolator.py:361-366), else None."""
    m = np.asarray(mask, dtype=float)
    if m.shape != (3, 3, 3):
        return None
    nm = np.linalg.norm(m[m != 0]) if np.any(m != 0) else 0.0
    if nm == 0:
        return None
    for kind, ref in (("xx", Operator.Dxx_mask), ("yy", Operator.Dyy_mask),
                      ("zz", Operator.Dzz_mask), ("xy", Operator.Dxy_mask),
                      ("xz", Operator.Dxz_mask), ("yz", Operator.Dyz_mask)):
        r = np.asarray(ref, dtype=float)
        r = r / np.linalg.norm(r[r != 0])
        if np.allclose(m / nm, r, atol=1e-12) or np.allclose(m / nm, -r, atol=1e-12):
            return kind
    return None


_ADMM_OPTIONS = ("adaptive_rho", "adaptive_rho_mu", "adaptive_rho_tau", "adaptive_rho_min",
                 "adaptive_rho_max", "admm_weight_final", "admm_weight_schedule", "admm_abs_tol",
                 "admm_rel_tol", "min_iterations", "active_set", "active_set_padding",
                 "active_set_min_size", "active_set_max_size", "active_set_hysteresis",
                 "active_set_refresh_interval", "model_update_tol", "return_history")


class _AdmmOps:
    """ADMM x-updates on the GPU system of an interpolator.

    The inequality rows act on the scaled unknown ``y = S^-1 c`` (S the
    column scaling of the ridge-augmented equality system, or I), so in node
    space they are ``Q S^-1``.  They join the system's data rows with a
    per-row weight; the x-update is then the system's own least-squares
    problem with right-hand side ``v`` on those rows, solved by PCG from the
    previous iterate to a relative tolerance ``step_rtol``."""

    def __init__(self, interp, system, Q, multigrid, step_rtol):
        dev = _lib.device()
        self.system = system
        self.multigrid = multigrid
        self.step_rtol = step_rtol
        self.iterations = 0
        self.base = list(interp._rows.values())
        if interp.apply_scaling_matrix:
            d = system.diag()
            self.S = torch.where(d > 0, 1.0 / torch.sqrt(d), torch.ones_like(d))
        else:
            self.S = torch.ones(interp.dof, dtype=torch.float64, device=dev)
        Q = sparse.csr_matrix(Q)
        Q.sum_duplicates()
        n_ie = Q.shape[0]
        counts = np.diff(Q.indptr)
        k = max(int(counts.max()) if n_ie else 1, 1)
        cols = np.zeros((n_ie, k), dtype=np.int64)
        coef = np.zeros((n_ie, k))
        pos = np.arange(Q.nnz) - np.repeat(Q.indptr[:-1], counts)
        rows = np.repeat(np.arange(n_ie), counts)
        cols[rows, pos] = Q.indices
        coef[rows, pos] = Q.data
        self.qcols = torch.as_tensor(cols, device=dev)
        Sinv = 1.0 / self.S[self.qcols]
        self.qcoef = torch.as_tensor(coef, device=dev) * Sinv  # Q S^-1
        self.n_ie, self.dof = n_ie, interp.dof
        self.base_b = torch.cat([blk.b for blk in self.base]) if self.base else \
            torch.zeros(0, dtype=torch.float64, device=dev)
        norm = torch.ones(n_ie, dtype=torch.float64, device=dev)
        self._qblk = fs.fixed_width_rows(self.qcols.to(torch.int32).contiguous(),
                                         self.qcoef.contiguous(),
                                         torch.zeros(n_ie, dtype=torch.float64, device=dev),
                                         torch.zeros(n_ie, dtype=torch.float64, device=dev),
                                         norm, interp.dof)

    def q(self, c):
        return (self.qcoef * c[self.qcols]).sum(dim=1)

    def set_weights(self, w):
        self._qblk.w = w.contiguous()
        self.system.set_rows(fs.concat_rows(self.base + [self._qblk], self.dof))
        d = self.system.diag()
        self.Sn = torch.where(d > 0, 1.0 / torch.sqrt(d), torch.ones_like(d))

    def step(self, v, c):
        rhs = self.system.rhs_for(torch.cat([self.base_b, v]))
        r = rhs - self.system.apply(c)
        thr = self.step_rtol * float(torch.linalg.vector_norm(self.Sn * rhs))
        if float(torch.linalg.vector_norm(self.Sn * r)) <= thr:
            return c
        dc, info = self.system.solve(rtol=0.0, atol=thr, rhs=r, multigrid=self.multigrid,
                                     check_every=8 if self.multigrid else 64)
        self.iterations += info["iterations"]
        return c + dc

    def y_norm(self, x):
        return float(torch.linalg.vector_norm(x / self.S))


class GeologicalInterpolator:
    """LI/_geological_interpolator.py:40-897 (data handling, projection,
    evaluation entry points, serialisation)."""

    def __init__(self, data=None, up_to_date=False):
        self._data = {}
        self.data = data
        self.clean()
        self.n_g = self.n_i = self.n_n = self.n_t = 0
        self.type = InterpolatorType.BASE
        self.up_to_date = up_to_date
        self.constraints = []
        self.valid = True
        self.dimensions = 3
        self.support = None
        self.bounding_box = None
        self.latest_diagnostics_report = None

    @property
    def data(self):
        return self._data

    @data.setter
    def data(self, data):
        if data is None:
            data = {}
        for k, v in data.items():
            self._data[k] = np.array(v)

    def clean(self):
        """LI/_geological_interpolator.py:864-885."""
        self.data = {"gradient": np.zeros((0, 7)), "value": np.zeros((0, 5)),
                     "normal": np.zeros((0, 7)), "tangent": np.zeros((0, 7)),
                     "interface": np.zeros((0, 5)), "inequality": np.zeros((0, 6)),
                     "inequality_pairs": np.zeros((0, 4))}
        self.up_to_date = False
        self.n_g = self.n_i = self.n_n = self.n_t = 0

    def __str__(self):
        return (f"{self.type} \n{self.n_g} gradient points\n{self.n_i} interface points\n"
                f"{self.n_n} normal points\n{self.n_t} tangent points\n"
                f"{self.n_g + self.n_i + self.n_n + self.n_t} total points\n")

    # -------------------------------------------------------- projection --
    def _project_points(self, points):
        return points if self.bounding_box is None else self.bounding_box.project(points)

    def _reproject_vectors(self, vectors):
        if self.bounding_box is None:
            return vectors
        return self.bounding_box.reproject_vectors(vectors)

    def _project_constraint_array(self, array, has_vector=False):
        """LI/_geological_interpolator.py:273-286."""
        if self.bounding_box is None or array.shape[0] == 0:
            return array
        d = self.dimensions
        array = array.copy()
        array[:, :d] = self.bounding_box.project(array[:, :d])
        if has_vector:
            array[:, d:2 * d] = self.bounding_box.project_vectors(array[:, d:2 * d])
        return array

    # ------------------------------------------------------------ set_* --
    def set_value_constraints(self, points):
        """(N,4|5) [x,y,z,value(,w)] (LI/_geological_interpolator.py:303-333)."""
        try:
            pts = self._project_constraint_array(cons.value_array(points, self.dimensions))
            self.data["value"] = pts.copy()
            self.n_i = pts.shape[0]
            self.up_to_date = False
        except ValidationError as e:
            raise type(e)(f"Failed to set value constraints: {e}") from e

    def set_gradient_constraints(self, points):
        """(N,6|7) [x,y,z,gx,gy,gz(,w)] (:335-367)."""
        try:
            pts = self._project_constraint_array(cons.gradient_array(points, self.dimensions),
                                                 has_vector=True)
            self.n_g = pts.shape[0]
            self.data["gradient"] = pts.copy()
            self.up_to_date = False
        except ValidationError as e:
            raise type(e)(f"Failed to set gradient constraints: {e}") from e

    def set_normal_constraints(self, points):
        """(N,6|7) [x,y,z,nx,ny,nz(,w)] (:369-400)."""
        try:
            pts = self._project_constraint_array(
                cons.gradient_array(points, self.dimensions, is_normal=True), has_vector=True)
            self.n_n = pts.shape[0]
            self.data["normal"] = pts.copy()
            self.up_to_date = False
        except ValidationError as e:
            raise type(e)(f"Failed to set normal constraints: {e}") from e

    def set_tangent_constraints(self, points):
        """(N,6|7) [x,y,z,tx,ty,tz(,w)] (:402-433)."""
        try:
            pts = self._project_constraint_array(cons.gradient_array(points, self.dimensions),
                                                 has_vector=True)
            self.n_t = pts.shape[0]
            self.data["tangent"] = pts.copy()
            self.up_to_date = False
        except ValidationError as e:
            raise type(e)(f"Failed to set tangent constraints: {e}") from e

    def set_interface_constraints(self, points):
        """(N,4|5) [x,y,z,id(,w)] (:435-462)."""
        try:
            pts = self._project_constraint_array(cons.interface_array(points, self.dimensions))
            self.data["interface"] = pts.copy()
            self.up_to_date = False
        except ValidationError as e:
            raise type(e)(f"Failed to set interface constraints: {e}") from e

    def set_value_inequality_constraints(self, points):
        """(N,5|6) [x,y,z,lower,upper(,w)], lower < upper (:464-493)."""
        try:
            pts = self._project_constraint_array(cons.inequality_array(points, self.dimensions))
            self.data["inequality"] = pts.copy()
            self.up_to_date = False
        except ValidationError as e:
            raise type(e)(f"Failed to set inequality value constraints: {e}") from e

    def set_inequality_pairs_constraints(self, points):
        """(N,4|5) [x,y,z,rock id(,w)] (:495-522)."""
        try:
            pts = self._project_constraint_array(
                cons.inequality_pair_array(points, self.dimensions))
            self.data["inequality_pairs"] = pts.copy()
            self.up_to_date = False
        except ValidationError as e:
            raise type(e)(f"Failed to set inequality pairs constraints: {e}") from e

    def get_value_constraints(self):
        return self.data["value"]

    def get_gradient_constraints(self):
        return self.data["gradient"]

    def get_tangent_constraints(self):
        return self.data["tangent"]

    def get_norm_constraints(self):
        return self.data["normal"]

    def get_interface_constraints(self):
        return self.data["interface"]

    def get_inequality_value_constraints(self):
        return self.data["inequality"]

    def get_inequality_pairs_constraints(self):
        return self.data["inequality_pairs"]

    def get_data_locations(self):
        return np.vstack([d[:, :3] for d in self.data.values()])

    # ------------------------------------------------------ diagnostics --
    def _outside_model_points_from_data(self):
        if self.support is None:
            return {}
        mapping = {"value": "value", "gradient": "gradient", "normal": "normal",
                   "tangent": "tangent", "interface": "interface",
                   "inequality": "inequality_value", "inequality_pairs": "inequality_pairs"}
        out = {}
        for key, fam in mapping.items():
            pts = self.data.get(key)
            if pts is None or pts.shape[0] == 0:
                out[fam] = 0
                continue
            out[fam] = int((~self.support.inside(pts[:, :3])).sum())
        return out

    def get_constraint_diagnostics_report(self, refresh=False):
        if self.latest_diagnostics_report is None or refresh:
            self.latest_diagnostics_report = self._build_constraint_diagnostics_report()
        return self.latest_diagnostics_report

    def setup(self, **kwargs):
        """Run setup and return a diagnostics report (:698-704)."""
        report = self.setup_interpolator(**kwargs)
        if isinstance(report, ConstraintDiagnosticsReport):
            self.latest_diagnostics_report = report
            return report
        return self.get_constraint_diagnostics_report(refresh=True)

    # --------------------------------------------------------- evaluate --
    def evaluate_value(self, locations):
        """World-coordinate evaluation (:723-730)."""
        return self._evaluate_value_local(self._project_points(locations))

    def evaluate_gradient(self, locations):
        """World-coordinate gradient (:732-741)."""
        return self._reproject_vectors(
            self._evaluate_gradient_local(self._project_points(locations)))

    # -------------------------------------------------------- serialise --
    def to_dict(self):
        """LI/_geological_interpolator.py:799-835."""
        def safe(v):
            if isinstance(v, np.ndarray):
                return v.tolist()
            if isinstance(v, (np.floating, np.integer)):
                return v.item()
            if isinstance(v, dict):
                return {k: safe(x) for k, x in v.items()}
            if isinstance(v, (list, tuple)):
                return [safe(x) for x in v]
            return v

        payload = {"type": self.type.value,
                   "data": safe({k: np.asarray(v) for k, v in self.data.items()}),
                   "up_to_date": self.up_to_date, "valid": self.valid}
        if self.support is not None and hasattr(self.support, "to_dict"):
            sd = safe(self.support.to_dict())
            if isinstance(sd, dict) and "nsteps" in sd and hasattr(self.support, "nsteps_cells"):
                sd["nsteps"] = safe(np.asarray(self.support.nsteps_cells))
            payload["support"] = sd
        return payload

    def to_json(self, indent: int = 2) -> str:
        return json.dumps(self.to_dict(), indent=indent)

    def to_yaml(self, file_path=None):
        import yaml

        if file_path is None:
            return yaml.safe_dump(self.to_dict(), sort_keys=False, allow_unicode=True)
        with open(file_path, "w") as f:
            yaml.safe_dump(self.to_dict(), f, sort_keys=False, allow_unicode=True)


class DiscreteInterpolator(GeologicalInterpolator):
    """Least-squares interpolator on a discrete support
    (LI/_discrete_interpolator.py:29-1405)."""

    def __init__(self, support, data=None, c=None, up_to_date=False):
        GeologicalInterpolator.__init__(self, data=data, up_to_date=up_to_date)
        self.support = support
        self.dimensions = support.dimension
        self.c = (np.array(c) if c is not None and np.array(c).shape[0] == support.n_nodes
                  else np.zeros(support.n_nodes))
        self.shape = "rectangular"
        self.c_ = 0
        self.solver = "cg"
        self.constraints = OrderedDict()
        self.ineq_constraints = {}
        self.interpolation_weights = {}
        self.type = InterpolatorType.BASE_DISCRETE
        self.apply_scaling_matrix = True
        self.add_ridge_regulatisation = True
        self.ridge_factor = 1e-8
        self.solver_history = None
        self.latest_solve_timing = {}
        self._rows = OrderedDict()       # name -> fs.FdRows (device)
        self._stencils = OrderedDict()   # name -> StencilFamily
        self._system = None
        self._system_key = None
        self._scale_sources = []         # pending regularisation_scale terms
        self._regularisation_scale = None
        if hasattr(support, "associateInterpolator"):
            support.associateInterpolator(self)

    # ----------------------------------------------------------- basics --
    def get_last_solve_timing(self) -> dict:
        return dict(self.latest_solve_timing)

    def set_nelements(self, nelements: int) -> int:
        return self.support.set_nelements(nelements)

    @property
    def n_elements(self) -> int:
        return self.support.n_elements

    @property
    def dof(self) -> int:
        return int(self.support.n_nodes)

    @property
    def region(self):
        return np.ones(self.support.n_nodes, dtype=bool)

    @property
    def region_map(self):
        return np.arange(self.support.n_nodes, dtype=int)

    def set_region(self, region=None):
        logger.info("Region masking has been removed; the interpolator always uses all "
                    "support nodes (%s DOF).", self.dof)

    def set_interpolation_weights(self, weights):
        for key in weights:
            self.up_to_date = False
            self.interpolation_weights[key] = weights[key]

    def _apply_isotropic_regularisation_weight(self, value, keys):
        if value is None:
            return
        for key in keys:
            self.interpolation_weights[key] = value

    def _apply_interpolation_weight_kwargs(self, kwargs, skip_keys=()):
        for key, value in kwargs.items():
            self.up_to_date = False
            if key in skip_keys:
                continue
            self.interpolation_weights[key] = value

    def reset(self):
        """LI/_discrete_interpolator.py:306-314."""
        self.constraints = OrderedDict()
        self._rows = OrderedDict()
        self._stencils = OrderedDict()
        self.c_ = 0
        self._scale_sources = []
        self._regularisation_scale = None

    @property
    def regularisation_scale(self):
        """ones(dof) + one ``1 + 10 exp(-d^2 / 2 sigma^2)`` term per normal
        family added (LI/_finite_difference_interpolator.py:436-444), d the
        distance to the nearest normal constraint; computed on the GPU on
        first use."""
        if self._regularisation_scale is None:
            s = torch.ones(self.dof, dtype=torch.float64, device=_lib.device())
            for pts, sigma in self._scale_sources:
                s += _regularisation_weighting(self.support, pts, sigma)
            self._regularisation_scale = s
        return self._regularisation_scale.cpu().numpy()

    @regularisation_scale.setter
    def regularisation_scale(self, value):
        self._regularisation_scale = torch.as_tensor(np.asarray(value, dtype=float),
                                                     device=_lib.device())
        self._scale_sources = []

    def _pre_solve(self):
        self.c = np.full(self.support.n_nodes, np.nan)
        return True

    def clear_constraints(self):
        self.constraints = OrderedDict()
        self._rows = OrderedDict()
        self._stencils = OrderedDict()
        self.ineq_constraints = {}

    # ------------------------------------------------------------- rows --
    def _unique_name(self, name):
        """Name under which a new family is stored: ``name`` itself, or the
        first free of ``name_k, name_(k+1), ...`` with k = 0, or one past the
        number after the first underscore of ``name`` (the rename of
        LI/_discrete_interpolator.py:375-380; where that loop would revisit a
        taken name, this takes the next number)."""
        if name not in self.constraints:
            return name
        parts = name.split("_")
        start = int(parts[1]) + 1 if len(parts) > 1 and parts[1].isdigit() else 0
        return next(c for c in (f"{name}_{k}" for k in itertools.count(start))
                    if c not in self.constraints)

    def _register_rows(self, blk: fs.FdRows, name: str):
        if blk.nrows == 0:
            name = self._unique_name(name)
            self._rows[name] = blk
            self.constraints[name] = _Constraint(matrix=blk, b=blk.b, w=blk.w)
            return name
        if bool(torch.isnan(blk.coef).any()) or bool(torch.isnan(blk.b).any()):
            logger.warning(f"Constraints contain nan not adding constraints: {name}")
            return None
        name = self._unique_name(name)
        self._rows[name] = blk
        self.constraints[name] = _Constraint(matrix=blk, b=blk.b, w=blk.w)
        return name

    def add_constraints_to_least_squares(self, A, B, idc, w=1.0, name="undefined"):
        """Generic rows (LI/_discrete_interpolator.py:316-389), normalised on the GPU."""
        A = np.array(A, dtype=float)
        B = np.array(B, dtype=float)
        idc = np.array(idc)
        if A.shape != idc.shape:
            logger.error(f"Cannot add constraints: A and indexes have different shape : {name}")
            return
        if A.ndim > 2:
            if isinstance(w, np.ndarray):
                w = np.tile(w, (A.shape[1]))
            A = A.reshape((A.shape[0] * A.shape[1], A.shape[2]))
            idc = idc.reshape((idc.shape[0] * idc.shape[1], idc.shape[2]))
            B = B.reshape(A.shape[0])
        if np.any(np.isnan(idc)) or np.any(np.isnan(A)) or np.any(np.isnan(B)):
            logger.warning(f"Constraints contain nan not adding constraints: {name}")
            return
        if isinstance(w, (float, int)):
            w = np.ones(A.shape[0]) * w
        if not isinstance(w, np.ndarray):
            raise TypeError("w must be a numpy array")
        if w.shape[0] != A.shape[0]:
            raise ValueError("Weight array does not match number of constraints")
        dev = _lib.device()
        R, k = A.shape
        coef = torch.from_numpy(np.ascontiguousarray(A)).to(dev)
        b = torch.from_numpy(np.ascontiguousarray(B)).to(dev)
        norm = torch.empty(R, dtype=torch.float64, device=dev)
        if R:
            _lib.call("lsamd_rows_normalise", R, k, _lib.ptr(coef), _lib.ptr(b), _lib.ptr(norm),
                      _lib.stream())
        cols = torch.from_numpy(np.ascontiguousarray(idc.astype(np.int32))).to(dev)
        blk = fs.fixed_width_rows(cols, coef, b,
                                  torch.from_numpy(np.ascontiguousarray(w, dtype=float)).to(dev),
                                  norm, self.dof)
        return self._register_rows(blk, name)

    # ------------------------------------------------------- inequalities --
    def add_inequality_constraints_to_matrix(self, A, bounds, idc, name="undefined"):
        """Rows ``l <= A c <= u`` over the node ids ``idc``, bounds (n, 2|3)
        (LI/_discrete_interpolator.py:467-507).  Kept on the host as a CSC
        block: these rows are few next to the grid, and the ADMM solver
        uploads them once per solve."""
        A = np.asarray(A, dtype=float)
        idc = np.asarray(idc, dtype=int)
        bounds = np.asarray(bounds, dtype=float)
        if A.ndim != 2 or idc.ndim != 2:
            raise ValueError("A and idc must be 2D arrays")
        if A.shape != idc.shape:
            raise ValueError("A and idc must have the same shape")
        if bounds.ndim != 2 or bounds.shape[0] != A.shape[0] or bounds.shape[1] not in (2, 3):
            raise ValueError("bounds must have shape (n_rows, 2) or (n_rows, 3)")
        if np.any(idc < 0) or np.any(idc >= self.dof):
            raise ValueError("Inequality constraint indices are out of range")
        rows = np.repeat(np.arange(idc.shape[0]), A.shape[1])
        self.ineq_constraints[name] = {
            "matrix": sparse.coo_matrix((A.ravel(), (rows, idc.ravel())),
                                        shape=(idc.shape[0], self.dof)).tocsc(),
            "bounds": bounds}

    def _element_rows(self, coords):
        """Shape-function values and node ids of the element holding each
        point, and the inside flags (get_element_for_location + elements)."""
        _, a, element, inside = self.support.get_element_for_location(coords)
        return a[inside], self._element_nodes(element[inside]), inside

    def _element_nodes(self, element):
        return self.support.elements[element]

    def add_value_inequality_constraints(self, w: float = 1.0):
        """``lower <= f(p) <= upper`` at each inequality point inside the
        support (LI/_discrete_interpolator.py:509-518)."""
        points = self.get_inequality_value_constraints()
        if points.shape[0] > 0:
            a, cols, inside = self._element_rows(points[:, :self.support.dimension])
            d = self.support.dimension
            self.add_inequality_constraints_to_matrix(a, points[inside, d:d + 2], cols,
                                                      "inequality_value")

    def add_inequality_pairs_constraints(self, w: float = 1.0, upper_bound=1.0,
                                         lower_bound=-np.inf, pairs=None):
        """``lower_bound <= f(p_i) - f(q_j) <= upper_bound`` for every point
        p_i of one rock id and q_j of another, per pair of ids (all ordered id
        pairs seen when ``pairs`` is not given)
        (LI/_discrete_interpolator.py:520-593)."""
        points = self.get_inequality_pairs_constraints()
        if points.shape[0] == 0:
            return
        d = self.support.dimension
        if not pairs:
            ids = np.unique(points[:, d])
            pairs = list(dict.fromkeys(tuple(sorted((i, j))) for i in ids for j in ids if i != j))
        for pair in pairs:
            up = points[points[:, d] == pair[0]]
            lo = points[points[:, d] == pair[1]]
            a_up, c_up, in_up = self._element_rows(up[:, :d])
            a_lo, c_lo, in_lo = self._element_rows(lo[:, :d])
            if (~in_up).sum() > 0:
                logger.warning(f"Upper points not in mesh {up[~in_up]}")
            if (~in_lo).sum() > 0:
                logger.warning(f"Lower points not in mesh {lo[~in_lo]}")
            # every (upper i, lower j), j major: the order of meshgrid(i, j)
            jj, ii = np.divmod(np.arange(a_up.shape[0] * a_lo.shape[0]), a_up.shape[0])
            a = np.hstack([a_up[ii], -a_lo[jj]])
            cols = np.hstack([c_up[ii], c_lo[jj]])
            bounds = np.zeros((a.shape[0], 2))
            bounds[:, 0] = lower_bound
            bounds[:, 1] = upper_bound
            self.add_inequality_constraints_to_matrix(
                a, bounds, cols, f"inequality_pairs_{pair[0]}_{pair[1]}")

    def add_inequality_feature(self, feature, lower=True, mask=None):
        """Bound every node by a feature evaluated there: from below
        (``lower``) or above, on the nodes of ``mask`` where the feature is
        defined (LI/_discrete_interpolator.py:595-636)."""
        value = feature(self.support.nodes)
        if mask is None:
            mask = np.ones(value.shape[0], dtype=bool)
        lo = np.full(value.shape[0], -np.inf)
        up = np.full(value.shape[0], np.inf)
        mask = np.logical_and(mask, ~np.isnan(value))
        if lower:
            lo[mask] = value[mask]
        else:
            up[mask] = value[mask]
        self.add_inequality_constraints_to_matrix(
            np.ones((value.shape[0], 1)), np.column_stack([lo, up]),
            np.arange(0, self.dof, dtype=int)[:, None], name="inequality_feature")

    def build_inequality_matrix(self):
        """(Q, bounds) stacked over the inequality families
        (LI/_discrete_interpolator.py:874-884)."""
        mats = [c["matrix"] for c in self.ineq_constraints.values()]
        if not mats:
            return sparse.csr_matrix((0, self.dof), dtype=float), np.zeros((0, 3))
        return sparse.vstack(mats), np.vstack([c["bounds"] for c in self.ineq_constraints.values()])

    # ------------------------------------------------------------ solve --
    def _normalise_solver_choice(self, solver):
        """LI/_solver_strategy.py:15-29."""
        if callable(solver):
            return solver
        if isinstance(solver, str) or solver is None:
            if solver not in ["cg", "lsmr", "admm", None]:
                logger.warning("Unknown solver %s using cg. Available solvers are cg, lsmr, "
                               "admm or a custom callable", solver)
                return "cg"
            return "cg" if solver is None else solver
        logger.warning("Unsupported solver type %s, using cg", type(solver).__name__)
        return "cg"

    def _system_rows(self):
        return fs.concat_rows(list(self._rows.values()), self.dof)

    def build_matrix(self):
        """The assembled system ``A = [w * a_hat]``, ``b = [w * b_hat]`` over
        the families in insertion order, as a scipy CSR matrix on the host
        (LI/_discrete_interpolator.py:684-755).  The GPU solvers do not use
        it; it is for callers that inspect the system or solve it themselves."""
        A, B = explicit.stack(((c["matrix"], c["b"], c["w"]) for c in self.constraints.values()),
                              self.dof)
        logger.info(f"Interpolation matrix is {A.shape[0]} x {A.shape[1]}")
        return A, B

    def _solve_callable(self, solver, timing, started) -> bool:
        """solve_with_callable (LI/_solver_strategy.py:32-45) on the explicit
        preprocessed system: ridge rows and column scaling
        (LI/_solver_pipeline.py:86-124), ``y = solver(A S, b)``, ``c = S y``
        (LI/_discrete_interpolator.py:1288-1313)."""
        t0 = perf_counter()
        A, b = self.build_matrix()
        timing["assembly_seconds"] = perf_counter() - t0
        timing["matrix_rows"] = int(A.shape[0])
        timing["matrix_cols"] = int(A.shape[1])
        timing["matrix_nnz"] = int(A.nnz)
        t1 = perf_counter()
        A, b, S = explicit.preprocess(A, b, self.ridge_factor, self.add_ridge_regulatisation,
                                      self.apply_scaling_matrix)
        timing["preprocess_seconds"] = perf_counter() - t1
        logger.warning("Using custom solver")
        t2 = perf_counter()
        c = np.asarray(solver(A.tocsr(), b))
        timing["solve_seconds"] = perf_counter() - t2
        self.c = S @ c if S is not None else c
        self.up_to_date = True
        timing["total_seconds"] = perf_counter() - started
        timing["up_to_date"] = True
        self.latest_solve_timing = timing
        return True

    def _build_system(self):
        raise NotImplementedError

    def solve_system(self, solver=None, tol=None, solver_kwargs=None) -> bool:
        """LI/_discrete_interpolator.py:1221-1319 on the GPU.

        ``solver_kwargs`` follow scipy's cg: rtol (default 1e-5), atol
        (default 0; ``tol`` sets atol unless both rtol and atol are given),
        maxiter (default 10*dof).  Extra keys: ``precision`` ('fp64' |
        'fp32': FP32 operator + FP64 refinement), ``inner_rtol``,
        ``preconditioner`` ('jacobi', the reference's column scaling, or
        'mg': geometric multigrid, FD only; same stopping test and minimiser).
        """
        if not self._pre_solve():
            raise ValueError("Pre solve failed")
        started = perf_counter()
        choice = self._normalise_solver_choice(solver)
        kw = dict(solver_kwargs or {})
        cn_iters = max(0, int(kw.pop("constant_norm_iterations", 0)))
        cn_weight = float(kw.pop("constant_norm_weight", 0.0))
        cn_target = kw.pop("constant_norm_target", None)
        if cn_target is not None:
            cn_target = float(cn_target)
            if cn_target <= 0.0:
                logger.warning("constant_norm_target must be > 0; disabling explicit target and "
                               "using auto mode")
                cn_target = None
        kw.pop("backend", None)
        timing = {"solver": "callable" if callable(choice) else choice, "backend": "hip"}
        self.solver_history = None
        t_ie = perf_counter()
        Q, bounds = self.build_inequality_matrix()
        timing["inequality_seconds"] = perf_counter() - t_ie
        timing["inequality_rows"] = int(Q.shape[0])
        if callable(choice):
            return self._solve_callable(choice, timing, started)
        t0 = perf_counter()
        system, nrows_total, nnz = self._build_system()
        torch.cuda.synchronize()
        timing["assembly_seconds"] = perf_counter() - t0
        timing["matrix_rows"] = int(nrows_total)
        timing["matrix_cols"] = int(self.dof)
        timing["matrix_nnz"] = int(nnz)
        timing["preprocess_seconds"] = 0.0
        if choice == "lsmr":
            return self._solve_lsmr(self._lsmr_system(system), tol, kw, timing, started)
        if choice == "admm":
            return self._solve_admm(system, Q, bounds, kw, timing, started,
                                    (cn_iters, cn_weight, cn_target))
        if ("atol" not in kw or "rtol" not in kw) and tol is not None:
            kw["atol"] = tol
        for k in ("x0", "M", "callback"):
            if kw.pop(k, None) is not None:
                logger.warning("solver_kwargs[%r] is not supported by the GPU solver; ignored", k)
        t1 = perf_counter()
        pre = str(kw.pop("preconditioner", "jacobi")).lower()
        if pre not in ("jacobi", "mg", "multigrid"):
            raise ValueError(f"unknown preconditioner {pre!r}: use 'jacobi' or 'mg'")
        mg = pre in ("mg", "multigrid") and isinstance(system, fs.FdSystem) and \
            not isinstance(system, fs.P1System)
        c, info = system.solve(rtol=kw.pop("rtol", 1e-5), atol=kw.pop("atol", 0.0),
                               maxiter=kw.pop("maxiter", None),
                               precision=kw.pop("precision", "fp64"),
                               inner_rtol=kw.pop("inner_rtol", 1e-4),
                               jacobi=self.apply_scaling_matrix, multigrid=mg)
        torch.cuda.synchronize()
        timing["solve_seconds"] = perf_counter() - t1
        timing.update({"iterations": info["iterations"], "refinement_sweeps": info["outer"],
                       "residual": info["residual"], "threshold": info["threshold"]})
        if not info["converged"]:
            logger.warning("CG reached iteration limit (%s) and did not converge; using last "
                           "iteration", info["iterations"])
        self.c = c.cpu().numpy()
        self.up_to_date = True
        timing["total_seconds"] = perf_counter() - started
        timing["up_to_date"] = True
        self.latest_solve_timing = timing
        return self.up_to_date

    def _solve_admm(self, system, Q, bounds, kw, timing, started, cn) -> bool:
        """solve_with_admm (LI/_solver_strategy.py:213-247) and
        _solve_with_admm (LI/_discrete_interpolator.py:1171-1219): the
        reference's ADMM on the preprocessed system, every x-update solved
        on the GPU operator, then the optional constant-norm polish.  Without
        inequality rows the reference's ADMM set-up raises, which
        _solve_with_admm logs and reports as a failed solve."""
        logger.info("Solving using admm")
        cn_kwargs = dict(kw)
        x0 = kw.pop("x0", None)
        kw.pop("linsys_solver", None)
        opts = {k: kw.pop(k) for k in list(kw) if k in _ADMM_OPTIONS}
        for k in ("batch_size", "batch_fraction", "random_seed"):
            kw.pop(k, None)
        admm_weight = kw.pop("admm_weight", 0.01)
        nmajor = kw.pop("nmajor", 200)
        step_rtol = float(kw.pop("step_rtol", 1e-10))
        pre = str(kw.pop("preconditioner", "jacobi")).lower()
        for k in kw:
            logger.info("solver_kwargs[%r] sets the reference's inner LSMR; the GPU x-update "
                        "is solved to step_rtol instead", k)
        t1 = perf_counter()
        try:
            if Q.shape[0] == 0:
                raise ValueError("nelements must be greater than 0")
            ops = _AdmmOps(self, system, Q, pre in ("mg", "multigrid"), step_rtol)
            if x0 is not None:
                y0 = torch.as_tensor(np.asarray(x0(self.support), dtype=float),
                                     device=_lib.device())
                c0 = y0 * ops.S
            else:
                c0 = torch.zeros(self.dof, dtype=torch.float64, device=_lib.device())
            from .admm import admm_solve
            b = torch.as_tensor(np.asarray(bounds, dtype=float), device=_lib.device())
            res = admm_solve(ops, b, c0, admm_weight=admm_weight, nmajor=nmajor, **opts)
        except ValueError as e:
            logger.error(f"ADMM solver failed: {e}")
            timing["solve_seconds"] = perf_counter() - t1
            return self._finish(False, timing, started)
        torch.cuda.synchronize()
        timing["solve_seconds"] = perf_counter() - t1
        timing["step_iterations"] = ops.iterations
        c, history = (res if isinstance(res, tuple) else (res, None))
        self.c = c.cpu().numpy()
        self.solver_history = history
        self._finish(True, timing, started)
        iters, weight, target = cn
        if iters > 0 and weight > 0.0:
            self._run_constant_norm_polish(cn_kwargs, iters, weight, target)
        return True

    def _finish(self, ok, timing, started) -> bool:
        self.up_to_date = bool(ok)
        timing["total_seconds"] = perf_counter() - started
        timing["up_to_date"] = bool(ok)
        self.latest_solve_timing = timing
        return self.up_to_date

    def _remove_constraints_with_prefix(self, prefix: str) -> None:
        for name in [n for n in self.constraints if n.startswith(prefix)]:
            self.constraints.pop(name, None)
            self._rows.pop(name, None)

    def _constant_norm_gradient_data(self):
        """Gradient of the current solution at every element barycentre,
        with the element gradient rows, node ids and volumes, keeping elements
        whose gradient norm is finite and above 1e-10
        (LI/_discrete_interpolator.py:886-936).  On the device."""
        sup = self.support
        try:
            _, grad, elements, inside = sup.get_element_gradient_for_location(sup.barycentre)
        except (AttributeError, TypeError, ValueError) as err:
            logger.debug("Unable to build constant-norm gradient rows: %s", err)
            return None
        inside = np.asarray(inside, dtype=bool)
        if not np.any(inside):
            return None
        dev = _lib.device()
        element_ids = np.asarray(elements, dtype=int)[inside]
        T = torch.as_tensor(np.asarray(grad, dtype=float)[inside], device=dev)
        idc = torch.as_tensor(np.asarray(self._element_nodes(element_ids), dtype=np.int64),
                              device=dev)
        c = torch.as_tensor(np.asarray(self.c, dtype=float), device=dev)
        gvec = torch.einsum("ijk,ik->ij", T, c[idc])
        gnorm = torch.linalg.vector_norm(gvec, dim=1)
        valid = torch.isfinite(gnorm) & (gnorm > 1e-10)
        if not bool(valid.any()):
            return None
        size = np.asarray(sup.element_size, dtype=float)
        if size.ndim == 0:
            vol = torch.full((int(valid.sum()),), float(size), dtype=torch.float64, device=dev)
        else:
            vol = torch.as_tensor(size[element_ids], device=dev)[valid].clamp(min=1e-12)
        return {"gradient": T[valid], "grad_vec": gvec[valid], "grad_norm": gnorm[valid],
                "idc": idc[valid], "volume": vol}

    def _run_constant_norm_polish(self, solver_kwargs, iterations, base_weight, target_norm):
        """Pull the gradient norm towards a constant: rows ``g_hat . grad f =
        target`` in every element, weight ``base_weight * (k + 1)`` in round k,
        each round an ADMM solve from the current solution
        (LI/_discrete_interpolator.py:938-1004)."""
        if iterations <= 0 or base_weight <= 0.0:
            return
        if not np.all(np.isfinite(self.c)):
            logger.warning("Skipping constant-norm polish because current solution is not finite")
            return
        prefix = "__constant_norm_polish__"
        self._remove_constraints_with_prefix(prefix)
        stable = dict(solver_kwargs or {})
        stable.pop("x0", None)
        for i in range(int(iterations)):
            data = self._constant_norm_gradient_data()
            if data is None:
                logger.warning("Constant-norm polish stopped: gradient rows unavailable")
                break
            if target_norm is None:
                target = float(np.median(data["grad_norm"].cpu().numpy()))
            else:
                target = float(target_norm)
            unit = data["grad_vec"] / data["grad_norm"][:, None]
            A = torch.einsum("ij,ijk->ik", unit, data["gradient"]) / data["volume"][:, None]
            b = torch.full((A.shape[0],), target, dtype=torch.float64,
                           device=A.device) / data["volume"]
            self._add_device_rows(A, b, data["idc"], float(base_weight) * float(i + 1),
                                  f"{prefix}{i}")
            seed = np.asarray(self.c, dtype=float).copy()
            kw = dict(stable)
            kw["x0"] = lambda _support, x0=seed: np.array(x0, copy=True)
            kw["constant_norm_iterations"] = 0
            if not self.solve_system("admm", solver_kwargs=kw):
                logger.warning("Constant-norm polish terminated because ADMM solve failed")
                break
            upd = self._constant_norm_gradient_data()
            if upd is not None:
                gn = upd["grad_norm"]
                logger.info("Constant-norm iter %d: target=%.3e, grad_norm mean=%.3e, std=%.3e",
                            i + 1, target, float(gn.mean()), float(gn.std(unbiased=False)))
        self._remove_constraints_with_prefix(prefix)

    def _add_device_rows(self, A, b, idc, w, name):
        """add_constraints_to_least_squares for rows already on the device;
        ``w`` one weight or one per row."""
        A = A.contiguous().clone()
        b = b.contiguous().clone()
        R, k = A.shape
        norm = torch.empty(R, dtype=torch.float64, device=A.device)
        if R:
            _lib.call("lsamd_rows_normalise", R, k, _lib.ptr(A), _lib.ptr(b), _lib.ptr(norm),
                      _lib.stream())
        if isinstance(w, torch.Tensor):
            wv = w.to(A.device, torch.float64).contiguous()
        else:
            wv = torch.full((R,), float(w), dtype=torch.float64, device=A.device)
        blk = fs.fixed_width_rows(idc.to(torch.int32).contiguous(), A, b, wv, norm, self.dof)
        return self._register_rows(blk, name)

    def _lsmr_system(self, system):
        """The system LSMR iterates on; FD systems hold their regulariser
        as rows already."""
        return system

    def _solve_lsmr(self, system, tol, kw, timing, started) -> bool:
        """solve_with_lsmr (LI/_solver_strategy.py:108-136): LSMR on the
        ridge-augmented, column-scaled rectangular system, matrix-free on the
        GPU; ``tol`` sets btol (and atol = 0) unless btol is given."""
        if "btol" not in kw and tol is not None:
            kw["btol"] = tol
            kw["atol"] = 0.0
        for k in ("damp", "show", "x0"):
            if kw.pop(k, None):
                logger.warning("solver_kwargs[%r] is not supported by the GPU LSMR; ignored", k)
        kw.pop("preconditioner", None)
        kw.pop("precision", None)
        t1 = perf_counter()
        c, info = system.lsmr(atol=kw.pop("atol", 1e-6), btol=kw.pop("btol", 1e-6),
                              conlim=kw.pop("conlim", 1e8), maxiter=kw.pop("maxiter", None),
                              scaling=self.apply_scaling_matrix)
        torch.cuda.synchronize()
        for k in kw:
            logger.warning("solver_kwargs[%r] is not supported by the GPU LSMR; ignored", k)
        timing["solve_seconds"] = perf_counter() - t1
        timing.update({"iterations": info["iterations"], "istop": info["istop"],
                       "normr": info["normr"], "normar": info["normar"]})
        if info["istop"] == 0:
            logger.warning("Solution to least squares problem is all zeros, check input data")
        elif info["istop"] in (3, 6):
            logger.warning("COND(A) seems to be greater than CONLIM, check input data")
        self.c = c.cpu().numpy()
        self.up_to_date = True
        timing["total_seconds"] = perf_counter() - started
        timing["up_to_date"] = True
        self.latest_solve_timing = timing
        return self.up_to_date

    def update(self) -> bool:
        """LI/_discrete_interpolator.py:1321-1340."""
        if self.solver is None:
            return False
        if not self.up_to_date:
            self.setup_interpolator()
            self.up_to_date = self.solve_system(self.solver)
            return self.up_to_date
        return bool(self.up_to_date)

    def _c_device(self):
        """Device copy of ``self.c`` as it is now.  The reference reads ``c`` on
        every evaluation, and callers may assign it or change it in place
        (from_dict, a saved model), so it is uploaded on every call."""
        c = np.ascontiguousarray(self.c, dtype=float)
        if c.shape != (self.dof,):
            raise ValueError(f"c has shape {c.shape}, expected ({self.dof},)")
        return torch.from_numpy(c).to(_lib.device())

    def to_dict(self):
        # LI/_discrete_interpolator.py:1381-1388: base-class keys win
        return {"type": self.type.name, "support": self.support.to_dict(),
                "c": np.asarray(self.c).tolist(), **super().to_dict()}

    # ------------------------------------------------------ diagnostics --
    @staticmethod
    def _family_of(name, inequality=False):
        """LI/_discrete_interpolator.py:179-201."""
        n = name.lower()
        for pre, fam in (("value", "value"), ("gradient", "gradient"), ("norm", "normal"),
                         ("tangent", "tangent"), ("interface", "interface"),
                         ("inequality_value", "inequality_value"),
                         ("inequality_pairs", "inequality_pairs"), ("fold", "fold")):
            if n.startswith(pre):
                return fam
        if n.startswith("d") and len(n) in (3, 4, 5, 6, 7, 8):
            return "regularisation"
        return "inequality" if inequality else "other"

    def _build_constraint_diagnostics_report(self):
        """LI/_discrete_interpolator.py:203-278."""
        outside = self._outside_model_points_from_data()
        raw = {"value": self.get_value_constraints().shape[0],
               "gradient": self.get_gradient_constraints().shape[0],
               "normal": self.get_norm_constraints().shape[0],
               "tangent": self.get_tangent_constraints().shape[0],
               "interface": self.get_interface_constraints().shape[0],
               "inequality_value": self.get_inequality_value_constraints().shape[0],
               "inequality_pairs": self.get_inequality_pairs_constraints().shape[0]}
        rows = defaultdict(int)
        wts = defaultdict(list)
        for name, con in self.constraints.items():
            fam = self._family_of(name)
            m = dict.__getitem__(con, "matrix")
            rows[fam] += int(m.shape[0])
            w = con["w"]
            if w is not None:
                w = np.atleast_1d(np.asarray(w, dtype=float))
                if w.size:
                    wts[fam].append(w)
        for name, con in self.ineq_constraints.items():
            rows[self._family_of(name, inequality=True)] += int(con["matrix"].shape[0])
        fams = {}
        for fam in sorted(set(raw) | set(rows)):
            rc = int(rows.get(fam, 0))
            src = int(raw.get(fam, 0))
            dropped = None
            if fam in ("value", "normal", "tangent", "inequality_value"):
                dropped = max(src - rc, 0)
            elif fam == "gradient":
                dropped = max(src * 2 - rc, 0)
            if wts.get(fam):
                allw = np.concatenate(wts[fam])
                mw, lo, hi = float(np.mean(allw)), float(np.min(allw)), float(np.max(allw))
            else:
                mw = lo = hi = None
            fams[fam] = ConstraintFamilyDiagnostics(fam, rc > 0, rc, dropped, mw, lo, hi, src,
                                                    int(outside.get(fam, 0)))
        n = int(self.support.n_nodes)
        cov = RegionCoverageDiagnostics(n, n, 0, 1.0 if n else 0.0)
        return ConstraintDiagnosticsReport(self.type.name, fams, cov, outside)

    def finalize_setup_diagnostics_report(self):
        self.latest_diagnostics_report = self._build_constraint_diagnostics_report()
        return self.latest_diagnostics_report

    # ------------------------------------------------------ evaluation --
    def _evaluate_value_local(self, locations):
        """LI/_discrete_interpolator.py:1342-1362: NaN coordinates -> 0."""
        self.update()
        pts = np.array(locations, dtype=float)
        if pts.shape[0] == 0:
            return np.zeros(0)
        v, _ = self._evaluate(pts, True, False)
        return v

    def _evaluate_gradient_local(self, locations):
        """LI/_discrete_interpolator.py:1364-1379."""
        self.update()
        pts = np.asarray(locations, dtype=float)
        if pts.shape[0] > 0:
            _, g = self._evaluate(pts, False, True)
            return g
        return np.zeros((0, 3))


def _regularisation_weighting(support, pts, sigma):
    """``1 + 10 exp(-d^2 / (2 sigma^2))`` at every node, d = distance to the
    nearest of ``pts`` (compute_weighting, LI/_finite_difference_interpolator.py:66-88)."""
    dev = _lib.device()
    p = torch.from_numpy(np.ascontiguousarray(pts[:, :3], dtype=float)).to(dev)
    d2 = torch.empty(support.n_nodes, dtype=torch.float64, device=dev)
    g = support._lsgrid()
    _lib.call("lsamd_fd_nearest_d2", ctypes.byref(g), _lib.ptr(p), int(p.shape[0]),
              _lib.ptr(d2), _lib.stream())
    return 1.0 + 10.0 * torch.exp(-d2 / (2.0 * float(sigma) ** 2))


class FiniteDifferenceInterpolator(DiscreteInterpolator):
    """LI/_finite_difference_interpolator.py:91-1205 on the MI355X."""

    def __init__(self, grid, data=None):
        self.shape = "rectangular"
        DiscreteInterpolator.__init__(self, grid, data=data)
        self.set_interpolation_weights({"dxy": 1.0, "dyz": 1.0, "dxz": 1.0, "dxx": 1.0,
                                        "dyy": 1.0, "dzz": 1.0, "dx": 1.0, "dy": 1.0,
                                        "dz": 1.0, "cpw": 1.0, "gpw": 1.0, "npw": 10.0,
                                        "tpw": 1.0, "ipw": 1.0})
        self.type = InterpolatorType.FINITE_DIFFERENCE
        self.use_regularisation_weight_scale = False
        self.regularisation_weight_sigma = None
        self.interface_pair_mode = "auto"
        self.interface_pair_threshold = 200
        self.interface_pair_fallback = "star"
        self.regularisation_matrix_free = False
        self.matrix_free_regularisation_blocks = {}

    def copy(self):
        return FiniteDifferenceInterpolator(self.support)

    # ------------------------------------------------------------ setup --
    def setup_interpolator(self, **kwargs):
        """LI/_finite_difference_interpolator.py:132-224."""
        self.reset()
        reg = kwargs.get("regularisation", None)
        directional = kwargs.get("directional_regularisation", None)
        isotropic = None
        if reg is not None:
            if np.isscalar(reg):
                isotropic = float(reg)
            elif isinstance(reg, dict) and ("isotropic" in reg or "weight" in reg) and \
                    "direction" not in reg and "vector" not in reg:
                isotropic = float(reg.get("isotropic", reg.get("weight")))
                directional = directional or reg.get("directional") or None
            else:
                directional = reg
        if directional:
            raise NotImplementedError("directional regularisation is not part of this package")
        self._apply_isotropic_regularisation_weight(
            isotropic, ("dxy", "dyz", "dxz", "dxx", "dyy", "dzz"))
        self._apply_interpolation_weight_kwargs(
            kwargs, skip_keys=("regularisation", "directional_regularisation"))
        operators = kwargs.get("operators",
                               self.support.get_operators(weights=self.interpolation_weights))
        self.use_regularisation_weight_scale = kwargs.get("use_regularisation_weight_scale", False)
        self.regularisation_weight_sigma = kwargs.get("regularisation_weight_sigma", None)
        self.regularisation_matrix_free = bool(kwargs.get("regularisation_matrix_free", False))
        if self.regularisation_matrix_free and self.apply_scaling_matrix:
            logger.warning("regularisation_matrix_free=True with apply_scaling_matrix=True: the "
                           "reference falls back to explicit assembly; the GPU operator is "
                           "matrix-free either way")
            self.regularisation_matrix_free = False
        self.interface_pair_mode = kwargs.get("interface_pair_mode", self.interface_pair_mode)
        self.interface_pair_threshold = int(kwargs.get("interface_pair_threshold",
                                                       self.interface_pair_threshold))
        self.interface_pair_fallback = kwargs.get("interface_pair_fallback",
                                                  self.interface_pair_fallback)
        self.add_norm_constraints(self.interpolation_weights["npw"])
        self.add_gradient_constraints(self.interpolation_weights["gpw"])
        self.add_value_constraints(self.interpolation_weights["cpw"])
        self.add_tangent_constraints(self.interpolation_weights["tpw"])
        self.add_interface_constraints(self.interpolation_weights["ipw"])
        self.add_value_inequality_constraints()
        self.add_inequality_pairs_constraints(
            pairs=kwargs.get("inequality_pairs", None),
            upper_bound=kwargs.get("inequality_pair_upper_bound", np.finfo(float).eps),
            lower_bound=kwargs.get("inequality_pair_lower_bound", -np.inf))
        for k, o in operators.items():
            self.assemble_inner(o[0], o[1], name=k)
        self.assemble_borders()
        return self.finalize_setup_diagnostics_report()

    def _element_nodes(self, element):
        """Corner node ids of the given cells, without building the table of
        every cell."""
        g = self.support
        return g.global_node_indices(g.cell_corner_indexes(
            g.global_index_to_cell_index(np.asarray(element))))

    # ------------------------------------------------------- data rows ----
    def _inside_points(self, points, ncols):
        """Device copy of the rows of ``points`` inside the grid, and the mask."""
        dev = _lib.device()
        pts = torch.from_numpy(np.ascontiguousarray(points[:, :ncols], dtype=float)).to(dev)
        _, ins, _, _ = fs.locate(self.support._lsgrid(), pts[:, :3].contiguous(), want_w=False)
        return pts[ins].contiguous(), ins

    def add_value_constraints(self, w=1.0):
        """Trilinear rows, b = value, weight cpw * w_pt (:236-269)."""
        points = self.get_value_constraints()
        if points.shape[0] > 0:
            pin, ins = self._inside_points(points, points.shape[1])
            wcol = self.support.dimension + 1 if points.shape[1] > 4 else -1
            blk = fs.build_rows(self.support._lsgrid(), _lib.ROWS_VALUE, pin, wcol, float(w),
                                self.dof)
            self._register_rows(blk, "value")
            if pin.shape[0] <= 0:
                logger.warning(f"{points.shape[0]} value constraints not added: outside of "
                               f"model bounding box")

    def add_norm_constraints(self, w=1.0):
        """Three rows per point T_d, b = n_d, weight npw/3 * w_pt (:403-476)."""
        points = self.get_norm_constraints()
        if points.shape[0] > 0:
            pin, ins = self._inside_points(points, points.shape[1])
            sigma = self.regularisation_weight_sigma
            if sigma is None:
                sigma = self.support.nsteps[0] * 10
            self._scale_sources.append((pin[:, :3].cpu().numpy(), sigma))
            self._regularisation_scale = None
            wcol = 2 * self.support.dimension if points.shape[1] > 6 else -1
            w = w / self.support.dimension
            blk = fs.build_rows(self.support._lsgrid(), _lib.ROWS_NORMAL, pin, wcol, float(w),
                                self.dof)
            n = pin.shape[0]
            for d in range(3):
                self._register_rows(_row_slice(blk, d * n, (d + 1) * n), f"norm_{d}")
            if n <= 0:
                logger.warning(f"{points.shape[0]} norm constraints not added: outside of "
                               f"model bounding box")
            self.up_to_date = False

    def add_gradient_constraints(self, w=1.0):
        """Strike and dip rows (:342-401); normalises the stored vectors in place."""
        points = self.get_gradient_constraints()
        if points.shape[0] > 0:
            pin, ins = self._inside_points(points, points.shape[1])
            blk = fs.build_rows(self.support._lsgrid(), _lib.ROWS_GRADIENT, pin, -1, float(w),
                                self.dof)
            insn = ins.cpu().numpy()
            norm = np.linalg.norm(points[~insn, 3:6], axis=1)
            points[~insn, 3:6] /= norm[:, None]
            points[insn, 3:6] = pin[:, 3:6].cpu().numpy()
            n = pin.shape[0]
            self._register_rows(_row_slice(blk, 0, n), "gradient")
            self._register_rows(_row_slice(blk, n, 2 * n), "gradient")
            if n <= 0:
                logger.warning(f" {points.shape[0]} norm constraints not added: outside of "
                               f"model bounding box")

    def add_tangent_constraints(self, w=1.0):
        """LI/_discrete_interpolator.py:668-682."""
        points = self.get_tangent_constraints()
        if points.shape[0] > 0:
            self.add_gradient_orthogonal_constraints(points[:, :3], points[:, 3:6], w)

    def add_gradient_orthogonal_constraints(self, points, vectors, w=1.0, b=0,
                                            name="gradient orthogonal"):
        """t_hat . grad = b rows (:478-539); ``vectors`` normalised in place."""
        points = np.asarray(points)
        if points.shape[0] > 0:
            arr = np.hstack([np.asarray(points[:, :3], dtype=float),
                             np.asarray(vectors[:, :3], dtype=float)])
            pin, ins = self._inside_points(arr, 6)
            blk = fs.build_rows(self.support._lsgrid(), _lib.ROWS_TANGENT, pin, -1, float(w),
                                self.dof)
            insn = ins.cpu().numpy()
            vout = np.asarray(vectors)
            norm = np.linalg.norm(vout[~insn, :3], axis=1)
            sel = np.nonzero(~insn)[0][norm > 0]
            vout[sel, :3] /= norm[norm > 0, None]
            vout[insn, :3] = pin[:, 3:6].cpu().numpy()
            if b != 0:
                bb = torch.full_like(blk.b, float(b))
                blk.b = torch.where(blk.norm > 0, bb / blk.norm, bb)
            self._register_rows(blk, name)
            if pin.shape[0] <= 0:
                logger.warning(f"{points.shape[0]} gradient constraints not added: outside of "
                               f"model bounding box")
            self.up_to_date = False

    def _interface_pair_indices(self, n: int) -> np.ndarray:
        """Point pairs tied by one interface id (LI/_finite_difference_interpolator.py:271-295):
        mode 'all' (every i < j), 'star' (0 with each other) or 'chain'
        (neighbours in order); 'auto' is 'all' up to the threshold and the
        fallback ('star' unless 'chain') above it."""
        if n < 2:
            return np.zeros((0, 2), dtype=int)
        mode = str(self.interface_pair_mode).lower()
        if mode == "auto":
            fallback = str(self.interface_pair_fallback).lower()
            large = n > max(2, int(self.interface_pair_threshold))
            mode = (fallback if fallback == "chain" else "star") if large else "all"
        make = _PAIR_MODES.get(mode)
        if make is None:
            raise ValueError(f"Unknown interface_pair_mode '{self.interface_pair_mode}'. Use one "
                             f"of: auto, all, star, chain")
        return make(n).astype(int)

    def add_interface_constraints(self, w=1.0):
        """Pairwise equal-value rows per interface id (:297-340), 16 entries per
        row, normalised on the GPU."""
        points = self.get_interface_constraints()
        if points.shape[0] > 1:
            dev = _lib.device()
            pts = torch.from_numpy(np.ascontiguousarray(points[:, :3], dtype=float)).to(dev)
            idc, ins, wt, _ = fs.locate(self.support._lsgrid(), pts, want_w=True)
            insn = ins.cpu().numpy()
            idc_in, w_in = idc[ins], wt[ins]
            ids = points[insn, 3]
            for uid in np.unique(points[np.logical_and(~np.isnan(points[:, 3]), insn), 3]):
                mask = np.nonzero(ids == uid)[0]
                pair = self._interface_pair_indices(len(mask))
                if pair.shape[0] == 0:
                    continue
                m = torch.from_numpy(mask).to(dev)
                p0 = torch.from_numpy(pair[:, 0]).to(dev)
                p1 = torch.from_numpy(pair[:, 1]).to(dev)
                A = torch.cat([w_in[m][p0], -w_in[m][p1]], dim=1).contiguous()
                I = torch.cat([idc_in[m][p0], idc_in[m][p1]], dim=1).to(torch.int32).contiguous()
                R = A.shape[0]
                b = torch.zeros(R, dtype=torch.float64, device=dev)
                norm = torch.empty(R, dtype=torch.float64, device=dev)
                _lib.call("lsamd_rows_normalise", R, 16, _lib.ptr(A), _lib.ptr(b),
                          _lib.ptr(norm), _lib.stream())
                wv = torch.full((R,), float(w), dtype=torch.float64, device=dev)
                self._register_rows(fs.fixed_width_rows(I, A, b, wv, norm, self.dof),
                                    f"interface_{uid}")

    # ------------------------------------------------------ regulariser ----
    def assemble_inner(self, operator, w, name="regularisation"):
        """Record an interior stencil family (:1146-1164); matrix-free.  A
        ``None`` operator (rectilinear grids) builds spacing-scaled rows."""
        if operator is None:
            self._assemble_rectilinear_operator(name, w)
            return
        kind = _mask_kind(operator)
        if kind is None:
            raise NotImplementedError(
                f"custom regularisation mask for '{name}' is not supported by the GPU operator")
        n = self.support.nsteps
        nrows = int(np.prod(n - 2)) if np.all(n >= 3) else 0
        name = self._unique_name(name)
        fam = StencilFamily(name, kind, float(w), nrows, self.dof, np.asarray(operator, float),
                            tuple(int(v) for v in n))
        self._add_stencil_family(fam)

    def _assemble_rectilinear_operator(self, name, w):
        """Second-derivative rows scaled by the local node spacing at every
        node interior along the axes of the family, as data rows on the GPU
        (LI/_finite_difference_interpolator.py:1166-1205)."""
        axes = {"dxx": (0, -1), "dyy": (1, -1), "dzz": (2, -1), "dxy": (0, 1), "dxz": (0, 2),
                "dyz": (1, 2)}
        if name not in axes:
            logger.warning(f"Unknown rectilinear operator name '{name}', skipping.")
            return
        from .supports import scaled_operator_rows

        nodes = (self.support.xnodes, self.support.ynodes, self.support.znodes)
        A, cols, centres = scaled_operator_rows(nodes, *axes[name])
        w_rows = torch.full((A.shape[0],), float(w), dtype=torch.float64, device=A.device)
        if self.use_regularisation_weight_scale:
            self.regularisation_scale  # materialise on the device
            w_rows = self._regularisation_scale[centres] * float(w)
        self._add_device_rows(A, torch.zeros(A.shape[0], dtype=torch.float64, device=A.device),
                              cols, w_rows, name)

    def assemble_borders(self):
        """Border layers 1 and n-2 with the axial masks, weights dx/dy/dz
        (:1052-1144); skipped when the weight is 0 or the axis has < 3 nodes."""
        n = self.support.nsteps
        masks = (Operator.Dxx_mask, Operator.Dyy_mask, Operator.Dzz_mask)
        for axis, key, kind in ((0, "dx", "bx"), (1, "dy", "by"), (2, "dz", "bz")):
            weight = self.interpolation_weights[key]
            if weight == 0 or n[axis] < 3:
                continue
            nrows = int(np.prod(n) // n[axis])
            for side in ("lower", "upper"):
                name = self._unique_name(f"{key}_{side}")
                fam = StencilFamily(name, kind, float(weight), nrows, self.dof,
                                    np.asarray(masks[axis], float), tuple(int(v) for v in n),
                                    (axis, side == "upper"))
                self._add_stencil_family(fam)

    def _add_stencil_family(self, fam):
        """Record a matrix-free family; its row weights are the family weight,
        times the per-node regularisation scale at each row's centre when
        use_regularisation_weight_scale is set (:670-674)."""
        self._stencils[fam.name] = fam

        def weights():
            w = np.full(fam.nrows, fam.weight)
            if self.use_regularisation_weight_scale:
                w = self.regularisation_scale[fam.centres()] * fam.weight
            return w

        self.constraints[fam.name] = _Constraint(matrix=fam, b=np.zeros(fam.nrows),
                                                 w=_Lazy(weights))

    def _reg_struct(self):
        r = _lib.FdReg()
        acc = defaultdict(float)
        for fam in self._stencils.values():
            acc[fam.kind] += fam.weight ** 2
        r.w2_xx, r.w2_yy, r.w2_zz = acc["xx"], acc["yy"], acc["zz"]
        r.w2_xy, r.w2_xz, r.w2_yz = acc["xy"], acc["xz"], acc["yz"]
        # lower and upper layers share one weight slot each
        r.w2_bx, r.w2_by, r.w2_bz = acc["bx"] / 2, acc["by"] / 2, acc["bz"] / 2
        r.ridge2 = float(self.ridge_factor) ** 2 if self.add_ridge_regulatisation else 0.0
        return r

    def _build_system(self):
        rows = self._system_rows()
        scale2 = None
        if self.use_regularisation_weight_scale:
            self.regularisation_scale  # materialise on the device
            scale2 = self._regularisation_scale ** 2
        system = fs.FdSystem(self.support._lsgrid(), self._reg_struct(), scale2)
        system.set_rows(rows)
        self._system = system
        nreg = sum(f.nrows for f in self._stencils.values())
        ndata = rows.nrows if rows is not None else 0
        nnz = (rows.nnz if rows is not None else 0) + sum(
            f.nrows * (3 if f.kind in ("xx", "yy", "zz", "bx", "by", "bz") else 4)
            for f in self._stencils.values())
        return system, ndata + nreg, nnz

    def _evaluate(self, pts, values, gradients):
        v, g = fs.evaluate(self.support._lsgrid(), self._c_device(), torch.from_numpy(pts),
                           values, gradients)
        return (v.cpu().numpy() if v is not None else None,
                g.cpu().numpy() if g is not None else None)


def _row_slice(blk: fs.FdRows, r0: int, r1: int) -> fs.FdRows:
    """Rows [r0, r1) of a fixed-width block."""
    k = int(blk.row_ptr[1] - blk.row_ptr[0]) if blk.nrows else 8
    rp = blk.row_ptr[r0:r1 + 1] - blk.row_ptr[r0] if blk.nrows else blk.row_ptr[:1]
    e0, e1 = r0 * k, r1 * k
    return fs.FdRows(rp, blk.cols[e0:e1], blk.coef[e0:e1], blk.b[r0:r1], blk.w[r0:r1],
                     blk.norm[r0:r1], blk.dof)


class EdgeJumpFamily:
    """The P1 edge-jump regulariser kept matrix-free: one normalised 8-entry
    row per face-sharing tet pair, uniform weight."""

    def __init__(self, name, weight, nrows, dof, grid=None):
        self.name, self.weight, self.nrows, self.dof = name, weight, nrows, dof
        self.grid = grid

    @property
    def shape(self):
        return (self.nrows, self.dof)

    def to_scipy(self):
        """The normalised rows from the GPU face table, in pair order."""
        return fs.p1_edge_rows(self.grid, 1.0, self.dof).to_scipy()


class P1Interpolator(DiscreteInterpolator):
    """Piecewise-linear interpolator on a TetMesh (LI/_p1interpolator.py:24-300)
    on the MI355X.  The edge-jump regulariser is applied as a 21-tap stencil
    (or as explicit rows when per-pair weights are requested)."""

    def __init__(self, mesh):
        self.shape = "rectangular"
        DiscreteInterpolator.__init__(self, mesh)
        self.support = mesh
        self.interpolation_weights = {"cgw": 0.1, "cpw": 1.0, "npw": 1.0, "gpw": 1.0,
                                      "tpw": 1.0, "ipw": 1.0}
        self.type = InterpolatorType.PIECEWISE_LINEAR
        self.use_regularisation_weight_scale = False
        self.regularisation_weight_sigma = None
        self._edge_families = OrderedDict()

    def reset(self):
        super().reset()
        self._edge_families = OrderedDict()

    def clear_constraints(self):
        super().clear_constraints()
        self._edge_families = OrderedDict()

    def _inside_points(self, points, ncols):
        """Device copy of the rows of ``points`` inside the mesh (strict box and
        a containing tet, as evaluate_shape) and the mask."""
        dev = _lib.device()
        pts = torch.from_numpy(np.ascontiguousarray(points[:, :ncols], dtype=float)).to(dev)
        _, _, ins, _, _ = fs.p1_locate(self.support._lsgrid(), pts[:, :3].contiguous(), False)
        return pts[ins].contiguous(), ins

    def _update_regularisation_scale_from_norm_constraints(self):
        """:50-68: += 1 + 10 exp(-d^2/2 sigma^2), d to the nearest normal point."""
        points = self.get_norm_constraints()
        if points.shape[0] == 0:
            return
        pin, ins = self._inside_points(points, 3)
        if pin.shape[0] == 0:
            return
        sigma = self.regularisation_weight_sigma
        if sigma is None:
            sigma = getattr(self.support, "nsteps", np.array([1.0], dtype=float))[0] * 10
        self._scale_sources.append((pin.cpu().numpy(), sigma))
        self._regularisation_scale = None

    # ------------------------------------------------------------ rows ----
    def add_gradient_constraints(self, w=1.0):
        """A no-op in the reference (:70-71)."""
        pass

    def add_norm_constraints(self, w=1.0):
        """Three rows grad(lambda)_d per point, b = n_d, weight w (:73-96)."""
        points = self.get_norm_constraints()
        if points.shape[0] > 0:
            pin, _ = self._inside_points(points, 6)
            blk = fs.p1_build_rows(self.support._lsgrid(), _lib.ROWS_NORMAL, pin, float(w),
                                   self.dof)
            self._register_rows(blk, "norm")
            self.up_to_date = False

    def add_value_constraints(self, w=1.0):
        """Barycentric rows, b = value, weight w (:98-113)."""
        points = self.get_value_constraints()
        if points.shape[0] > 0:
            pin, _ = self._inside_points(points, 4)
            blk = fs.p1_build_rows(self.support._lsgrid(), _lib.ROWS_VALUE, pin, float(w),
                                   self.dof)
            self._register_rows(blk, "value")
            self.up_to_date = False

    def add_gradient_orthogonal_constraints(self, points, vectors, w=1.0, b=0,
                                            name="undefined gradient orthogonal constraint"):
        """vector . grad = b rows weighted by w * tet volume (:248-296);
        ``vectors`` is normalised in place."""
        points = np.asarray(points)
        if points.shape[0] > 0:
            vectors = np.asarray(vectors)
            norm = np.linalg.norm(vectors, axis=1)
            vectors[norm > 0, :] /= norm[norm > 0, None]
            arr = np.hstack([np.asarray(points[:, :3], dtype=float),
                             np.asarray(vectors[:, :3], dtype=float)])
            pin, ins = self._inside_points(arr, 6)
            blk = fs.p1_build_rows(self.support._lsgrid(), _lib.ROWS_TANGENT, pin, float(w),
                                   self.dof, bval=float(b))
            self._register_rows(blk, name)
            if pin.shape[0] <= 0:
                logger.warning(f"{points.shape[0]} gradient constraints not added: outside of "
                               f"model bounding box")
            self.up_to_date = False

    def add_interface_constraints(self, w: float = 1):
        raise NotImplementedError

    def minimise_edge_jumps(self, w=0.1, vector_func=None, vector=None, name="edge jump"):
        """(n . grad)|T1 - (n . grad)|T2 = 0 over every shared face (:115-166)."""
        if vector_func is not None or vector is not None:
            raise NotImplementedError("directional regularisation is not part of this package")
        grid = self.support._lsgrid()
        P = fs.p1_n_pairs(grid)
        if self.use_regularisation_weight_scale:
            self.regularisation_scale  # materialise on the device
            blk = fs.p1_edge_rows(grid, float(w), self.dof, self._regularisation_scale)
            self._register_rows(blk, name)
        else:
            name = self._unique_name(name)
            fam = EdgeJumpFamily(name, float(w), P, self.dof, grid)
            self._edge_families[name] = fam
            self.constraints[name] = _Constraint(matrix=fam, b=np.zeros(P), w=np.full(P, float(w)))
        self.up_to_date = False

    def get_regularisation_sample_points(self):
        f = self.support.shared_elements
        return self.support.nodes[f].mean(axis=1)

    # ------------------------------------------------------------ setup --
    def setup_interpolator(self, **kwargs):
        """LI/_p1interpolator.py:180-246."""
        self.reset()
        reg = kwargs.get("regularisation", None)
        directional = kwargs.get("directional_regularisation", None)
        isotropic = None
        if reg is not None:
            if np.isscalar(reg):
                isotropic = float(reg)
            elif isinstance(reg, dict) and ("isotropic" in reg or "weight" in reg) and \
                    "direction" not in reg and "vector" not in reg:
                isotropic = float(reg.get("isotropic", reg.get("weight")))
                directional = directional or reg.get("directional") or None
            else:
                directional = reg
        if directional:
            raise NotImplementedError("directional regularisation is not part of this package")
        self._apply_isotropic_regularisation_weight(isotropic, ("cgw",))
        self._apply_interpolation_weight_kwargs(
            kwargs, skip_keys=("regularisation", "directional_regularisation"))
        self.use_regularisation_weight_scale = kwargs.get("use_regularisation_weight_scale", False)
        self.regularisation_weight_sigma = kwargs.get("regularisation_weight_sigma", None)
        if self.use_regularisation_weight_scale:
            self._update_regularisation_scale_from_norm_constraints()
        if self.interpolation_weights["cgw"] > 0.0:
            self.up_to_date = False
            self.minimise_edge_jumps(self.interpolation_weights["cgw"])
        self.add_gradient_constraints(self.interpolation_weights["gpw"])
        self.add_norm_constraints(self.interpolation_weights["npw"])
        self.add_value_constraints(self.interpolation_weights["cpw"])
        self.add_tangent_constraints(self.interpolation_weights["tpw"])
        self.add_value_inequality_constraints()
        self.add_inequality_pairs_constraints()
        return self.finalize_setup_diagnostics_report()

    def add_tangent_constraints(self, w=1.0):
        """LI/_discrete_interpolator.py:668-682."""
        points = self.get_tangent_constraints()
        if points.shape[0] > 0:
            self.add_gradient_orthogonal_constraints(points[:, :3], points[:, 3:6], w)

    # ------------------------------------------------------------ solve --
    def _build_system(self):
        rows = self._system_rows()
        w2 = sum(f.weight ** 2 for f in self._edge_families.values())
        ridge = float(self.ridge_factor) if self.add_ridge_regulatisation else 0.0
        system = fs.P1System(self.support._lsgrid(), float(np.sqrt(w2)), ridge)
        system.set_rows(rows)
        self._system = system
        nreg = sum(f.nrows for f in self._edge_families.values())
        ndata = rows.nrows if rows is not None else 0
        nnz = (rows.nnz if rows is not None else 0) + 8 * nreg
        return system, ndata + nreg, nnz

    def _lsmr_system(self, system):
        """LSMR needs the edge-jump operator as rows (LI/_p1interpolator.py:115-166
        adds them to the explicit matrix that solve_with_lsmr iterates on,
        LI/_solver_strategy.py:108-136).  The families share one set of rows
        and differ only in weight, so they are merged into one block of
        weight sqrt(sum w^2): A^T A, A^T b and |b| are unchanged, and so are
        the LSMR iterates.  Data rows, those edge rows and the ridge go to a
        grid system whose finite-difference regulariser is zero."""
        grid = self.support._lsgrid()
        blocks = list(self._rows.values())
        if self._edge_families:
            w2 = sum(f.weight ** 2 for f in self._edge_families.values())
            blocks.append(fs.p1_edge_rows(grid, float(np.sqrt(w2)), self.dof))
        ridge = float(self.ridge_factor) if self.add_ridge_regulatisation else 0.0
        zero = {k: 0.0 for k in ("dxx", "dyy", "dzz", "dxy", "dxz", "dyz")}
        ls = fs.FdSystem(grid, fs.reg_struct(zero, grid.nnodes, ridge))
        ls.set_rows(fs.concat_rows(blocks, self.dof))
        return ls

    def _evaluate(self, pts, values, gradients):
        v, g = fs.p1_evaluate(self.support._lsgrid(), self._c_device(), torch.from_numpy(pts),
                              values, gradients)
        v = v.cpu().numpy() if v is not None else None
        if v is not None:
            v[np.any(np.isnan(pts[:, :3]), axis=1)] = 0.0
        return v, (g.cpu().numpy() if g is not None else None)

    def copy(self):
        return P1Interpolator(self.support)
