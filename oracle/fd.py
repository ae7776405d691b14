"""CPU restatement of the reference finite-difference interpolation path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Reference paths are
relative to the LoopStructural checkout; ``LI/`` and ``LC/`` as in the package
docstring.
"""
from __future__ import annotations

import numpy as np
from scipy import sparse
from scipy.sparse import linalg as spla

EPS = np.finfo(float).eps

# Corner order of a cell (x fastest): 000,100,010,110,001,101,011,111.
# LC/supports/_3d_base_structured.py:369-412
CORNER_DX = np.array([0, 1, 0, 1, 0, 1, 0, 1])
CORNER_DY = np.array([0, 0, 1, 1, 0, 0, 1, 1])
CORNER_DZ = np.array([0, 0, 0, 0, 1, 1, 1, 1])


class Grid:
    """Geometry of a ``StructuredGrid``: ``nsteps`` cells -> ``nsteps+1`` nodes.

    LC/supports/_3d_base_structured.py:24-76 (nodes = cells + 1 at :66),
    maximum at :187-189, nodes at :233-245 (F-order, x fastest).
    """

    def __init__(self, origin, nsteps_cells, step_vector):
        self.origin = np.array(origin, dtype=float)
        self.step = np.array(step_vector, dtype=float)
        self.ncells = np.array(nsteps_cells, dtype=int)
        self.nnodes = self.ncells + 1

    @property
    def maximum(self):
        return self.origin + self.ncells * self.step

    @property
    def n_nodes(self):
        return int(np.prod(self.nnodes))

    def nodes(self):
        mx = self.maximum
        x = np.linspace(self.origin[0], mx[0], self.nnodes[0])
        y = np.linspace(self.origin[1], mx[1], self.nnodes[1])
        z = np.linspace(self.origin[2], mx[2], self.nnodes[2])
        xx, yy, zz = np.meshgrid(x, y, z, indexing="ij")
        return np.array([xx.flatten(order="F"), yy.flatten(order="F"), zz.flatten(order="F")]).T

    def node_index(self, g):
        """Global node id -> (i, j, k).  _3d_base_structured.py:464-485."""
        g = np.asarray(g)
        nx, ny = self.nnodes[0], self.nnodes[1]
        return np.stack([g % nx, g // nx % ny, g // nx // ny], axis=-1)


# ---------------------------------------------------------------- A2-A4 ----
def inside(grid: Grid, pos):
    """_3d_base_structured.py:287-293 (tolerance eps*32*max(extent,1))."""
    extent = np.maximum(np.abs(grid.maximum - grid.origin), 1.0)
    tol = EPS * 32.0 * extent
    return np.all((pos >= (grid.origin - tol)) & (pos <= (grid.maximum + tol)), axis=1)


def cell_index(grid: Grid, pos):
    """_3d_base_structured.py:251-282: floor((p-o)/h) with a true divide, clipped."""
    pos = np.asarray(pos, dtype=float)
    ins = inside(grid, pos)
    ci = np.zeros((pos.shape[0], 3), dtype=int)
    for a in range(3):
        x = pos[:, a] - grid.origin[a]
        ci[ins, a] = np.floor(x[ins] / grid.step[a]).astype(int)
        ci[ins, a] = np.clip(ci[ins, a], 0, grid.ncells[a] - 1)
    return ci, ins


def cell_corners(grid: Grid, pos):
    """_3d_base_structured.py:386-412: 8 corner node ids, -1 outside."""
    ci, ins = cell_index(grid, pos)
    nx, ny = grid.nnodes[0], grid.nnodes[1]
    offsets = np.array([0, 1, nx, nx + 1, nx * ny, nx * ny + 1, nx * ny + nx, nx * ny + nx + 1],
                       dtype=np.intp)
    g = ci[:, 0] + nx * ci[:, 1] + nx * ny * ci[:, 2]
    idc = g[:, None] + offsets[None, :]
    idc[~ins] = -1
    return idc, ins


def local_coords(grid: Grid, pos):
    """_3d_structured_grid.py:126-146: clip((p - (o + h*cell))/h, 0, 1)."""
    pos = np.asarray(pos, dtype=float)
    ci, _ = cell_index(grid, pos)
    cell_origin = grid.origin[None, :] + grid.step[None, :] * ci
    lc = (pos - cell_origin) / grid.step[None, :]
    return np.clip(lc, 0.0, 1.0)


def trilinear(lc):
    """_3d_structured_grid.py:90-124 (same products, same association)."""
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


def element_gradient(grid: Grid, pos):
    """_3d_structured_grid.py:393-454: T[:, d, k] = dN_k/dx_d / h_d."""
    T = element_gradient_from_local(local_coords(grid, pos))
    T[:, 0, :] /= grid.step[None, 0]
    T[:, 1, :] /= grid.step[None, 1]
    T[:, 2, :] /= grid.step[None, 2]
    return T


def element_gradient_from_local(lc):
    """dN_k/d(local)_d of the trilinear shape functions, _3d_structured_grid.py:425-453."""
    x, y, z = lc[:, 0], lc[:, 1], lc[:, 2]
    T = np.zeros((lc.shape[0], 3, 8))
    T[:, 0, 0] = (1 - z) * (y - 1)
    T[:, 0, 1] = (1 - y) * (1 - z)
    T[:, 0, 2] = -y * (1 - z)
    T[:, 0, 4] = -(1 - y) * z
    T[:, 0, 5] = (1 - y) * z
    T[:, 0, 6] = -y * z
    T[:, 0, 3] = y * (1 - z)
    T[:, 0, 7] = y * z
    T[:, 1, 0] = (x - 1) * (1 - z)
    T[:, 1, 1] = -x * (1 - z)
    T[:, 1, 2] = (1 - x) * (1 - z)
    T[:, 1, 4] = -(1 - x) * z
    T[:, 1, 5] = -x * z
    T[:, 1, 6] = (1 - x) * z
    T[:, 1, 3] = x * (1 - z)
    T[:, 1, 7] = x * z
    T[:, 2, 0] = -(1 - x) * (1 - y)
    T[:, 2, 1] = -x * (1 - y)
    T[:, 2, 2] = -(1 - x) * y
    T[:, 2, 4] = (1 - x) * (1 - y)
    T[:, 2, 5] = x * (1 - y)
    T[:, 2, 6] = (1 - x) * y
    T[:, 2, 3] = -x * y
    T[:, 2, 7] = x * y
    return T


def evaluate_value(grid: Grid, c, pos):
    """_3d_structured_grid.py:295-322 + _discrete_interpolator.py:1342-1362.

    Outside -> NaN; NaN coordinates -> 0.0; row sum = numpy pairwise order.
    """
    pos = np.asarray(pos, dtype=float)
    out = np.zeros(pos.shape[0])
    nanmask = np.any(np.isnan(pos), axis=1)
    p = pos[~nanmask]
    if p.shape[0] > 0:
        idc, ins = cell_corners(grid, p)
        v = np.full(idc.shape, np.nan)
        v[ins, :] = trilinear(local_coords(grid, p[ins]))
        v[ins, :] *= c[idc[ins, :]]
        out[~nanmask] = np.sum(v, axis=1)
    return out


def evaluate_gradient(grid: Grid, c, pos):
    """_3d_structured_grid.py:324-391: outside -> 0 (not NaN)."""
    pos = np.asarray(pos, dtype=float)
    idc, ins = cell_corners(grid, pos)
    T = np.zeros((idc.shape[0], 3, 8))
    T[ins] = element_gradient(grid, pos[ins])
    for d in range(3):
        T[ins, d, :] *= c[idc[ins, :]]
    return np.array([np.sum(T[:, 0, :], axis=1), np.sum(T[:, 1, :], axis=1),
                     np.sum(T[:, 2, :], axis=1)]).T


# ------------------------------------------------------------------- A5 ----
def get_vectors(normal):
    """LC/math/_maths.py:276-298 via :126-160 and :301-332.

    Returns (strike, dip), each (3, N).  ``normal`` is normalised in place,
    as in the reference.
    """
    length = np.linalg.norm(normal, axis=1)[:, None]
    normal /= length
    nv = np.array(normal)
    nv /= np.linalg.norm(nv, axis=1)[:, None]
    strike = np.rad2deg(-np.arctan2(nv[:, 1], nv[:, 0]))
    s = np.deg2rad(strike)
    strike_vec = np.array([np.sin(-s), -np.cos(-s), np.zeros(s.shape[0])])
    strike_vec /= np.linalg.norm(strike_vec, axis=0)[None, :]
    dip_vec = np.cross(strike_vec, normal, axisa=0, axisb=1).T
    dip_vec /= np.linalg.norm(dip_vec, axis=0)[None, :]
    return strike_vec * length.T, dip_vec * length.T


# ------------------------------------------------------------- A6, A8 ----
# Operator masks, (3,3,3) indexed [dz+1][dy+1][dx+1].
# LC/math/finite_difference_stencil.py:12-38 (= LI/_operator.py:11-44)
_Z = np.zeros((3, 3))
DXX = np.array([_Z, [[0, 0, 0], [1, -2, 1], [0, 0, 0]], _Z], dtype=float)
DYY = DXX.swapaxes(1, 2)
DZZ = DXX.swapaxes(0, 2)
DXY = np.array([_Z, [[-0.25, 0, 0.25], [0, 0, 0], [0.25, 0, -0.25]], _Z]) / np.sqrt(2)
DXZ = DXY.swapaxes(0, 1)
DYZ = DXY.swapaxes(0, 2)

# full neighbour mask: flatten index m = a*9 + b*3 + c <-> (dx=c-1, dy=b-1, dz=a-1)
# LI/_finite_difference_interpolator.py:541-634
_FULL = np.array([[c - 1 for a in range(3) for b in range(3) for c in range(3)],
                  [b - 1 for a in range(3) for b in range(3) for c in range(3)],
                  [a - 1 for a in range(3) for b in range(3) for c in range(3)]])


def normalise_rows(A, B):
    """LI/_discrete_interpolator.py:361-366: a/|a|, b/|a| where |a|>0."""
    A = np.array(A, dtype=float)
    B = np.array(B, dtype=float)
    length = np.linalg.norm(A, axis=1)
    B[length > 0] /= length[length > 0]
    A[length > 0, :] /= length[length > 0, None]
    return A, B


class LSSystem:
    """Ordered constraint families {name: (A csc, b, w)} as in
    LI/_discrete_interpolator.py:316-389 (dict insertion order = row order)."""

    def __init__(self, dof):
        self.dof = dof
        self.constraints = {}

    def add(self, A, B, idc, w, name):
        """add_constraints_to_least_squares, LI/_discrete_interpolator.py:316-389."""
        A = np.array(A, dtype=float)
        B = np.array(B, dtype=float)
        idc = np.array(idc)
        if A.shape != idc.shape:
            return
        if A.ndim > 2:
            if isinstance(w, np.ndarray):
                w = np.tile(w, (A.shape[1]))
            A = A.reshape((A.shape[0] * A.shape[1], A.shape[2]))
            idc = idc.reshape((idc.shape[0] * idc.shape[1], idc.shape[2]))
            B = B.reshape(A.shape[0])
        n_rows = A.shape[0]
        if np.any(np.isnan(idc)) or np.any(np.isnan(A)) or np.any(np.isnan(B)):
            return
        A, B = normalise_rows(A, B)
        if isinstance(w, (float, int)):
            w = np.ones(A.shape[0]) * w
        if not isinstance(w, np.ndarray):
            raise TypeError("w must be a numpy array")
        if w.shape[0] != A.shape[0]:
            raise ValueError("Weight array does not match number of constraints")
        base = name
        while name in self.constraints:
            count = 0
            if "_" in name:
                count = int(name.split("_")[1]) + 1
            name = base + f"_{count}"
        rows = np.tile(np.arange(n_rows), (A.shape[-1], 1)).T
        M = sparse.coo_matrix((A.flatten(), (rows.flatten(), idc.flatten())),
                              shape=(n_rows, self.dof)).tocsc()
        self.constraints[name] = {"matrix": M, "b": B.flatten(), "w": w}

    def build_matrix(self):
        """LI/_discrete_interpolator.py:684-705."""
        mats, bs = [], []
        for c in self.constraints.values():
            if len(c["w"]) == 0:
                continue
            mats.append(c["matrix"].multiply(c["w"][:, None]))
            bs.append(c["b"] * c["w"])
        if not mats:
            return sparse.csr_matrix((0, self.dof)), np.zeros(0)
        return sparse.vstack(mats).tocsr(), np.hstack(bs)


def default_weights():
    """LI/_finite_difference_interpolator.py:102-119."""
    return {"dxy": 1.0, "dyz": 1.0, "dxz": 1.0, "dxx": 1.0, "dyy": 1.0, "dzz": 1.0,
            "dx": 1.0, "dy": 1.0, "dz": 1.0, "cpw": 1.0, "gpw": 1.0, "npw": 10.0,
            "tpw": 1.0, "ipw": 1.0}


def _stencil_rows(grid: Grid, mask, centres):
    """_assemble_operator (LI/_finite_difference_interpolator.py:643-692) +
    neighbour_global_indexes (LC/supports/_3d_structured_grid.py:164-293)."""
    active = mask.flatten() != 0
    off = _FULL[:, active]
    vals = mask.flatten()[active]
    nx, ny = grid.nnodes[0], grid.nnodes[1]
    nb = centres[:, None, :] + off.T[None, :, :]
    idc = nb[:, :, 0] + nx * nb[:, :, 1] + nx * ny * nb[:, :, 2]
    a = np.tile(vals, (centres.shape[0], 1))
    return a, idc.astype(np.int64)


def assemble_fd(grid: Grid, data: dict, weights: dict | None = None, regularisation=None,
                scale=None):
    """FiniteDifferenceInterpolator.setup_interpolator, explicit path:
    LI/_finite_difference_interpolator.py:132-224 (family order), data rows
    :236-476, regulariser :643-692, borders :1052-1144.

    ``data`` keys: value (N,5), normal (N,7), gradient (N,7), tangent (N,7),
    interface (N,5).  ``scale`` is the optional per-node regularisation
    scale (use_regularisation_weight_scale).  Returns an LSSystem.
    """
    wts = default_weights()
    if weights:
        wts.update(weights)
    if regularisation is not None:
        for k in ("dxy", "dyz", "dxz", "dxx", "dyy", "dzz"):
            wts[k] = float(regularisation)
    sysm = LSSystem(grid.n_nodes)
    empty = {"value": np.zeros((0, 5)), "normal": np.zeros((0, 7)), "gradient": np.zeros((0, 7)),
             "tangent": np.zeros((0, 7)), "interface": np.zeros((0, 5))}
    d = {**empty, **{k: np.array(v, dtype=float) for k, v in data.items()}}

    # norm_0..2: LI/_finite_difference_interpolator.py:403-476
    pts = d["normal"]
    if pts.shape[0] > 0:
        idc, ins = cell_corners(grid, pts[:, :3])
        T = element_gradient(grid, pts[ins, :3])
        pw = np.ones(int(np.sum(ins)))
        if pts.shape[1] > 6:
            pw = pts[ins, 6]
        w = wts["npw"] / 3
        cw = float(w) * pw
        for k in range(3):
            sysm.add(T[:, k, :], pts[ins, 3 + k], idc[ins, :], w=cw, name=f"norm_{k}")

    # gradient (strike, dip): :342-401
    pts = d["gradient"]
    if pts.shape[0] > 0:
        idc, ins = cell_corners(grid, pts[:, :3])
        T = element_gradient(grid, pts[ins, :3])
        norm = np.linalg.norm(pts[:, 3:6], axis=1)
        pts[:, 3:6] /= norm[:, None]
        T /= norm[ins, None, None]
        sv, dv = get_vectors(pts[ins, 3:6])
        A = np.einsum("ij,ijk->ik", sv.T, T)
        B = np.zeros(int(np.sum(ins)))
        sysm.add(A, B, idc[ins, :], w=wts["gpw"], name="gradient")
        A = np.einsum("ij,ijk->ik", dv.T, T)
        sysm.add(A, B, idc[ins, :], w=wts["gpw"], name="gradient")

    # value: :236-269
    pts = d["value"]
    if pts.shape[0] > 0:
        idc, ins = cell_corners(grid, pts[:, :3])
        a = trilinear(local_coords(grid, pts[ins, :3]))
        sysm.add(a, pts[ins, 3], idc[ins, :], w=wts["cpw"] * pts[ins, 4], name="value")

    # tangent = gradient orthogonal: LI/_discrete_interpolator.py:668-682 -> :478-539
    pts = d["tangent"]
    if pts.shape[0] > 0:
        vectors = pts[:, 3:6]
        idc, ins = cell_corners(grid, pts[:, :3])
        norm = np.linalg.norm(vectors, axis=1)
        vectors[norm > 0, :] /= norm[norm > 0, None]
        T = element_gradient(grid, pts[ins, :3])
        ni = norm[ins]
        T[ni > 0] /= ni[ni > 0, None, None]
        A = np.einsum("ij,ijk->ik", vectors[ins, :3], T)
        sysm.add(A, np.zeros(int(np.sum(ins))), idc[ins, :], w=wts["tpw"],
                 name="gradient orthogonal")

    # interface: :297-340 with _interface_pair_indices :271-295 (auto, thr 200, star)
    pts = d["interface"]
    if pts.shape[0] > 1:
        idc, ins = cell_corners(grid, pts[:, :3])
        idc = idc[ins]
        A = trilinear(local_coords(grid, pts[ins, :3]))
        for uid in np.unique(pts[np.logical_and(~np.isnan(pts[:, 3]), ins), 3]):
            m = pts[ins, 3] == uid
            n = int(np.sum(m))
            if n < 2:
                continue
            if n > 200:
                pair = np.column_stack([np.zeros(n - 1, dtype=int), np.arange(1, n)])
            else:
                ii, jj = np.triu_indices(n, k=1)
                pair = np.column_stack([ii, jj])
            IA = np.hstack([A[m][pair[:, 0]], -A[m][pair[:, 1]]])
            II = np.hstack([idc[m][pair[:, 0]], idc[m][pair[:, 1]]])
            sysm.add(IA, np.zeros(IA.shape[0]), II, w=wts["ipw"], name=f"interface_{uid}")

    # interior regulariser families in get_operators order:
    # LC/supports/_3d_structured_grid.py:489-510
    ops = [("dxy", DXY, wts["dxy"] / 4), ("dyz", DYZ, wts["dyz"] / 4),
           ("dxz", DXZ, wts["dxz"] / 4), ("dxx", DXX, wts["dxx"] / 1),
           ("dyy", DYY, wts["dyy"] / 1), ("dzz", DZZ, wts["dzz"] / 1)]
    gidx = grid.node_index(np.arange(grid.n_nodes))
    n = grid.nnodes
    emask = np.all((gidx > 0) & (gidx < n[None, :] - 1), axis=1)
    centres = gidx[emask]
    cid = np.nonzero(emask)[0]
    for name, mask, w in ops:
        if centres.shape[0] == 0:
            continue
        a, idc = _stencil_rows(grid, mask, centres)
        rw = scale[cid] * w if scale is not None else w
        sysm.add(a, np.zeros(a.shape[0]), idc, w=rw, name=name)

    # borders: LI/_finite_difference_interpolator.py:1052-1144, :636-641
    borders = [(DXX, wts["dx"], "dx_lower", 0, False), (DXX, wts["dx"], "dx_upper", 0, True),
               (DYY, wts["dy"], "dy_lower", 1, False), (DYY, wts["dy"], "dy_upper", 1, True),
               (DZZ, wts["dz"], "dz_lower", 2, False), (DZZ, wts["dz"], "dz_upper", 2, True)]
    for mask, w, name, axis, upper in borders:
        if n[axis] < 3 or w == 0:
            continue
        b = n[axis] - 2 if upper else 1
        sel = gidx[:, axis] == b
        if not np.any(sel):
            continue
        a, idc = _stencil_rows(grid, mask, gidx[sel])
        rw = scale[np.nonzero(sel)[0]] * w if scale is not None else w
        sysm.add(a, np.zeros(a.shape[0]), idc, w=rw, name=name)
    return sysm


# ---------------------------------------------------------------- A10-A13 ----
def preprocess(A, b, ridge=1e-8, add_ridge=True, scaling=True):
    """LI/_solver_pipeline.py:86-124 and compute_column_scaling_matrix
    (LI/_discrete_interpolator.py:815-833)."""
    A = sparse.csr_matrix(A)
    if add_ridge:
        A = sparse.vstack([A, sparse.eye(A.shape[1]) * ridge]).tocsr()
        b = np.hstack([b, np.zeros(A.shape[1])])
    S = None
    if scaling:
        col = spla.norm(A, axis=0)
        sf = np.ones(A.shape[1])
        m = col > 0
        sf[m] = 1.0 / col[m]
        S = sparse.diags(sf)
        A = (A @ S).tocsr()
    return A, b, S


def normal_equations(sysm: LSSystem, ridge=1e-8, add_ridge=True):
    """Unscaled N = A^T A (ridge included) and rhs = A^T b."""
    A, b = sysm.build_matrix()
    A, b, _ = preprocess(A, b, ridge=ridge, add_ridge=add_ridge, scaling=False)
    return (A.T @ A).tocsr(), A.T @ b


def solve_cg(sysm: LSSystem, rtol=1e-5, atol=0.0, maxiter=None, ridge=1e-8, x0=None,
             return_info=False):
    """solve_system('cg') = _solver_strategy.solve_with_cg (LI/_solver_strategy.py:48-71)
    on the preprocessed system, then c = S y (LI/_discrete_interpolator.py:1312-1313)."""
    A, b = sysm.build_matrix()
    A, b, S = preprocess(A, b, ridge=ridge)
    kw = {"rtol": rtol, "atol": atol}
    if maxiter is not None:
        kw["maxiter"] = maxiter
    if x0 is not None:
        kw["x0"] = x0
    it = [0]

    def cb(_):
        it[0] += 1

    y, info = spla.cg(A.T @ A, A.T @ b, callback=cb, **kw)
    c = S @ y
    if return_info:
        return c, info, it[0]
    return c


def solve_lsmr(sysm: LSSystem, tol=None, ridge=1e-8, **kw):
    """solve_with_lsmr, LI/_solver_strategy.py:108-136."""
    A, b = sysm.build_matrix()
    A, b, S = preprocess(A, b, ridge=ridge)
    if tol is not None and "btol" not in kw:
        kw["btol"] = tol
        kw["atol"] = 0.0
    y = spla.lsmr(A, b, **kw)[0]
    return S @ y


def solve_exact(sysm: LSSystem, ridge=1e-8):
    """Unique LS minimiser of the ridge-regularised system via spsolve on the
    normal equations (library direct solve)."""
    N, rhs = normal_equations(sysm, ridge=ridge)
    return spla.spsolve(N.tocsc(), rhs)


def jacobi_pcg(N, rhs, rtol=1e-10, atol=0.0, maxiter=None, x0=None):
    """Jacobi-preconditioned CG on N c = rhs with M = diag(N)^-1.

    Equal in exact arithmetic to scipy's cg on the column-scaled system
    (S N S) y = S rhs, c = S y, S = diag(N)^-1/2, including the stopping test
    |S r| < max(atol, rtol |S rhs|).  Used to pin the GPU solver's iterate
    sequence.  Returns (c, iterations, converged)."""
    d = N.diagonal()
    minv = np.where(d > 0, 1.0 / np.where(d > 0, d, 1), 1.0)
    x = np.zeros_like(rhs) if x0 is None else np.array(x0, dtype=float)
    r = rhs - N @ x
    z = minv * r
    rho = float(r @ z)
    bn = float(np.sqrt(rhs @ (minv * rhs)))
    thr = max(atol, rtol * bn)
    if maxiter is None:
        maxiter = 10 * rhs.shape[0]
    p = None
    rho_prev = None
    for it in range(maxiter):
        if np.sqrt(rho) < thr:
            return x, it, True
        if it == 0:
            p = z.copy()
        else:
            p = z + (rho / rho_prev) * p
        q = N @ p
        alpha = rho / float(p @ q)
        x = x + alpha * p
        r = r - alpha * q
        rho_prev = rho
        z = minv * r
        rho = float(r @ z)
    return x, maxiter, False


# --------------------------------------------------------- synthetic data ----
def field_c1(p):
    """f = z + 0.1 sin(2 pi x) cos(2 pi y) and its gradient."""
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    f = z + 0.1 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    g = np.stack([0.2 * np.pi * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y),
                  -0.2 * np.pi * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y),
                  np.ones_like(x)], axis=1)
    return f, g


def compute_weighting(grid_points, constraint_points, alpha=10.0, sigma=1.0):
    """LI/_finite_difference_interpolator.py:66-88 (scipy KDTree)."""
    from scipy.spatial import KDTree

    distances, _ = KDTree(constraint_points).query(grid_points, k=1)
    return 1 + alpha * np.exp(-(distances ** 2) / (2 * sigma ** 2))


def regularisation_scale(grid: Grid, normal_points, sigma=None):
    """reset() ones + one compute_weighting term for the inside normal points
    (LI/_discrete_interpolator.py:313, LI/_finite_difference_interpolator.py:436-444)."""
    s = np.ones(grid.n_nodes)
    if normal_points is not None and len(normal_points):
        ins = inside(grid, normal_points[:, :3])
        if sigma is None:
            sigma = grid.nnodes[0] * 10
        s = s + compute_weighting(grid.nodes(), normal_points[ins, :3], sigma=sigma)
    return s
