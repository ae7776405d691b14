"""Device-side finite-difference least-squares system (thin owner of the C handle).

``FdRows`` is one family of least-squares rows held in HBM (CSR with int32
columns, normalised FP64 coefficients, normalised rhs, weights and the
pre-normalisation row norms).  ``FdSystem`` owns an ``lsamd_fd_system``
handle: the matrix-free normal operator of the regulariser plus the data rows,
its Jacobi diagonal, rhs and the PCG driver.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

F64 = torch.float64
# stencil forms of the autotuner, in the order of lsamd_fd_system_stencil_info
STENCIL_FORMS = (8, 16, 32, 1008, 1016, 1032, 2008, 2016, 2032)


@dataclass
class FdRows:
    """A block of least-squares rows on the device (one reference family).

    Mirrors ``constraints[name] = {matrix, b, w}``
    (LI/_discrete_interpolator.py:383-389); coefficients and ``b`` are
    already row-normalised (:361-366) and ``w`` is not yet applied.
    """

    row_ptr: torch.Tensor  # (R+1,) int64
    cols: torch.Tensor     # (nnz,) int32
    coef: torch.Tensor     # (nnz,) float64
    b: torch.Tensor        # (R,) float64
    w: torch.Tensor        # (R,) float64
    norm: torch.Tensor     # (R,) float64, |a| before normalisation
    dof: int

    @property
    def nrows(self) -> int:
        return int(self.b.shape[0])

    @property
    def shape(self):
        return (self.nrows, self.dof)

    @property
    def nnz(self) -> int:
        return int(self.cols.shape[0])

    def to_scipy(self):
        """Materialise the normalised rows as a scipy CSR matrix (inspection only)."""
        from scipy import sparse

        return sparse.csr_matrix((self.coef.cpu().numpy(), self.cols.cpu().numpy().astype(np.int64),
                                  self.row_ptr.cpu().numpy()), shape=self.shape)


def fixed_width_rows(cols, coef, b, w, norm, dof) -> FdRows:
    """Wrap (R,k) column/coefficient tensors as an FdRows block."""
    R, k = cols.shape
    dev = cols.device
    row_ptr = torch.arange(0, (R + 1) * k, k, dtype=torch.int64, device=dev)
    return FdRows(row_ptr, cols.reshape(-1).contiguous(), coef.reshape(-1).contiguous(),
                  b.contiguous(), w.contiguous(), norm.contiguous(), dof)


def concat_rows(blocks, dof) -> FdRows | None:
    blocks = [b for b in blocks if b is not None and b.nrows > 0]
    if not blocks:
        return None
    ptrs, off = [], 0
    for i, b in enumerate(blocks):
        p = b.row_ptr[:-1] + off if i < len(blocks) - 1 else b.row_ptr + off
        ptrs.append(p)
        off += b.nnz
    return FdRows(torch.cat(ptrs), torch.cat([b.cols for b in blocks]),
                  torch.cat([b.coef for b in blocks]), torch.cat([b.b for b in blocks]),
                  torch.cat([b.w for b in blocks]), torch.cat([b.norm for b in blocks]), dof)


def reg_struct(w: dict, nnodes, ridge: float | None) -> _lib.FdReg:
    """Squared regulariser weights (LC/supports/_3d_structured_grid.py:489-510:
    cross families use dxy/4; LI/_finite_difference_interpolator.py:1141-1143:
    border families skipped when their weight is 0 or the axis has < 3 nodes)."""
    r = _lib.FdReg()
    r.w2_xx = float(w["dxx"]) ** 2
    r.w2_yy = float(w["dyy"]) ** 2
    r.w2_zz = float(w["dzz"]) ** 2
    r.w2_xy = (float(w["dxy"]) / 4) ** 2
    r.w2_xz = (float(w["dxz"]) / 4) ** 2
    r.w2_yz = (float(w["dyz"]) / 4) ** 2
    for ax, key, name in ((0, "dx", "w2_bx"), (1, "dy", "w2_by"), (2, "dz", "w2_bz")):
        bw = float(w.get(key, 0.0))
        setattr(r, name, bw * bw if (bw != 0 and int(nnodes[ax]) >= 3) else 0.0)
    r.ridge2 = float(ridge) ** 2 if ridge else 0.0
    return r


class FdSystem:
    """Owner of one ``lsamd_fd_system`` handle."""

    def __init__(self, grid: _lib.Grid, reg: _lib.FdReg, scale2: torch.Tensor | None = None):
        self.dev = _lib.device()
        self.grid = grid
        self.reg = reg
        self.n = int(grid.nnodes[0] * grid.nnodes[1] * grid.nnodes[2])
        self._scale2 = None
        if scale2 is not None:
            self._scale2 = scale2.to(self.dev, F64).contiguous()
        h = C.c_void_p()
        _lib.call("lsamd_fd_system_create", C.byref(grid), C.byref(reg), _lib.ptr(self._scale2),
                  C.byref(h))
        self._h = h
        self.rows: FdRows | None = None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().lsamd_fd_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_rows(self, rows: FdRows | None):
        self.rows = rows
        s = _lib.stream()
        if rows is None or rows.nrows == 0:
            _lib.call("lsamd_fd_system_set_rows", self._h, 0, 0, None, None, None, None, None, s)
            return
        _lib.call("lsamd_fd_system_set_rows", self._h, rows.nrows, rows.nnz,
                  _lib.ptr(rows.row_ptr), _lib.ptr(rows.cols), _lib.ptr(rows.coef),
                  _lib.ptr(rows.b), _lib.ptr(rows.w), s)

    def apply(self, x: torch.Tensor) -> torch.Tensor:
        prec = _lib.FP32 if x.dtype == torch.float32 else _lib.FP64
        x = x.contiguous()
        y = torch.empty_like(x)
        _lib.call("lsamd_fd_system_apply", self._h, prec, _lib.ptr(x), _lib.ptr(y), _lib.stream())
        return y

    def rhs(self) -> torch.Tensor:
        out = torch.empty(self.n, dtype=F64, device=self.dev)
        _lib.call("lsamd_fd_system_rhs", self._h, _lib.ptr(out), _lib.stream())
        return out

    def diag(self) -> torch.Tensor:
        out = torch.empty(self.n, dtype=F64, device=self.dev)
        _lib.call("lsamd_fd_system_diag", self._h, _lib.ptr(out), _lib.stream())
        return out

    def rhs_for(self, b: torch.Tensor) -> torch.Tensor:
        """A^T W^2 b for other observed values b of the system's rows (one
        value per row, the rows' own coefficients and weights)."""
        nrows = self.rows.nrows if self.rows is not None else 0
        if b.numel() != nrows:
            raise ValueError(f"rhs_for: {b.numel()} values for {nrows} rows")
        out = torch.empty(self.n, dtype=F64, device=self.dev)
        b = b.to(self.dev, F64).contiguous()
        _lib.call("lsamd_fd_system_rhs_rows", self._h, _lib.ptr(b), _lib.ptr(out), _lib.stream())
        return out

    def solve_multi(self, rhs: torch.Tensor, rtol=1e-5, atol=0.0, maxiter=None, check_every=8):
        """Multigrid solves of N c_r = rhs[r] for every row of rhs (R, n) at
        once: the V-cycle and outer-iteration launches serve up to 64 right-hand
        sides each.  Same stopping test per right-hand side as solve()."""
        rhs = rhs.to(self.dev, F64).contiguous()
        if rhs.dim() != 2 or rhs.shape[1] != self.n:
            raise ValueError(f"solve_multi: rhs must be (R, {self.n})")
        nr = rhs.shape[0]
        o = _lib.SolveOpts()
        o.rtol, o.atol = float(rtol), float(atol)
        o.maxiter = int(maxiter) if maxiter else 0
        o.precision = _lib.FP64
        o.check_every = int(check_every)
        o.jacobi = 2
        infos = (_lib.SolveInfo * max(nr, 1))()
        c = torch.empty((nr, self.n), dtype=F64, device=self.dev)
        _lib.call("lsamd_fd_system_solve_multi", self._h, C.byref(o), int(nr), _lib.ptr(rhs),
                  _lib.ptr(c), infos, _lib.stream())
        return c, [{"iterations": int(i.iterations), "converged": bool(i.converged),
                    "residual": float(i.residual), "threshold": float(i.threshold)}
                   for i in list(infos)[:nr]]

    def solve(self, rtol=1e-5, atol=0.0, maxiter=None, precision="fp64", rhs=None,
              check_every=64, inner_rtol=1e-4, max_outer=50, jacobi=True, multigrid=False):
        """PCG on N c = rhs.  jacobi: the reference's column scaling as the
        preconditioner (False: plain CG).  multigrid: geometric-multigrid
        preconditioned flexible CG (FP32 V-cycle, FP64 outer iteration; FD
        systems only); same stopping test, same minimiser, far fewer
        iterations."""
        o = _lib.SolveOpts()
        o.rtol = float(rtol)
        o.atol = float(atol)
        o.maxiter = int(maxiter) if maxiter else 0
        o.precision = _lib.FP32 if precision == "fp32" else _lib.FP64
        o.check_every = int(check_every)
        o.inner_rtol = float(inner_rtol)
        o.max_outer = int(max_outer)
        o.jacobi = 2 if multigrid else (1 if jacobi else 0)
        info = _lib.SolveInfo()
        c = torch.empty(self.n, dtype=F64, device=self.dev)
        if rhs is not None:
            rhs = rhs.to(self.dev, F64).contiguous()
        _lib.call("lsamd_fd_system_solve", self._h, C.byref(o), _lib.ptr(rhs), _lib.ptr(c),
                  C.byref(info), _lib.stream())
        return c, {"iterations": int(info.iterations), "outer": int(info.outer),
                   "converged": bool(info.converged), "residual": float(info.residual),
                   "threshold": float(info.threshold)}

    def lsmr(self, atol=1e-6, btol=1e-6, conlim=1e8, maxiter=None, check_every=64,
             scaling=True):
        """LSMR on the rectangular least-squares system (ridge rows and the
        column scaling included), as scipy.sparse.linalg.lsmr with the same
        atol/btol/conlim/maxiter meaning; returns c = S y and scipy's
        (istop, itn, normr, normar, normA, condA, normx) as a dict."""
        o = _lib.LsmrOpts()
        o.atol, o.btol = float(atol), float(btol)
        o.conlim = float(conlim) if conlim else 0.0
        o.maxiter = int(maxiter) if maxiter else 0
        o.check_every = int(check_every)
        o.scaling = 1 if scaling else 0
        info = _lib.LsmrInfo()
        c = torch.empty(self.n, dtype=F64, device=self.dev)
        _lib.call("lsamd_fd_system_lsmr", self._h, C.byref(o), _lib.ptr(c), C.byref(info),
                  _lib.stream())
        return c, {"istop": int(info.istop), "iterations": int(info.iterations),
                   "normr": float(info.normr), "normar": float(info.normar),
                   "normA": float(info.normA), "condA": float(info.condA),
                   "normx": float(info.normx)}

    def time_stencil(self, precision="fp64", iters=20) -> float:
        """Mean duration (ms) of the stencil kernel alone, HIP events on the
        current stream."""
        ms = C.c_double(0.0)
        _lib.call("lsamd_fd_system_time_stencil", self._h,
                  _lib.FP32 if precision == "fp32" else _lib.FP64, int(iters), C.byref(ms),
                  _lib.stream())
        return float(ms.value)

    def evaluate_adjoint(self, pts: torch.Tensor, gv: torch.Tensor) -> torch.Tensor:
        """A20: E^T gv on this handle's grid (E the trilinear evaluation matrix
        of pts), bit-identical to lsamd_fd_evaluate_adjoint; the sort buffers
        stay cached on the handle, so repeated calls do not allocate."""
        pts = pts.to(self.dev, F64).contiguous()
        gv = gv.to(self.dev, F64).contiguous()
        out = torch.empty(self.n, dtype=F64, device=self.dev)
        _lib.call("lsamd_fd_system_evaluate_adjoint", self._h, _lib.ptr(pts), int(pts.shape[0]),
                  _lib.ptr(gv), _lib.ptr(out), _lib.stream())
        return out

    def time_multi(self, kernel: str, iters=20):
        """Average launch time (ms) and algorithmic bytes of one fine-level
        kernel of the multi-rhs multigrid ('cheb', 'update' or 'stencil') on
        the widest lane group built (64, else 32), after a solve_multi with
        such a group; the group's work is left undefined."""
        k = {"cheb": 0, "update": 1, "stencil": 2}[kernel]
        ms, nb = C.c_double(0.0), C.c_double(0.0)
        _lib.call("lsamd_fd_system_time_multi", self._h, k, int(iters), C.byref(ms),
                  C.byref(nb), _lib.stream())
        return float(ms.value), float(nb.value)

    def stencil_info(self, precision="fp64"):
        """Autotuned stencil form (z-chunk 8/16/32 on 64 x 8 tiles, 1008/1016/
        1032 for those with LDS-DMA staged planes, 2008/2016/2032 on 32 x 16
        tiles; 0: per-node kernel) and the timings of every form (us, -1 =
        skipped)."""
        zc = C.c_int32(0)
        us = (C.c_float * 9)()
        _lib.call("lsamd_fd_system_stencil_info", self._h,
                  _lib.FP32 if precision == "fp32" else _lib.FP64, C.byref(zc), us)
        return {"zc": int(zc.value),
                "candidates_us": dict(zip(STENCIL_FORMS, list(us)))}

    def adjoint_rows(self, c: torch.Tensor, lam: torch.Tensor):
        R = self.rows.nrows if self.rows is not None else 0
        gb = torch.empty(R, dtype=F64, device=self.dev)
        gw = torch.empty(R, dtype=F64, device=self.dev)
        if R:
            _lib.call("lsamd_rows_adjoint", self._h, _lib.ptr(c.contiguous()),
                      _lib.ptr(lam.contiguous()), _lib.ptr(gb), _lib.ptr(gw), _lib.stream())
        return gb, gw


# ------------------------------------------------------- kernel wrappers ----
def locate(grid: _lib.Grid, pts: torch.Tensor, want_w=True, want_T=False):
    """A2-A4 on the device: returns (idc int64 (n,8), inside bool, w (n,8), T (n,3,8))."""
    pts = pts.to(_lib.device(), F64).contiguous()
    n = pts.shape[0]
    idc = torch.empty((n, 8), dtype=torch.int64, device=pts.device)
    ins = torch.empty(n, dtype=torch.uint8, device=pts.device)
    w = torch.empty((n, 8), dtype=F64, device=pts.device) if want_w else None
    T = torch.empty((n, 3, 8), dtype=F64, device=pts.device) if want_T else None
    _lib.call("lsamd_fd_locate", C.byref(grid), _lib.ptr(pts), n, _lib.ptr(idc), _lib.ptr(w),
              _lib.ptr(T), _lib.ptr(ins), _lib.stream())
    return idc, ins.bool(), w, T


def build_rows(grid: _lib.Grid, kind: int, pts: torch.Tensor, weight_col: int, wscale: float,
               dof: int) -> FdRows:
    """A5-A7 rows for points already known to be inside.  ``pts`` is updated in
    place for GRADIENT/TANGENT (normalised vectors), as in the reference."""
    n, ld = pts.shape
    per = {_lib.ROWS_VALUE: 1, _lib.ROWS_NORMAL: 3, _lib.ROWS_GRADIENT: 2,
           _lib.ROWS_TANGENT: 1}[kind]
    R = per * n
    dev = pts.device
    cols = torch.empty((R, 8), dtype=torch.int32, device=dev)
    coef = torch.empty((R, 8), dtype=F64, device=dev)
    b = torch.empty(R, dtype=F64, device=dev)
    w = torch.empty(R, dtype=F64, device=dev)
    norm = torch.empty(R, dtype=F64, device=dev)
    if n:
        _lib.call("lsamd_fd_build_rows", C.byref(grid), kind, _lib.ptr(pts), n, ld, weight_col,
                  float(wscale), _lib.ptr(cols), _lib.ptr(coef), _lib.ptr(b), _lib.ptr(w),
                  _lib.ptr(norm), _lib.stream())
    return fixed_width_rows(cols, coef, b, w, norm, dof)


def evaluate(grid: _lib.Grid, c: torch.Tensor, pts: torch.Tensor, values=True, gradients=False):
    pts = pts.to(_lib.device(), F64).contiguous()
    c = c.to(pts.device, F64).contiguous()
    n = pts.shape[0]
    v = torch.empty(n, dtype=F64, device=pts.device) if values else None
    g = torch.empty((n, 3), dtype=F64, device=pts.device) if gradients else None
    _lib.call("lsamd_fd_evaluate", C.byref(grid), _lib.ptr(c), _lib.ptr(pts), n, _lib.ptr(v),
              _lib.ptr(g), _lib.stream())
    return v, g


# ------------------------------------------------------------------- P1 ----
class P1System(FdSystem):
    """``lsamd_fd_system`` handle whose regulariser is the P1 edge-jump
    operator cgw^2 E^T E (LI/_p1interpolator.py:115-166), applied by the
    21-tap table stencil; data rows, solve, apply and adjoint as FdSystem."""

    def __init__(self, grid: _lib.Grid, cgw: float, ridge: float | None):
        self.dev = _lib.device()
        self.grid = grid
        self.reg = None
        self.n = int(grid.nnodes[0] * grid.nnodes[1] * grid.nnodes[2])
        self._scale2 = None
        h = C.c_void_p()
        _lib.call("lsamd_p1_system_create", C.byref(grid), float(cgw),
                  float(ridge) ** 2 if ridge else 0.0, C.byref(h))
        self._h = h
        self.rows = None


def p1_locate(grid: _lib.Grid, pts: torch.Tensor, want_bc=True, want_grad=False):
    """A16: (tet int64 (n,), nodes (n,4), inside bool, bc (n,4), grad (n,3,4))."""
    pts = pts.to(_lib.device(), F64).contiguous()
    n = pts.shape[0]
    d = pts.device
    tet = torch.empty(n, dtype=torch.int64, device=d)
    nodes = torch.empty((n, 4), dtype=torch.int64, device=d)
    ins = torch.empty(n, dtype=torch.uint8, device=d)
    bc = torch.empty((n, 4), dtype=F64, device=d) if want_bc else None
    gr = torch.empty((n, 3, 4), dtype=F64, device=d) if want_grad else None
    _lib.call("lsamd_p1_locate", C.byref(grid), _lib.ptr(pts), n, _lib.ptr(tet), _lib.ptr(nodes),
              _lib.ptr(bc), _lib.ptr(gr), _lib.ptr(ins), _lib.stream())
    return tet, nodes, ins.bool(), bc, gr


def p1_elements(grid: _lib.Grid, want_neighbours=False):
    ntet = 5 * int((grid.nnodes[0] - 1) * (grid.nnodes[1] - 1) * (grid.nnodes[2] - 1))
    dev = _lib.device()
    el = torch.empty((ntet, 4), dtype=torch.int64, device=dev)
    nb = torch.empty((ntet, 4), dtype=torch.int64, device=dev) if want_neighbours else None
    _lib.call("lsamd_p1_elements", C.byref(grid), _lib.ptr(el), _lib.ptr(nb), _lib.stream())
    return el, nb


def p1_tet_gradients(grid: _lib.Grid, tets: torch.Tensor, want_size=False):
    tets = tets.to(_lib.device(), torch.int64).contiguous()
    n = tets.shape[0]
    g = torch.empty((n, 3, 4), dtype=F64, device=tets.device)
    s = torch.empty(n, dtype=F64, device=tets.device) if want_size else None
    if n:
        _lib.call("lsamd_p1_tet_gradients", C.byref(grid), _lib.ptr(tets), n, _lib.ptr(g),
                  _lib.ptr(s), _lib.stream())
    return g, s


def p1_n_pairs(grid: _lib.Grid) -> int:
    n = int(_lib.load().lsamd_p1_n_pairs(C.byref(grid)))
    if n < 0:
        _lib.check(_lib.LSAMD_ERR_ARG)
    return n


def p1_pairs(grid: _lib.Grid, want_faces=False):
    """A15 face table: pairs (P,2) (larger tet first) and faces (P,3)."""
    P = p1_n_pairs(grid)
    dev = _lib.device()
    pairs = torch.empty((P, 2), dtype=torch.int64, device=dev)
    faces = torch.empty((P, 3), dtype=torch.int64, device=dev) if want_faces else None
    if P:
        _lib.call("lsamd_p1_pairs", C.byref(grid), _lib.ptr(pairs), _lib.ptr(faces), None, None,
                  None, None, None, 0.0, _lib.stream())
    return pairs, faces


def p1_edge_rows(grid: _lib.Grid, cgw: float, dof: int, scale: torch.Tensor | None = None) -> FdRows:
    """A17 edge-jump rows (8 entries each, normalised, b = 0)."""
    P = p1_n_pairs(grid)
    dev = _lib.device()
    cols = torch.empty((P, 8), dtype=torch.int32, device=dev)
    coef = torch.empty((P, 8), dtype=F64, device=dev)
    w = torch.empty(P, dtype=F64, device=dev)
    norm = torch.empty(P, dtype=F64, device=dev)
    if scale is not None:
        scale = scale.to(dev, F64).contiguous()
    if P:
        _lib.call("lsamd_p1_pairs", C.byref(grid), None, None, _lib.ptr(cols), _lib.ptr(coef),
                  _lib.ptr(w), _lib.ptr(norm), _lib.ptr(scale), float(cgw), _lib.stream())
    b = torch.zeros(P, dtype=F64, device=dev)
    return fixed_width_rows(cols, coef, b, w, norm, dof)


def p1_build_rows(grid: _lib.Grid, kind: int, pts: torch.Tensor, wscale: float, dof: int,
                  bval: float = 0.0) -> FdRows:
    """A17 rows (4 entries) for points already known to be inside."""
    pts = pts.to(_lib.device(), F64).contiguous()
    n, ld = pts.shape
    per = 3 if kind == _lib.ROWS_NORMAL else 1
    R = per * n
    dev = pts.device
    cols = torch.empty((R, 4), dtype=torch.int32, device=dev)
    coef = torch.empty((R, 4), dtype=F64, device=dev)
    b = torch.empty(R, dtype=F64, device=dev)
    w = torch.empty(R, dtype=F64, device=dev)
    norm = torch.empty(R, dtype=F64, device=dev)
    if n:
        _lib.call("lsamd_p1_build_rows", C.byref(grid), kind, _lib.ptr(pts), n, ld, float(wscale),
                  float(bval), _lib.ptr(cols), _lib.ptr(coef), _lib.ptr(b), _lib.ptr(w),
                  _lib.ptr(norm), _lib.stream())
    return fixed_width_rows(cols, coef, b, w, norm, dof)


def p1_evaluate(grid: _lib.Grid, c: torch.Tensor, pts: torch.Tensor, values=True, gradients=False):
    pts = pts.to(_lib.device(), F64).contiguous()
    c = c.to(pts.device, F64).contiguous()
    n = pts.shape[0]
    v = torch.empty(n, dtype=F64, device=pts.device) if values else None
    g = torch.empty((n, 3), dtype=F64, device=pts.device) if gradients else None
    if n:
        _lib.call("lsamd_p1_evaluate", C.byref(grid), _lib.ptr(c), _lib.ptr(pts), n, _lib.ptr(v),
                  _lib.ptr(g), _lib.stream())
    return v, g


# ----------------------------------------------------------------- batch ----
class BatchSystem:
    """An ensemble of FD systems of different grid sizes solved together
    (``lsamd_batch_*``): node vectors concatenated per system, one shared
    regulariser class table, per-system PCG scalars and stopping tests."""

    def __init__(self, grids, reg: _lib.FdReg):
        self.dev = _lib.device()
        self.grids = list(grids)
        self.nsys = len(self.grids)
        arr = (_lib.Grid * self.nsys)(*self.grids)
        h = C.c_void_p()
        _lib.call("lsamd_batch_create", self.nsys, arr, C.byref(reg), C.byref(h))
        self._h = h
        off = (C.c_int64 * (self.nsys + 1))()
        _lib.call("lsamd_batch_offsets", self._h, off)
        self.offsets = [int(v) for v in off]
        self.n = self.offsets[-1]
        self.rows = None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().lsamd_fd_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_rows(self, blocks):
        """``blocks[s]``: the FdRows of system s (local node ids)."""
        shifted, sys_rows = [], [0]
        for s, blk in enumerate(blocks):
            if blk is not None and blk.nrows:
                shifted.append(FdRows(blk.row_ptr, blk.cols + self.offsets[s], blk.coef, blk.b,
                                      blk.w, blk.norm, self.n))
            sys_rows.append(sys_rows[-1] + (blk.nrows if blk is not None else 0))
        rows = concat_rows(shifted, self.n)
        self.rows = rows
        sr = (C.c_int64 * (self.nsys + 1))(*sys_rows)
        s = _lib.stream()
        if rows is None:
            _lib.call("lsamd_batch_set_rows", self._h, sr, 0, 0, None, None, None, None, None, s)
            return
        _lib.call("lsamd_batch_set_rows", self._h, sr, rows.nrows, rows.nnz, _lib.ptr(rows.row_ptr),
                  _lib.ptr(rows.cols), _lib.ptr(rows.coef), _lib.ptr(rows.b), _lib.ptr(rows.w), s)

    def apply(self, x: torch.Tensor) -> torch.Tensor:
        prec = _lib.FP32 if x.dtype == torch.float32 else _lib.FP64
        x = x.contiguous()
        y = torch.empty_like(x)
        _lib.call("lsamd_fd_system_apply", self._h, prec, _lib.ptr(x), _lib.ptr(y), _lib.stream())
        return y

    def rhs(self) -> torch.Tensor:
        out = torch.empty(self.n, dtype=F64, device=self.dev)
        _lib.call("lsamd_fd_system_rhs", self._h, _lib.ptr(out), _lib.stream())
        return out

    def diag(self) -> torch.Tensor:
        out = torch.empty(self.n, dtype=F64, device=self.dev)
        _lib.call("lsamd_fd_system_diag", self._h, _lib.ptr(out), _lib.stream())
        return out

    def solve(self, rtol=1e-5, atol=0.0, maxiter=None, precision="fp64", rhs=None,
              check_every=64, inner_rtol=1e-4, max_outer=50):
        o = _lib.SolveOpts()
        o.rtol, o.atol = float(rtol), float(atol)
        o.maxiter = int(maxiter) if maxiter else 0
        o.precision = _lib.FP32 if precision == "fp32" else _lib.FP64
        o.check_every = int(check_every)
        o.inner_rtol = float(inner_rtol)
        o.max_outer = int(max_outer)
        o.jacobi = 1
        infos = (_lib.SolveInfo * self.nsys)()
        c = torch.empty(self.n, dtype=F64, device=self.dev)
        if rhs is not None:
            rhs = rhs.to(self.dev, F64).contiguous()
        _lib.call("lsamd_batch_solve", self._h, C.byref(o), _lib.ptr(rhs), _lib.ptr(c), infos,
                  _lib.stream())
        return c, [{"iterations": int(i.iterations), "outer": int(i.outer),
                    "converged": bool(i.converged), "residual": float(i.residual),
                    "threshold": float(i.threshold)} for i in infos]

    def split(self, v: torch.Tensor):
        return [v[self.offsets[s]:self.offsets[s + 1]] for s in range(self.nsys)]

    def time_stencil(self, precision="fp32", iters=20) -> float:
        ms = C.c_double(0.0)
        _lib.call("lsamd_fd_system_time_stencil", self._h,
                  _lib.FP32 if precision == "fp32" else _lib.FP64, int(iters), C.byref(ms),
                  _lib.stream())
        return float(ms.value)


def solve_mg_batch(systems, rtol=1e-5, atol=0.0, maxiter=None, check_every=8):
    """Multigrid-preconditioned solves of independent FdSystems (distinct
    handles, rows set) in one launch sequence per iteration for all of them
    (``lsamd_fd_systems_solve_mg``); each keeps its own hierarchy, stopping
    test and iteration count.  Returns (list of c, list of infos)."""
    n = len(systems)
    o = _lib.SolveOpts()
    o.rtol, o.atol = float(rtol), float(atol)
    o.maxiter = int(maxiter) if maxiter else 0
    o.precision = _lib.FP64
    o.check_every = int(check_every)
    o.jacobi = 2
    dev = _lib.device()
    outs = [torch.empty(sy.n, dtype=F64, device=dev) for sy in systems]
    handles = (C.c_void_p * n)(*[sy._h.value for sy in systems])
    ptrs = (C.c_void_p * n)(*[t.data_ptr() for t in outs])
    infos = (_lib.SolveInfo * n)()
    _lib.call("lsamd_fd_systems_solve_mg", handles, n, C.byref(o), ptrs, infos, _lib.stream())
    return outs, [{"iterations": int(i.iterations), "outer": int(i.outer),
                   "converged": bool(i.converged), "residual": float(i.residual),
                   "threshold": float(i.threshold)} for i in infos]


class BatchedMultigrid:
    """Uneven ensemble solved by batched multigrid: one handle per system
    (its own hierarchy), all systems advanced together by one launch sequence
    per iteration.  ``prepare`` creates the handles and builds the grid-only
    part of every hierarchy; ``solve`` sets each system's rows, brings its
    hierarchy up to date (host threads, one stream each) and runs the
    batched solve."""

    def __init__(self, grids, reg: _lib.FdReg, workers: int = 12):
        self.dev = _lib.device()
        self.grids = list(grids)
        self.reg = reg
        self.workers = max(1, int(workers))
        self.systems = []
        self.last_timing = {}

    def close(self):
        for h in self.systems:
            h.close()
        self.systems = []

    def _threaded(self, fn):
        import threading

        idx = list(range(len(self.systems)))
        lock = threading.Lock()
        errors = []

        def work():
            stream = torch.cuda.Stream(device=self.dev)
            with torch.cuda.stream(stream):
                while True:
                    with lock:
                        if not idx or errors:
                            break
                        s = idx.pop(0)
                    try:
                        fn(s)
                    except Exception as e:  # surfaced on the calling thread
                        with lock:
                            errors.append(e)
                        break
            stream.synchronize()

        threads = [threading.Thread(target=work) for _ in range(min(self.workers, len(idx)))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]

    def prepare(self, blocks):
        """One handle per system, its rows and its full hierarchy."""
        if not self.systems:
            self.systems = [FdSystem(g, self.reg) for g in self.grids]

        def build(s):
            self.systems[s].set_rows(blocks[s])
            _lib.call("lsamd_fd_system_mg_build", self.systems[s]._h, _lib.stream())

        self._threaded(build)
        torch.cuda.synchronize()

    def solve(self, blocks, rtol=1e-5, atol=0.0, maxiter=None):
        """``blocks[s]``: FdRows of system s.  Returns (list of c, list of infos)."""
        t0 = time.perf_counter()
        self.prepare(blocks)
        t1 = time.perf_counter()
        out = solve_mg_batch(self.systems, rtol=rtol, atol=atol, maxiter=maxiter)
        torch.cuda.synchronize()
        self.last_timing = {"rows_and_hierarchies_s": t1 - t0,
                            "batched_solve_s": time.perf_counter() - t1}
        return out


class EnsembleSolver:
    """Multigrid solves of the independent FD systems of an uneven ensemble,
    several at a time.

    Systems are grouped by grid shape; each shape gets one handle, whose
    coarse grids, tuned stencils and level vectors are built once (a set-up
    solve of one iteration) and reused for every system of that shape (only
    the row-dependent parts are rebuilt per system).  Shape groups are spread
    over ``workers`` host threads, largest first, each thread driving its own
    stream, so solves of different shapes overlap on the GPU.  The solves
    are the single-system multigrid solves: same stopping test and minimiser
    as ``FdSystem.solve(multigrid=True)``."""

    def __init__(self, grids, reg: _lib.FdReg, workers: int = 12):
        self.dev = _lib.device()
        self.grids = list(grids)
        self.reg = reg
        self.workers = max(1, int(workers))
        self.shapes = {}
        for s, g in enumerate(self.grids):
            key = (tuple(float(v) for v in g.origin), tuple(float(v) for v in g.step),
                   tuple(int(v) for v in g.nnodes))
            self.shapes.setdefault(key, []).append(s)
        self.handles = {}
        self.last_timing = {}

    def close(self):
        for h in self.handles.values():
            h.close()
        self.handles = {}

    def prepare(self, blocks, rtol=1e-10):
        """Create one handle per shape and build its hierarchy (one-iteration
        solve with the rows of the shape's first system)."""
        for key, members in self.shapes.items():
            if key in self.handles:
                continue
            h = FdSystem(self.grids[members[0]], self.reg)
            h.set_rows(blocks[members[0]])
            h.solve(rtol=rtol, maxiter=1, multigrid=True, check_every=8)
            self.handles[key] = h
        torch.cuda.synchronize()

    def solve(self, blocks, rtol=1e-5, atol=0.0, maxiter=None):
        """``blocks[s]``: FdRows of system s.  Returns (list of c, list of infos)."""
        import threading

        self.prepare(blocks, rtol)
        out = [None] * len(self.grids)
        infos = [None] * len(self.grids)
        order = sorted(self.shapes, key=lambda k: -sum(int(np.prod(self.grids[s].nnodes))
                                                        for s in self.shapes[k]))
        lock = threading.Lock()
        queue = list(order)
        errors = []
        spent = [0.0, 0.0]  # host-thread seconds in set_rows, in solve

        def work():
            stream = torch.cuda.Stream(device=self.dev)
            with torch.cuda.stream(stream):
                while True:
                    with lock:
                        if not queue or errors:
                            break
                        key = queue.pop(0)
                    try:
                        h = self.handles[key]
                        for s in self.shapes[key]:
                            t0 = time.perf_counter()
                            h.set_rows(blocks[s])
                            t1 = time.perf_counter()
                            c, info = h.solve(rtol=rtol, atol=atol, maxiter=maxiter,
                                              multigrid=True, check_every=8)
                            t2 = time.perf_counter()
                            out[s], infos[s] = c, info
                            with lock:
                                spent[0] += t1 - t0
                                spent[1] += t2 - t1
                    except Exception as e:  # surfaced on the calling thread
                        with lock:
                            errors.append(e)
                        break
            stream.synchronize()

        threads = [threading.Thread(target=work) for _ in range(min(self.workers, len(order)))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        torch.cuda.synchronize()
        self.last_timing = {"set_rows_thread_s": spent[0], "solve_thread_s": spent[1]}
        return out, infos
