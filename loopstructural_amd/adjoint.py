"""Differentiable finite-difference interpolation (adjoint through the solve).

``FDProblem`` fixes a grid, the constraint locations and the weights of the
reference FD system; its parameters are the constraint values, per-point
value weights, normal vectors and per-point normal weights.  ``solve`` maps
them to the nodal solution ``c`` of ``N c = rhs`` through
``FDSolveFunction`` (a ``torch.autograd.Function``); ``evaluate`` maps ``c`` to
values at arbitrary points (trilinear, differentiable in ``c``).

Backward uses one more solve with the same SPD operator, ``N lambda = dL/dc``,
then the per-row contraction of ``lsamd_rows_adjoint``: for a row with
normalised coefficients a, normalised rhs b_hat = b/|a| and weight w,
``dL/db = w^2 (lambda.a) / |a|`` and ``dL/dw = 2 w (lambda.a)(b_hat - a.c)``.
Rows follow the reference: value rows weight cpw * w_pt
(LI/_finite_difference_interpolator.py:236-269), normal rows one per axis
with weight npw/3 * w_pt (:403-476).  The adjoint of a linear solve needs no
stored CG iterates: only the converged ``c`` is kept.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import fd_system as fs

F64 = torch.float64


class FDProblem:
    def __init__(self, origin, nsteps_cells, step_vector, value_points, normal_points,
                 regularisation=0.1, cpw=1.0, npw=10.0, border=1.0, ridge=1e-8,
                 precision="fp64", rtol=1e-10, inner_rtol=1e-3):
        self.dev = _lib.device()
        nn = np.asarray(nsteps_cells, dtype=int) + 1
        self.grid = _lib.make_grid(origin, step_vector, nn)
        self.n = int(np.prod(nn))
        w = {k: regularisation for k in ("dxx", "dyy", "dzz", "dxy", "dxz", "dyz")}
        w.update(dx=border, dy=border, dz=border)
        self.system = fs.FdSystem(self.grid, fs.reg_struct(w, nn, ridge))
        self.cpw, self.npw = float(cpw), float(npw)
        self.precision, self.rtol, self.inner_rtol = precision, rtol, inner_rtol
        self.vp = torch.as_tensor(np.asarray(value_points, dtype=float)[:, :3]).to(self.dev)
        self.np_ = torch.as_tensor(np.asarray(normal_points, dtype=float)[:, :3]).to(self.dev)
        _, self.v_in, _, _ = fs.locate(self.grid, self.vp.contiguous(), want_w=False)
        _, self.n_in, _, _ = fs.locate(self.grid, self.np_.contiguous(), want_w=False)
        self._rows_key = None

    def build_rows(self, values, value_w, normals, normal_w) -> fs.FdRows:
        """Normal rows (3 per inside point, axis-major) then value rows."""
        pv = torch.cat([self.vp, values.detach().to(F64)[:, None],
                        value_w.detach().to(F64)[:, None]], 1)[self.v_in].contiguous()
        pn = torch.cat([self.np_, normals.detach().to(F64),
                        normal_w.detach().to(F64)[:, None]], 1)[self.n_in].contiguous()
        bn = fs.build_rows(self.grid, _lib.ROWS_NORMAL, pn, 6, self.npw / 3, self.n)
        bv = fs.build_rows(self.grid, _lib.ROWS_VALUE, pv, 4, self.cpw, self.n)
        return fs.concat_rows([bn, bv], self.n)

    def _solve(self, rhs=None):
        """Forward and adjoint solves: precision 'fp64' (Jacobi PCG), 'fp32'
        (FP32 operator, FP64 refinement) or 'mg' (multigrid-preconditioned
        flexible CG; the hierarchy is built once per set of rows)."""
        if self.precision == "mg":
            return self.system.solve(rtol=self.rtol, rhs=rhs, multigrid=True, check_every=8)
        return self.system.solve(rtol=self.rtol, precision=self.precision, rhs=rhs,
                                 inner_rtol=self.inner_rtol, check_every=128)

    def solve(self, values, value_w, normals, normal_w):
        return FDSolveFunction.apply(self, values, value_w, normals, normal_w)

    def evaluate(self, c, points):
        pts = torch.as_tensor(np.asarray(points, dtype=float)).to(self.dev)
        return FDEvaluateFunction.apply(c, self.system, pts)


class FDSolveFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, problem, values, value_w, normals, normal_w):
        rows = problem.build_rows(values, value_w, normals, normal_w)
        problem.system.set_rows(rows)
        c, info = problem._solve()
        ctx.problem = problem
        ctx.rows = rows
        ctx.info = info
        ctx.save_for_backward(c)
        return c

    @staticmethod
    def backward(ctx, grad_c):
        (c,) = ctx.saved_tensors
        p = ctx.problem
        if p.system.rows is not ctx.rows:  # another forward ran since
            p.system.set_rows(ctx.rows)
        lam, _ = p._solve(rhs=grad_c.detach().to(F64).contiguous())
        gb, gw = p.system.adjoint_rows(c, lam)
        m = int(p.n_in.sum())
        norm = ctx.rows.norm
        gbn = (gb[:3 * m] / norm[:3 * m]).view(3, m)
        gwn = gw[:3 * m].view(3, m)
        gbv = gb[3 * m:] / norm[3 * m:]
        gwv = gw[3 * m:]
        dv = torch.zeros(p.vp.shape[0], dtype=F64, device=p.dev)
        dvw = torch.zeros_like(dv)
        dv[p.v_in] = gbv
        dvw[p.v_in] = p.cpw * gwv
        dn = torch.zeros((p.np_.shape[0], 3), dtype=F64, device=p.dev)
        dnw = torch.zeros(p.np_.shape[0], dtype=F64, device=p.dev)
        dn[p.n_in] = gbn.t()
        dnw[p.n_in] = (p.npw / 3) * gwn.sum(0)
        return None, dv, dvw, dn, dnw


class FDEvaluateFunction(torch.autograd.Function):
    """Trilinear evaluation v = E c (bit-exact kernel) with backward E^T g on
    the problem's system handle (sort buffers cached there)."""

    @staticmethod
    def forward(ctx, c, system, pts):
        v, _ = fs.evaluate(system.grid, c.detach(), pts, True, False)
        ctx.system = system
        ctx.save_for_backward(pts)
        return v

    @staticmethod
    def backward(ctx, gv):
        (pts,) = ctx.saved_tensors
        return ctx.system.evaluate_adjoint(pts, gv.detach()), None, None
