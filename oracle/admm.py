"""CPU restatement of the reference's inequality-constrained solve.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``value_inequality_rows`` / ``pair_inequality_rows``: the inequality rows of
  LI/_discrete_interpolator.py:509-593 (shape functions at the points, no row
  normalisation, no weights; pairs over every (upper, lower) point
  combination).
* ``stack_inequalities``: build_inequality_matrix (:874-884).
* ``admm_solve``: LI/loopsolver/admm_solver.py:59-463 with the projection of
  LI/loopsolver/admm_method.py:4-25: the major loop, rho schedule, adaptive
  rho, active set and the primal/dual stopping test, with the inner step
  taken by scipy ``lsmr`` warm-started from the previous step (the
  reference's default) or solved exactly (``inner='exact'``).
* ``constrained_minimiser``: brute force over active sets (each row free, at
  its lower bound or at its upper bound) of the equality-constrained least
  squares KKT system; the feasible candidate of least objective is the
  minimiser of the convex problem.  It pins ``admm_solve``.
"""
from __future__ import annotations

import itertools

import numpy as np
from scipy import sparse
from scipy.sparse import linalg as spla

from . import fd as ofd


def value_inequality_rows(grid: ofd.Grid, points):
    """Trilinear rows at the inside points, bounds = columns 3:5
    (LI/_discrete_interpolator.py:509-518 on a StructuredGrid)."""
    pts = np.asarray(points, dtype=float)
    idc, ins = ofd.cell_corners(grid, pts[:, :3])
    a = ofd.trilinear(ofd.local_coords(grid, pts[:, :3]))
    return a[ins], idc[ins], pts[ins, 3:5]


def pair_inequality_rows(grid: ofd.Grid, points, lower_bound=-np.inf, upper_bound=1.0,
                         pairs=None):
    """[N(p_i), -N(q_j)] rows for points p of the smaller id and q of the
    larger of each id pair, i fastest (LI/_discrete_interpolator.py:520-593).
    Returns a list of (name, a, idc, bounds)."""
    pts = np.asarray(points, dtype=float)
    if not pairs:
        seen = {}
        for i in np.unique(pts[:, 3]):
            for j in np.unique(pts[:, 3]):
                if i != j and tuple(sorted([i, j])) not in seen:
                    seen[tuple(sorted([i, j]))] = len(seen)
        pairs = list(seen)
    out = []
    for p0, p1 in pairs:
        up = pts[pts[:, 3] == p0]
        lo = pts[pts[:, 3] == p1]
        iu, insu = ofd.cell_corners(grid, up[:, :3])
        il, insl = ofd.cell_corners(grid, lo[:, :3])
        au = ofd.trilinear(ofd.local_coords(grid, up[:, :3]))[insu]
        al = ofd.trilinear(ofd.local_coords(grid, lo[:, :3]))[insl]
        iu, il = iu[insu], il[insl]
        rows = []
        for j in range(al.shape[0]):
            for i in range(au.shape[0]):
                rows.append((np.concatenate([au[i], -al[j]]), np.concatenate([iu[i], il[j]])))
        a = np.array([r[0] for r in rows]).reshape(-1, 16)
        idc = np.array([r[1] for r in rows]).reshape(-1, 16)
        b = np.zeros((a.shape[0], 2))
        b[:, 0] = lower_bound
        b[:, 1] = upper_bound
        out.append((f"inequality_pairs_{p0}_{p1}", a, idc, b))
    return out


def stack_inequalities(families, dof):
    """families: (a (n,k), idc (n,k), bounds (n,2)) -> (Q csr, bounds)."""
    mats, bnds = [], []
    for a, idc, b in families:
        rows = np.tile(np.arange(idc.shape[0]), (a.shape[1], 1)).T
        mats.append(sparse.coo_matrix((a.ravel(), (rows.ravel(), idc.ravel())),
                                      shape=(idc.shape[0], dof)).tocsc())
        bnds.append(np.asarray(b, dtype=float))
    if not mats:
        return sparse.csr_matrix((0, dof)), np.zeros((0, 2))
    return sparse.vstack(mats).tocsr(), np.vstack(bnds)


def _project(z, u, xmin, xmax, x):
    """admm_method.py:13-25."""
    arg = x + u
    inside = np.logical_and(arg >= xmin, arg <= xmax)
    z = z.copy()
    z[inside] = arg[inside]
    below = np.logical_and(arg < xmin, ~inside)
    above = np.logical_and(arg > xmax, ~inside)
    z[below] = xmin[below]
    z[above] = xmax[above]
    u = u + x - z
    return z, u


def admm_solve(A, b, Q, bounds, x0, admm_weight=0.01, nmajor=200, inner="lsmr",
               inner_kwargs=None, admm_abs_tol=1e-4, admm_rel_tol=1e-3, min_iterations=5,
               admm_weight_final=None, admm_weight_schedule="geometric", adaptive_rho=False,
               adaptive_rho_mu=10.0, adaptive_rho_tau=2.0, adaptive_rho_min=1e-4,
               adaptive_rho_max=1e3, model_update_tol=0.0, active_set=False,
               active_set_padding=1e-3, return_history=False):
    """admm_solver.py:59-463 (core path; active set without size limits)."""
    A = sparse.csr_matrix(A)
    Q = sparse.csr_matrix(Q)
    n_ie = Q.shape[0]
    if n_ie < 1:
        raise ValueError("nelements must be greater than 0")
    xmin, xmax = bounds[:, 0], bounds[:, 1]
    z = np.zeros(n_ie)
    u = np.zeros(n_ie)
    model = np.array(x0, dtype=float)
    if admm_weight_final is not None:
        if admm_weight_schedule == "linear":
            rhos = np.linspace(admm_weight, admm_weight_final, nmajor)
        else:
            rhos = np.geomspace(admm_weight, admm_weight_final, nmajor)
    else:
        rhos = np.full(nmajor, float(admm_weight))
    adaptive = adaptive_rho and admm_weight_final is None
    rho = float(rhos[0])
    kw = dict(inner_kwargs or {})
    inner_x0 = None
    active = None
    history = []
    for it in range(nmajor):
        if rhos[it] != rho:
            u *= rho / rhos[it]
            rho = float(rhos[it])
        matrix = sparse.vstack([A, Q * rho]).tocsr()
        z_prev = z.copy()
        Mx = matrix @ model
        rhs = np.zeros(matrix.shape[0])
        rhs[:A.shape[0]] = b - Mx[:A.shape[0]]
        qx = Mx[A.shape[0]:] / rho
        z, u = _project(z, u, xmin, xmax, qx)
        rhs[A.shape[0]:] = -rho * (qx - (z - u))
        solve_matrix, solve_rhs = matrix, rhs
        if active_set:
            viol = np.maximum(xmin - qx, 0.0) + np.maximum(qx - xmax, 0.0)
            dist = np.minimum(np.abs(qx - xmin), np.abs(qx - xmax))
            act = (viol > 0.0) | (dist <= active_set_padding)
            if active is not None:
                act = act | active
            if not act.any():
                act[np.argmax(viol)] = True
            active = act
            solve_matrix = sparse.vstack([A, (Q * rho)[active]]).tocsr()
            solve_rhs = np.concatenate([rhs[:A.shape[0]], rhs[A.shape[0]:][active]])
        if inner == "exact":
            N = (solve_matrix.T @ solve_matrix).tocsc()
            dx = spla.spsolve(N, solve_matrix.T @ solve_rhs)
        else:
            args = dict(kw)
            if inner_x0 is not None:
                args["x0"] = inner_x0
            dx = spla.lsmr(solve_matrix, solve_rhs, **args)[0]
        dx_norm = float(np.linalg.norm(dx))
        model_norm = float(np.linalg.norm(model))
        model = model + dx
        inner_x0 = dx
        primal = float(np.linalg.norm(qx - z))
        dual = float(rho * np.linalg.norm(z - z_prev))
        eps_pri = np.sqrt(n_ie) * admm_abs_tol + admm_rel_tol * max(np.linalg.norm(qx),
                                                                     np.linalg.norm(z))
        eps_dual = np.sqrt(n_ie) * admm_abs_tol + admm_rel_tol * np.linalg.norm(rho * u)
        history.append({"iteration": it + 1, "primal_norm": primal, "dual_norm": dual})
        if it + 1 >= min_iterations and primal <= eps_pri and dual <= eps_dual:
            break
        if model_update_tol > 0 and it + 1 >= min_iterations and \
                dx_norm <= model_update_tol * max(model_norm, 1.0):
            break
        if adaptive and it + 1 >= min_iterations:
            new = rho
            if primal > adaptive_rho_mu * dual:
                new = min(rho * adaptive_rho_tau, adaptive_rho_max)
            elif dual > adaptive_rho_mu * primal:
                new = max(rho / adaptive_rho_tau, adaptive_rho_min)
            if new != rho:
                u *= rho / new
                rho = float(new)
    return (model, history) if return_history else model


def constrained_minimiser(A, b, Q, bounds, tol=1e-10):
    """argmin |A y - b|^2 s.t. lower <= Q y <= upper, by enumerating active
    sets (3^m candidates; small m only)."""
    A = sparse.csr_matrix(A)
    Q = np.asarray(sparse.csr_matrix(Q).todense())
    N = (A.T @ A).toarray()
    g = A.T @ b
    n, m = N.shape[0], Q.shape[0]
    best, best_f = None, np.inf
    for state in itertools.product((0, 1, 2), repeat=m):
        act = [i for i in range(m) if state[i]]
        vals = [bounds[i, 0] if state[i] == 1 else bounds[i, 1] for i in act]
        if any(not np.isfinite(v) for v in vals):
            continue
        k = len(act)
        K = np.zeros((n + k, n + k))
        K[:n, :n] = N
        K[:n, n:] = Q[act].T
        K[n:, :n] = Q[act]
        r = np.concatenate([g, vals])
        try:
            y = np.linalg.solve(K, r)[:n]
        except np.linalg.LinAlgError:
            continue
        q = Q @ y
        scale = max(1.0, float(np.max(np.abs(q)))) if q.size else 1.0
        if np.all(q >= bounds[:, 0] - tol * scale) and np.all(q <= bounds[:, 1] + tol * scale):
            f = float(np.sum((A @ y - b) ** 2))
            if f < best_f:
                best, best_f = y, f
    return best
