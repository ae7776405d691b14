"""Inequality-constrained least squares by ADMM, with GPU x-updates.

The reference solves  min |A y - b|^2  subject to  l <= Q y <= u  with the
ADMM loop of LI/loopsolver/admm_solver.py:59-463 (driven from
LI/_solver_strategy.py:213-247): each major iteration projects ``Q y + u``
onto the bounds (LI/loopsolver/admm_method.py:4-25), updates the scaled dual
``u``, and takes the least-squares step on the stacked system
``[A; rho Q] y' ~ [b; rho (z - u)]``.  ``A`` here is the ridge-augmented,
column-scaled system ``[A; ridge I] S`` and ``y = S^-1 c``.

This module keeps that outer loop (rho schedule, adaptive rho, active set,
stopping tests, history) on torch device tensors.  The step is delegated to
an ``ops`` object that solves it on the GPU operator:

* ``ops.q(c)``          -> ``Q y`` for the node vector ``c = S y``;
* ``ops.set_weights(w)`` sets the per-row weight of the Q rows (rho, or 0 for
  rows outside the active set);
* ``ops.step(v, c)``    -> the ``c`` minimising
  ``|A c - b|^2 + sum_i w_i^2 (Q S^-1 c - v)_i^2`` (the exact x-update);
* ``ops.y_norm(x)``     -> ``|S^-1 x|``.

The reference takes the step with an inexact inner LSMR (btol 1e-6) warm
started from the previous step; this module solves it to a tight tolerance,
so its iterates are those of the same ADMM with exact steps.
"""
from __future__ import annotations

import math

import numpy as np
import torch


class Projection:
    """z/u state of the scaled ADMM (LI/loopsolver/admm_method.py:4-25)."""

    def __init__(self, n: int, device):
        if n < 1:
            raise ValueError("nelements must be greater than 0")
        self.z = torch.zeros(n, dtype=torch.float64, device=device)
        self.u = torch.zeros(n, dtype=torch.float64, device=device)

    def iterate(self, xmin, xmax, x):
        """z <- projection of x + u onto [xmin, xmax] (z kept where the
        argument compares false with both bounds, e.g. NaN); u <- u + x - z;
        returns z - u."""
        arg = x + self.u
        inside = (arg >= xmin) & (arg <= xmax)
        below = (arg < xmin) & ~inside
        above = (arg > xmax) & ~inside
        z = torch.where(inside, arg, self.z)
        z = torch.where(below, xmin, z)
        self.z = torch.where(above, xmax, z)
        self.u = self.u + x - self.z
        return self.z - self.u


def _rho_schedule(admm_weight, admm_weight_final, schedule, nmajor):
    if admm_weight_final is None:
        return np.full(nmajor, float(admm_weight))
    if admm_weight_final <= 0:
        raise ValueError("admm_weight_final must be > 0")
    if admm_weight <= 0:
        raise ValueError("admm_weight must be > 0")
    if schedule == "linear":
        return np.linspace(float(admm_weight), float(admm_weight_final), nmajor)
    return np.geomspace(float(admm_weight), float(admm_weight_final), nmajor)


def _active_rows(qx, xmin, xmax, prev, padding, hysteresis, min_size, max_size):
    """Rows near or beyond a bound (admm_solver.py:212-236)."""
    lower_gap = xmin - qx
    upper_gap = qx - xmax
    violation = torch.clamp(lower_gap, min=0.0) + torch.clamp(upper_gap, min=0.0)
    dist = torch.minimum((qx - xmin).abs(), (qx - xmax).abs())
    active = (violation > 0.0) | (dist <= padding)
    if prev is not None and hysteresis:
        active = active | prev
    n = qx.shape[0]
    score = violation + torch.clamp(padding - dist, min=0.0)
    if min_size > 0 and int(active.sum()) < min_size:
        k = int(min(min_size, n))
        if k > 0:
            active[torch.topk(score, k).indices] = True
    if max_size > 0 and int(active.sum()) > max_size:
        k = int(min(max_size, n))
        keep = torch.zeros_like(active)
        keep[torch.topk(score, k).indices] = True
        active = keep
    if int(active.sum()) == 0 and n > 0:
        active[int(torch.argmax(violation))] = True
    return active


def admm_solve(ops, bounds, c0, admm_weight=0.1, nmajor=200, adaptive_rho=False,
               adaptive_rho_mu=10.0, adaptive_rho_tau=2.0, adaptive_rho_min=1e-4,
               adaptive_rho_max=1e3, admm_weight_final=None, admm_weight_schedule="geometric",
               active_set=False, active_set_padding=1e-3, active_set_min_size=0,
               active_set_max_size=0, active_set_hysteresis=True, active_set_refresh_interval=1,
               admm_abs_tol=1e-4, admm_rel_tol=1e-3, model_update_tol=0.0, min_iterations=5,
               return_history=False):
    """Run the ADMM major iterations from ``c0``; returns ``c`` (and the
    history list when ``return_history``).  ``bounds`` is (n_ie, 2|3)
    [lower, upper(, 1)] on the device.  Argument meaning, defaults and
    stopping tests are those of admm_solver.py:59-463."""
    dev = c0.device
    if bounds.shape[1] not in (2, 3):
        raise ValueError("Bounds must have two columns")
    n_ie = bounds.shape[0]
    nmajor = int(nmajor)
    if nmajor <= 0:
        raise ValueError("nmajor must be positive")
    if model_update_tol < 0:
        raise ValueError("model_update_tol must be >= 0")
    proj = Projection(n_ie, dev)
    xmin, xmax = bounds[:, 0].contiguous(), bounds[:, 1].contiguous()
    rhos = _rho_schedule(admm_weight, admm_weight_final, admm_weight_schedule, nmajor)
    adaptive = bool(adaptive_rho) and admm_weight_final is None
    c = c0.clone()
    rho = float(rhos[0])
    active = None
    ones = torch.ones(n_ie, dtype=torch.float64, device=dev)
    ops.set_weights(rho * ones)
    history = []
    sqrt_n = math.sqrt(n_ie)
    for it in range(nmajor):
        target = float(rhos[it])
        reweight = False
        if target != rho:
            if rho > 0:
                proj.u *= rho / target
            rho = target
            reweight = True
        z_prev = proj.z.clone()
        qx = ops.q(c)
        v = proj.iterate(xmin, xmax, qx)
        if active_set:
            refresh = (active is None or int(active_set_refresh_interval) <= 1
                       or it % int(active_set_refresh_interval) == 0)
            if refresh:
                active = _active_rows(qx, xmin, xmax, active, active_set_padding,
                                      active_set_hysteresis, active_set_min_size,
                                      active_set_max_size)
                reweight = True
        if reweight or (active_set and it == 0):
            ops.set_weights(rho * (active.to(torch.float64) if active_set else ones))
        model_norm_before = ops.y_norm(c)
        c_new = ops.step(v, c)
        dx_norm = ops.y_norm(c_new - c)
        c = c_new
        primal = float(torch.linalg.vector_norm(qx - proj.z))
        dual = rho * float(torch.linalg.vector_norm(proj.z - z_prev))
        qx_norm = float(torch.linalg.vector_norm(qx))
        z_norm = float(torch.linalg.vector_norm(proj.z))
        u_norm = float(torch.linalg.vector_norm(rho * proj.u))
        eps_pri = sqrt_n * admm_abs_tol + admm_rel_tol * max(qx_norm, z_norm)
        eps_dual = sqrt_n * admm_abs_tol + admm_rel_tol * u_norm
        if return_history:
            history.append({"iteration": it + 1, "primal_norm": primal, "dual_norm": dual,
                            "eps_pri": eps_pri, "eps_dual": eps_dual, "rho": rho,
                            "active_count": int(active.sum()) if active_set and active is not None
                            else int(n_ie)})
        if it + 1 >= min_iterations and primal <= eps_pri and dual <= eps_dual:
            break
        if model_update_tol > 0.0 and it + 1 >= min_iterations:
            if dx_norm <= model_update_tol * max(model_norm_before, 1.0):
                if return_history:
                    history[-1].update(stopped_on_model_update=True, dx_norm=dx_norm,
                                       model_norm=model_norm_before)
                break
        if adaptive and it + 1 >= min_iterations:
            new = rho
            if primal > adaptive_rho_mu * dual:
                new = min(rho * adaptive_rho_tau, adaptive_rho_max)
            elif dual > adaptive_rho_mu * primal:
                new = max(rho / adaptive_rho_tau, adaptive_rho_min)
            if new != rho:
                if rho > 0:
                    proj.u *= rho / new
                rho = float(new)
                ops.set_weights(rho * (active.to(torch.float64) if active_set and
                                       active is not None else ones))
    if return_history:
        return c, history
    return c
