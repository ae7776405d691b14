"""Explicit sparse form of the least-squares system.

The GPU solvers never build the system matrix: data rows live on the device
as CSR blocks, and the regulariser is a stencil.  Callers written against the
reference may still inspect ``constraints[name]['matrix']``, call
``build_matrix()`` or hand the preprocessed system to their own solver
(``solve_system(solver=callable)``).  This module materialises those objects
on the host, on request only:

* data rows: copied from the device blocks (already row-normalised);
* FD regulariser families: one row of the normalised mask per centre node,
  centres in ascending node order (interior nodes, or one border layer), as
  ``_assemble_operator`` emits them (LI/_finite_difference_interpolator.py:643-692);
* P1 edge-jump rows: the GPU face table's rows (lsamd_p1_pairs).

Preprocessing follows LI/_solver_pipeline.py:86-124: ridge rows
``ridge_factor * I`` appended, then column scaling ``S = diag(1/|A[:, j]|)``
(LI/_discrete_interpolator.py:815-833).
"""
from __future__ import annotations

import numpy as np
from scipy import sparse


def mask_offsets(mask):
    """Non-zero taps of a (3,3,3) mask indexed [dz+1][dy+1][dx+1], in the
    mask's flattened order: (dx, dy, dz, value)."""
    m = np.asarray(mask, dtype=float)
    flat = np.flatnonzero(m.ravel())
    dz, dy, dx = np.unravel_index(flat, (3, 3, 3))
    return dx - 1, dy - 1, dz - 1, m.ravel()[flat]


def interior_centres(nnodes):
    """Ids of the nodes strictly inside on every axis, ascending."""
    nx, ny, nz = (int(v) for v in nnodes)
    if min(nx, ny, nz) < 3:
        return np.zeros(0, dtype=np.int64)
    i = np.arange(1, nx - 1, dtype=np.int64)
    j = np.arange(1, ny - 1, dtype=np.int64)
    k = np.arange(1, nz - 1, dtype=np.int64)
    return (k[:, None, None] * (nx * ny) + j[None, :, None] * nx + i[None, None, :]).ravel()


def layer_centres(nnodes, axis, upper):
    """Ids of the nodes whose index on ``axis`` is 1 (lower) or n-2 (upper),
    ascending (LI/_finite_difference_interpolator.py:636-641)."""
    n = [int(v) for v in nnodes]
    if n[axis] < 3:
        return np.zeros(0, dtype=np.int64)
    ranges = [np.arange(v, dtype=np.int64) for v in n]
    ranges[axis] = np.array([n[axis] - 2 if upper else 1], dtype=np.int64)
    i, j, k = ranges
    return (k[:, None, None] * (n[0] * n[1]) + j[None, :, None] * n[0] + i[None, None, :]).ravel()


def stencil_matrix(nnodes, mask, centres) -> sparse.csc_matrix:
    """Rows of the normalised ``mask`` centred at ``centres`` (len(centres) x
    n_nodes), one row per centre in the given order."""
    nx, ny, nz = (int(v) for v in nnodes)
    n = nx * ny * nz
    dx, dy, dz, vals = mask_offsets(mask)
    nrm = np.sqrt(np.sum(vals * vals))
    a = vals / nrm if nrm > 0 else vals
    R = centres.shape[0]
    cols = centres[:, None] + (dx + nx * dy + nx * ny * dz)[None, :]
    rows = np.repeat(np.arange(R, dtype=np.int64), vals.shape[0])
    return sparse.csc_matrix((np.tile(a, R), (rows, cols.ravel())), shape=(R, n))


def stack(families, dof):
    """``[w * A_f]`` and ``[w * b_f]`` over (A, b, w) families in order, empty
    families skipped (LI/_discrete_interpolator.py:684-705)."""
    mats, bs = [], []
    for A, b, w in families:
        w = np.asarray(w, dtype=float)
        if w.shape[0] == 0:
            continue
        mats.append(sparse.csr_matrix(A).multiply(w[:, None]))
        bs.append(np.asarray(b, dtype=float) * w)
    if not mats:
        return sparse.csr_matrix((0, dof), dtype=float), np.zeros(0)
    return sparse.vstack(mats).tocsr(), np.hstack(bs)


def preprocess(A, b, ridge_factor, add_ridge=True, scaling=True):
    """Ridge rows, then column scaling; returns (A S, b, S) with S None when
    scaling is off."""
    A = sparse.csr_matrix(A)
    b = np.asarray(b, dtype=float)
    if add_ridge:
        A = sparse.vstack([A, sparse.identity(A.shape[1], format="csr") * ridge_factor]).tocsr()
        b = np.concatenate([b, np.zeros(A.shape[1])])
    S = None
    if scaling:
        norms = np.sqrt(np.asarray(A.multiply(A).sum(axis=0)).ravel())
        f = np.ones(A.shape[1])
        f[norms > 0] = 1.0 / norms[norms > 0]
        S = sparse.diags(f)
        A = (A @ S).tocsr()
    return A, b, S
