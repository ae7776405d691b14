"""CPU restatement of the FD path on a rectilinear grid.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows LC/supports/_3d_rectilinear_grid.py (cell lookup by searchsorted,
strict inside test, unclipped local coordinates, shape gradients divided by
each cell's size, spacing-scaled second-derivative rows) and the FD
interpolator's use of it: data rows as on a uniform grid
(LI/_finite_difference_interpolator.py:236-476), interior regulariser rows
from ``build_scaled_operator_rows`` (:1166-1205), border rows with the uniform
masks at layers 1 and n-2 (:1052-1144, :636-641).
"""
from __future__ import annotations

import numpy as np

from . import fd as ofd


class RectGrid:
    def __init__(self, xn, yn, zn):
        self.axis = [np.asarray(a, dtype=float) for a in (xn, yn, zn)]
        self.nnodes = np.array([len(a) for a in self.axis])
        self.n_nodes = int(np.prod(self.nnodes))
        self.origin = np.array([a[0] for a in self.axis])
        self.maximum = np.array([a[-1] for a in self.axis])

    def nodes(self):
        xx, yy, zz = np.meshgrid(*self.axis, indexing="ij")
        return np.column_stack([xx.ravel(order="F"), yy.ravel(order="F"), zz.ravel(order="F")])

    def inside(self, pos):
        """:115-116, strict."""
        return np.all((pos > self.origin[None, :]) & (pos < self.maximum[None, :]), axis=1)

    def cell_index(self, pos):
        """:118-130."""
        ci = np.zeros((pos.shape[0], 3), dtype=int)
        for d, a in enumerate(self.axis):
            ci[:, d] = np.clip(np.searchsorted(a, pos[:, d], side="right") - 1, 0, len(a) - 2)
        return ci, self.inside(pos)

    def local_coords(self, pos):
        """:145-159."""
        ci, _ = self.cell_index(pos)
        out = np.zeros(pos.shape)
        for d, a in enumerate(self.axis):
            x0 = a[ci[:, d]]
            h = a[np.minimum(ci[:, d] + 1, len(a) - 1)] - x0
            h = np.where(h == 0, 1.0, h)
            out[:, d] = (pos[:, d] - x0) / h
        return out

    def corners(self, pos):
        """position_to_cell_corners (LC/supports/_3d_base_structured.py:386-412)."""
        ci, ins = self.cell_index(pos)
        nx, ny = self.nnodes[0], self.nnodes[1]
        g = ci[:, 0] + nx * ci[:, 1] + nx * ny * ci[:, 2]
        off = np.array([0, 1, nx, nx + 1, nx * ny, nx * ny + 1, nx * ny + nx, nx * ny + nx + 1])
        idc = g[:, None] + off[None, :]
        idc[~ins] = -1
        return idc, ins

    def element_gradient(self, pos):
        """:161-215: the uniform-grid T with local coordinates, then / h."""
        c = self.local_coords(pos)
        ci, _ = self.cell_index(pos)
        T = ofd.element_gradient_from_local(c)
        for d, a in enumerate(self.axis):
            h = a[ci[:, d] + 1] - a[ci[:, d]]
            T[:, d, :] /= np.where(h == 0, 1.0, h)[:, None]
        return T

    def evaluate_value(self, c, pos):
        """StructuredGrid.evaluate_value (_3d_structured_grid.py:295-322)."""
        idc, ins = self.corners(pos)
        v = np.full(pos.shape[0], np.nan)
        w = ofd.trilinear(self.local_coords(pos[ins]))
        v[ins] = np.sum(w * c[idc[ins]], axis=1)
        return v

    def evaluate_gradient(self, c, pos):
        """StructuredGrid.evaluate_gradient (:324-391): products T * c, then
        np.sum over the 8 corners; 0 outside."""
        idc, ins = self.corners(pos)
        T = np.zeros((pos.shape[0], 3, 8))
        T[ins] = self.element_gradient(pos[ins])
        for d in range(3):
            T[ins, d, :] *= c[idc[ins, :]]
        return np.array([np.sum(T[:, 0, :], axis=1), np.sum(T[:, 1, :], axis=1),
                         np.sum(T[:, 2, :], axis=1)]).T

    def scaled_rows(self, axis, cross=-1):
        """build_scaled_operator_rows, :232-315."""
        ijk = np.stack(np.meshgrid(*[np.arange(n) for n in self.nnodes], indexing="ij"),
                       -1).reshape(-1, 3, order="C")
        ijk = np.column_stack([ijk[:, 0], ijk[:, 1], ijk[:, 2]])
        gid = ijk[:, 0] + self.nnodes[0] * ijk[:, 1] + self.nnodes[0] * self.nnodes[1] * ijk[:, 2]
        order = np.argsort(gid)
        ijk = ijk[order]
        pure = cross < 0 or cross == axis
        m = np.ones(ijk.shape[0], dtype=bool)
        for a in ((axis,) if pure else (axis, cross)):
            m &= (ijk[:, a] > 0) & (ijk[:, a] < self.nnodes[a] - 1)
        c = ijk[m]

        def gi(x):
            return x[..., 0] + self.nnodes[0] * x[..., 1] + self.nnodes[0] * self.nnodes[1] * x[..., 2]

        rows = gi(c)
        if pure:
            a = self.axis[axis]
            i = c[:, axis]
            hL, hR = a[i] - a[i - 1], a[i + 1] - a[i]
            cm, cp = 2.0 / (hL * (hL + hR)), 2.0 / (hR * (hL + hR))
            A = np.column_stack([cm, -(cm + cp), cp])
            cols = np.repeat(c[:, None, :], 3, axis=1)
            cols[:, 0, axis] -= 1
            cols[:, 2, axis] += 1
        else:
            a, b = self.axis[axis], self.axis[cross]
            hx = a[c[:, axis] + 1] - a[c[:, axis] - 1]
            hy = b[c[:, cross] + 1] - b[c[:, cross] - 1]
            A = np.array([1.0, -1.0, -1.0, 1.0])[None, :] / (hx * hy)[:, None]
            cols = np.repeat(c[:, None, :], 4, axis=1)
            for k, (da, dc) in enumerate(((1, 1), (1, -1), (-1, 1), (-1, -1))):
                cols[:, k, axis] += da
                cols[:, k, cross] += dc
        return A, gi(cols), rows


def assemble(grid: RectGrid, data: dict, weights=None, regularisation=None):
    """setup_interpolator on a rectilinear grid with value / normal data:
    families in the reference order (norm_0..2, value, then the operators in
    RectilinearGrid.get_operators order dxx, dyy, dzz, dxy, dyz, dxz
    (LC/supports/_3d_rectilinear_grid.py:217-230), borders)."""
    wts = ofd.default_weights()
    if weights:
        wts.update(weights)
    if regularisation is not None:
        for k in ("dxy", "dyz", "dxz", "dxx", "dyy", "dzz"):
            wts[k] = float(regularisation)
    s = ofd.LSSystem(grid.n_nodes)
    normals = data.get("normal", np.zeros((0, 7)))
    if normals.shape[0]:
        idc, ins = grid.corners(normals[:, :3])
        T = grid.element_gradient(normals[ins, :3])
        w = wts["npw"] / 3 * normals[ins, 6]
        for d in range(3):
            s.add(T[:, d, :], normals[ins, 3 + d], idc[ins], w=w, name=f"norm_{d}")
    values = data.get("value", np.zeros((0, 5)))
    if values.shape[0]:
        idc, ins = grid.corners(values[:, :3])
        a = ofd.trilinear(grid.local_coords(values[ins, :3]))
        s.add(a, values[ins, 3], idc[ins], w=wts["cpw"] * values[ins, 4], name="value")
    for name, ax, cx, w in (("dxx", 0, -1, wts["dxx"]), ("dyy", 1, -1, wts["dyy"]),
                            ("dzz", 2, -1, wts["dzz"]), ("dxy", 0, 1, wts["dxy"] / 4),
                            ("dyz", 1, 2, wts["dyz"] / 4), ("dxz", 0, 2, wts["dxz"] / 4)):
        A, cols, _ = grid.scaled_rows(ax, cx)
        s.add(A, np.zeros(A.shape[0]), cols, w=w, name=name)
    og = ofd.Grid((0, 0, 0), grid.nnodes - 1, (1.0, 1.0, 1.0))
    gidx = og.node_index(np.arange(og.n_nodes))
    for mask, key, axis, upper in ((ofd.DXX, "dx", 0, False), (ofd.DXX, "dx", 0, True),
                                   (ofd.DYY, "dy", 1, False), (ofd.DYY, "dy", 1, True),
                                   (ofd.DZZ, "dz", 2, False), (ofd.DZZ, "dz", 2, True)):
        if grid.nnodes[axis] < 3 or wts[key] == 0:
            continue
        sel = gidx[:, axis] == (grid.nnodes[axis] - 2 if upper else 1)
        a, idc = ofd._stencil_rows(og, mask, gidx[sel])
        s.add(a, np.zeros(a.shape[0]), idc, w=wts[key],
              name=f"{key}_{'upper' if upper else 'lower'}")
    return s
