"""CPU restatement of the reference P1 (piecewise-linear, 5-tet) path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Follows
LC/supports/_3d_structured_tetra.py and LI/_p1interpolator.py with the same
numpy operations in the same order, so that tet selection, barycentric
coordinates and evaluation reproduce the reference's floating-point results.

Pinning (tests/test_oracle_p1.py): closed-form barycentric coordinates and
element gradients (partition of unity, linear reproduction, M^-1 identities),
the 5-tet decomposition covering each cell exactly (volumes sum to the cell),
neighbour symmetry and face sharing brute-forced from element node sets, the
edge-jump operator annihilating every globally linear field and agreeing with
the reference's own coo/tril construction, and scalar triple products equal
to numpy's einsum on this machine.
"""
from __future__ import annotations

import numpy as np
from scipy import sparse
from scipy.sparse import coo_matrix, tril

from .fd import EPS, LSSystem

# LC/supports/_3d_structured_tetra.py:28-34 (corner ids in the cell's
# 000,100,010,110,001,101,011,111 order)
TETRA_MASK_EVEN = np.array([[7, 1, 2, 4], [6, 2, 4, 7], [5, 1, 4, 7], [0, 1, 2, 4], [3, 1, 2, 7]])
TETRA_MASK_ODD = np.array([[0, 6, 5, 3], [7, 3, 5, 6], [4, 0, 5, 6], [2, 0, 3, 6], [1, 0, 3, 5]])
_XC = np.array([0, 1, 0, 1, 0, 1, 0, 1])
_YC = np.array([0, 0, 1, 1, 0, 0, 1, 1])
_ZC = np.array([0, 0, 0, 0, 1, 1, 1, 1])
_I = np.array([[-1.0, 1.0, 0.0, 0.0], [-1.0, 0.0, 1.0, 0.0], [-1.0, 0.0, 0.0, 1.0]])


def dot3(a, b):
    """np.einsum('ikj,ikj->ik') over a length-3 axis: numpy 2.2's
    sum-of-products loop adds the products as (p0 + p2) + p1."""
    p = a * b
    return (p[..., 0] + p[..., 2]) + p[..., 1]


def cross3(a, b):
    """np.cross(a, b, axisa=-1, axisb=-1) (component formula)."""
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


class TetMesh:
    """Geometry and topology of ``TetMesh(origin, nsteps_cells, step_vector)``:
    5 tets per cell, cells x fastest, tet id = local + 5*cell
    (LC/supports/_3d_structured_tetra.py:16-118,376-419)."""

    def __init__(self, origin, nsteps_cells, step_vector):
        self.origin = np.array(origin, dtype=float)
        self.step = np.array(step_vector, dtype=float)
        self.ncells = np.array(nsteps_cells, dtype=int)
        self.nnodes = self.ncells + 1
        self._elements = None
        self._faces = None

    @property
    def n_nodes(self):
        return int(np.prod(self.nnodes))

    @property
    def n_cells(self):
        return int(np.prod(self.ncells))

    @property
    def ntetra(self):
        return 5 * self.n_cells

    @property
    def maximum(self):
        return self.origin + self.ncells * self.step

    def nodes(self):
        """_3d_base_structured.py:233-245 (linspace, F order)."""
        mx = self.origin + self.ncells * self.step
        x = np.linspace(self.origin[0], mx[0], self.nnodes[0])
        y = np.linspace(self.origin[1], mx[1], self.nnodes[1])
        z = np.linspace(self.origin[2], mx[2], self.nnodes[2])
        xx, yy, zz = np.meshgrid(x, y, z, indexing="ij")
        return np.array([xx.flatten(order="F"), yy.flatten(order="F"), zz.flatten(order="F")]).T

    def _cells(self):
        x, y, z = (np.arange(0, n) for n in self.ncells)
        zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
        return np.array([xx.flatten(), yy.flatten(), zz.flatten()]).T

    def corner_indexes(self, cell):
        """cell_corner_indexes, _3d_base_structured.py:353-384 -> (N,8,3)."""
        return np.stack([cell[:, None, 0] + _XC[None, :], cell[:, None, 1] + _YC[None, :],
                         cell[:, None, 2] + _ZC[None, :]], axis=-1)

    def global_node(self, idx):
        nx, ny = self.nnodes[0], self.nnodes[1]
        return idx[..., 0] + nx * idx[..., 1] + nx * ny * idx[..., 2]

    def node_position(self, idx):
        """node_indexes_to_position, _3d_base_structured.py:431-439."""
        return np.stack([self.origin[a] + self.step[a] * idx[..., a] for a in range(3)], axis=-1)

    @property
    def elements(self):
        """get_elements, _3d_structured_tetra.py:376-399 -> (ntetra, 4) int64."""
        if self._elements is None:
            cells = self._cells()
            gi = self.global_node(self.corner_indexes(cells))
            even = np.sum(cells, axis=1) % 2 == 0
            t = np.zeros((cells.shape[0], 5, 4), dtype=np.int64)
            t[even] = gi[even][:, TETRA_MASK_EVEN]
            t[~even] = gi[~even][:, TETRA_MASK_ODD]
            self._elements = t.reshape(-1, 4)
        return self._elements

    def tetra_global_index(self, cell, local):
        """_3d_structured_tetra.py:401-419."""
        nx, ny = self.ncells[0], self.ncells[1]
        return local + cell[:, 0] * 5 + nx * cell[:, 1] * 5 + nx * ny * cell[:, 2] * 5

    # ---------------------------------------------------------- A16 ----
    def inside(self, pos):
        """Strict box test, _3d_structured_tetra.py:264-272."""
        ins = np.ones(pos.shape[0], dtype=bool)
        for i in range(3):
            ins *= pos[:, i] > self.origin[i]
            ins *= pos[:, i] < self.origin[i] + self.step[i] * self.ncells[i]
        return ins

    def cell_index(self, pos):
        """position_to_cell_index (_3d_base_structured.py:251-282) evaluated
        with the TetMesh's strict ``inside`` (the method it dispatches to)."""
        ins = self.inside(pos)
        ci = np.zeros((pos.shape[0], 3), dtype=int)
        for a in range(3):
            x = pos[:, a] - self.origin[a]
            ci[ins, a] = np.floor(x[ins] / self.step[a]).astype(int)
            ci[ins, a] = np.clip(ci[ins, a], 0, self.ncells[a] - 1)
        return ci, ins

    def element_for_location(self, pos):
        """get_element_for_location, _3d_structured_tetra.py:274-365.

        Returns (vertices (N,4,3), bc (N,4), tet (N,), inside (N,)); NaN / -1
        outside.  A point with non-negative coordinates in several tets takes
        the last of them (the dict(zip) at :331-337)."""
        pos = np.array(pos, dtype=float)[:, :3]
        cell, ins = self.cell_index(pos)
        even = np.sum(cell, axis=1) % 2 == 0
        corners = self.corner_indexes(cell)
        nodes = self.node_position(corners)
        vert = np.full((pos.shape[0], 5, 4, 3), np.nan)
        vert[even] = nodes[even][:, TETRA_MASK_EVEN, :]
        vert[~even] = nodes[~even][:, TETRA_MASK_ODD, :]
        vap = pos[:, None, :] - vert[:, :, 0, :]
        vbp = pos[:, None, :] - vert[:, :, 1, :]
        vab = vert[:, :, 1, :] - vert[:, :, 0, :]
        vac = vert[:, :, 2, :] - vert[:, :, 0, :]
        vad = vert[:, :, 3, :] - vert[:, :, 0, :]
        vbc = vert[:, :, 2, :] - vert[:, :, 1, :]
        vbd = vert[:, :, 3, :] - vert[:, :, 1, :]
        va = dot3(vbp, cross3(vbd, vbc)) / 6.0
        vb = dot3(vap, cross3(vac, vad)) / 6.0
        vc = dot3(vap, cross3(vad, vab)) / 6.0
        vd = dot3(vap, cross3(vab, vac)) / 6.0
        v = dot3(vab, cross3(vac, vad)) / 6.0
        c = np.stack([va / v, vb / v, vc / v, vd / v], axis=-1)
        mask = np.all(c >= 0, axis=2)
        i, j = np.where(mask)
        pairs = dict(zip(i, j))
        mask[:] = False
        mask[list(pairs.keys()), list(pairs.values())] = True
        ins = np.logical_and(ins, np.any(mask, axis=1))
        ins = np.logical_and(ins, self.inside(pos))
        mask[~ins, :] = False
        vr = np.full((pos.shape[0], 4, 3), np.nan)
        vr[ins] = vert[mask]
        cr = np.full((pos.shape[0], 4), np.nan)
        cr[ins] = c[mask]
        tr = np.full(pos.shape[0], -1, dtype=int)
        local = np.tile(np.arange(5)[None, :], (mask.shape[0], 1))[mask]
        tr[ins] = self.tetra_global_index(cell[ins], local)
        return vr, cr, tr, ins

    def element_gradients(self, elements=None):
        """get_element_gradients, _3d_structured_tetra.py:421-486: inv(M) @ I
        with M rows v1-v0, v2-v0, v3-v0 of every tet, then indexed."""
        cells = self._cells()
        even = np.sum(cells, axis=1) % 2 == 0
        nodes = self.node_position(self.corner_indexes(cells))
        pts = np.zeros((cells.shape[0], 5, 4, 3))
        pts[even] = nodes[even][:, TETRA_MASK_EVEN, :]
        pts[~even] = nodes[~even][:, TETRA_MASK_ODD, :]
        ps = pts.reshape(-1, 4, 3)
        g = _grad_from_vertices(ps)
        if elements is None:
            return g
        return g[elements]

    def element_gradient_for_location(self, pos):
        """get_element_gradient_for_location, :502-545 (NaN outside)."""
        vert, _, tet, ins = self.element_for_location(pos)
        g = np.full((vert.shape[0], 3, 4), np.nan)
        if np.any(ins):
            g[ins] = _grad_from_vertices(vert[ins])
        return vert, g, tet, ins

    def element_size(self):
        """|det(v1-v0, v2-v0, v3-v0)| / 6, :80-94."""
        e = self.elements
        nd = self.nodes()
        vecs = nd[e[:, :4], :][:, 1:, :] - nd[e[:, :4], :][:, 0, None, :]
        return np.abs(np.linalg.det(vecs)) / 6

    def barycentre(self):
        return np.sum(self.nodes()[self.elements][:, :, :], axis=1) / 4.0

    # ---------------------------------------------------------- A18 ----
    def evaluate_value(self, pos, c):
        """_3d_structured_tetra.py:208-229: NaN outside (strict)."""
        out = np.full(pos.shape[0], np.nan)
        _, bc, tet, ins = self.element_for_location(pos)
        out[ins] = np.sum(bc[ins, :] * c[self.elements[tet[ins]]], axis=1)
        return out

    def evaluate_gradient(self, pos, c):
        """_3d_structured_tetra.py:231-262: NaN outside."""
        out = np.full((pos.shape[0], 3), np.nan)
        _, g, tet, ins = self.element_gradient_for_location(pos)
        out[ins, :] = (g[ins, :, :] * c[self.elements[tet[ins]][:, None, :]]).sum(2)
        return out

    # ---------------------------------------------------------- A15 ----
    def neighbours(self):
        """get_neighbours, _3d_structured_tetra.py:626-740 -> (ntetra, 4), -1 = none."""
        t = np.arange(0, self.ntetra)
        nb = np.full((self.ntetra, 4), -9999, dtype=np.int64)
        nb[t % 5 == 0, :] = t[t % 5 == 0, None] + np.arange(1, 5)[None, :]
        nb[t % 5 != 0, 0] = np.tile(t[t % 5 == 0], (4, 1)).flatten(order="F")
        cx = self.ncells
        cell = t // 5
        ci = np.stack([cell % cx[0], cell // cx[0] % cx[1], cell // cx[0] // cx[1]], axis=0)
        odd = (np.sum(ci, axis=0) % 2 == 1)
        for local, odd_flag, m in NEIGHBOUR_TABLE:
            sel = np.logical_and(t % 5 == local, odd if odd_flag else ~odd)
            c = ci[:, sel]
            nc = np.stack([c[a][:, None] + m[:, a] for a in range(3)], axis=-1)
            ok = np.all((nc >= 0) & (nc < cx[None, None, :]), axis=2)
            g = (nc[:, :, 0] + nc[:, :, 1] * cx[0] + nc[:, :, 2] * cx[0] * cx[1]) * 5 + m[:, 3]
            g[~ok] = -1
            nb[sel, 1:] = g
        return nb

    def face_table(self):
        """_init_face_table, _3d_structured_tetra.py:120-184 ->
        (pairs (P,2) with pairs[:,0] > pairs[:,1], faces (P,3) node ids)."""
        if self._faces is not None:
            return self._faces
        ne = self.ntetra
        rows = np.tile(np.arange(ne)[:, None], (1, 4))
        el = self.elements
        nb = self.neighbours()
        en = coo_matrix((np.ones(ne * 4), (rows.ravel(), el.ravel())),
                        shape=(ne, self.n_nodes), dtype=bool).tocsr()
        n1 = np.tile(np.arange(nb.shape[0], dtype=int)[:, None], (1, 4)).flatten()
        n2 = nb.flatten()
        n1, n2 = n1[n2 >= 0], n2[n2 >= 0]
        pairs_m = coo_matrix((np.ones(n1.shape[0]), (n1, n2))).tocsr()
        i, j = tril(pairs_m).nonzero()
        rel = np.stack([i, j], axis=1)
        faces = sparse.csr_matrix(en[rel[:, 0], :].multiply(en[rel[:, 1], :]))
        keep = np.array(np.sum(faces, axis=1) == 3).flatten()
        shared = faces[keep, :]
        r, col = shared.nonzero()
        # the reference's two argsorts of an already sorted array (:170-171):
        # they may permute the 3 nodes of a face, which flips the sign of its
        # normal and hence of the (normalised) edge-jump row, nothing else
        r = r[r.argsort()]
        col = col[r.argsort()]
        idx = np.full((shared.shape[0], 3), -1, dtype=int)
        idx[r.reshape(-1, 3)[:, 0], :] = col.reshape(-1, 3)
        self._faces = (rel[keep], idx)
        return self._faces

    def shared_element_norm(self):
        """(v1-v0) x (v2-v0) of each shared face, :186-194 (unnormalised)."""
        _, f = self.face_table()
        nd = self.nodes()
        v1 = nd[f[:, 1], :] - nd[f[:, 0], :]
        v2 = nd[f[:, 2], :] - nd[f[:, 0], :]
        return np.cross(v1, v2, axisa=1, axisb=1)


# _3d_structured_tetra.py:652-708: (local tet, odd cell?, rows (dx,dy,dz,local neighbour))
NEIGHBOUR_TABLE = [
    (1, True, np.array([[1, 0, 0, 1], [0, 1, 0, 2], [0, 0, 1, 4]])),
    (2, True, np.array([[-1, 0, 0, 2], [0, -1, 0, 1], [0, 0, 1, 3]])),
    (3, True, np.array([[-1, 0, 0, 4], [0, 1, 0, 3], [0, 0, -1, 1]])),
    (4, True, np.array([[1, 0, 0, 3], [0, -1, 0, 4], [0, 0, -1, 2]])),
    (1, False, np.array([[-1, 0, 0, 1], [0, 1, 0, 2], [0, 0, 1, 3]])),
    (2, False, np.array([[1, 0, 0, 2], [0, -1, 0, 1], [0, 0, 1, 4]])),
    (3, False, np.array([[-1, 0, 0, 4], [0, -1, 0, 3], [0, 0, -1, 2]])),
    (4, False, np.array([[1, 0, 0, 3], [0, 1, 0, 4], [0, 0, -1, 1]])),
]


def _grad_from_vertices(ps):
    """inv(M) @ I with numpy's (3,3,N) -> swapaxes construction (:449-486)."""
    m = np.array([[ps[:, r, c] - ps[:, 0, c] for c in range(3)] for r in range(1, 4)])
    m = np.swapaxes(m, 0, 2)
    g = np.linalg.inv(m).swapaxes(1, 2)
    return g @ _I


# ------------------------------------------------------------- A17 ----
def default_weights():
    """LI/_p1interpolator.py:37-44."""
    return {"cgw": 0.1, "cpw": 1.0, "npw": 1.0, "gpw": 1.0, "tpw": 1.0, "ipw": 1.0}


def edge_jump_rows(mesh: TetMesh):
    """minimise_edge_jumps, LI/_p1interpolator.py:115-166 (unnormalised):
    [n.grad(T1), -n.grad(T2)] over [nodes(T1), nodes(T2)], T1 = pairs[:,0]."""
    rel, _ = mesh.face_table()
    norm = mesh.shared_element_norm()
    g = mesh.element_gradients()
    dt = g[rel[:, 0]]
    dn = g[rel[:, 1]]
    ct = np.einsum("ij,ijk->ik", norm, dt)
    cn = -np.einsum("ij,ijk->ik", norm, dn)
    a = np.hstack([ct, cn])
    idc = np.hstack([mesh.elements[rel[:, 0]], mesh.elements[rel[:, 1]]])
    return a, idc


def assemble_p1(mesh: TetMesh, data: dict, weights: dict | None = None, regularisation=None,
                scale=None):
    """P1Interpolator.setup_interpolator (LI/_p1interpolator.py:180-246).

    data: {"value": (N,4+) [x y z v ...], "normal": (N,6+), "tangent": (N,6+)}.
    Family order: edge jump, norm, value, tangent; gradient constraints add
    nothing (:70-71) and per-point weights are ignored (:80-81,104-105).
    ``scale`` (n_nodes,): use_regularisation_weight_scale (:150-154)."""
    w = default_weights()
    if regularisation is not None:
        w["cgw"] = regularisation
    if weights:
        w.update(weights)
    s = LSSystem(mesh.n_nodes)
    if w["cgw"] > 0:
        a, idc = edge_jump_rows(mesh)
        rw = w["cgw"]
        if scale is not None:
            rel, _ = mesh.face_table()
            st = scale[mesh.elements[rel[:, 0]]].mean(axis=1)
            sn = scale[mesh.elements[rel[:, 1]]].mean(axis=1)
            rw = 0.5 * (st + sn) * w["cgw"]
        s.add(a, np.zeros(a.shape[0]), idc, rw, "edge jump")
    if "normal" in data and len(data["normal"]):
        p = np.asarray(data["normal"], dtype=float)
        _, tet, ins = mesh.element_for_location(p[:, :3])[1:]
        g = mesh.element_gradients(tet)
        wt = np.ones(int(ins.sum())) * w["npw"]
        el = np.tile(mesh.elements[tet[ins]], (3, 1, 1)).swapaxes(0, 1)
        s.add(g[ins, :, :], p[ins, 3:6], el, wt, "norm")
    if "value" in data and len(data["value"]):
        p = np.asarray(data["value"], dtype=float)
        _, bc, tet, ins = mesh.element_for_location(p[:, :3])
        wt = np.ones(int(ins.sum())) * w["cpw"]
        s.add(bc[ins, :], p[ins, 3], mesh.elements[tet[ins], :], wt, "value")
    if "tangent" in data and len(data["tangent"]):
        p = np.asarray(data["tangent"], dtype=float)
        _, tet, ins = mesh.element_for_location(p[:, :3])[1:]
        g = mesh.element_gradients(tet)
        size = mesh.element_size()[tet[ins]]
        wt = np.ones(size.shape[0]) * w["tpw"] * size
        vec = p[:, 3:6].copy()
        nrm = np.linalg.norm(vec, axis=1)
        vec[nrm > 0, :] /= nrm[nrm > 0, None]
        a = np.einsum("ij,ijk->ik", vec[ins, :3], g[ins, :, :])
        s.add(a, np.zeros(a.shape[0]), mesh.elements[tet[ins], :], wt,
              "undefined gradient orthogonal constraint")
    return s
