"""Pins of the CPU oracle for the P1 path (oracle/p1.py).

Each check is independent of the oracle's own formulas: numpy's summation
orders, closed forms of barycentric coordinates and P1 gradients, brute-force
topology from element node sets, invariants of the edge-jump operator, and a
linear-field reproduction through the full least-squares solve.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import sparse

from oracle import p1


def _mesh(cells=(4, 3, 5), step=(0.3, 0.5, 0.2), origin=(1.0, -2.0, 0.5)):
    return p1.TetMesh(origin, cells, step)


# ------------------------------------------------------------ arithmetic ----
def test_dot3_is_numpy_einsum_order():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((20000, 5, 3)) * np.exp(rng.uniform(-20, 20, (20000, 5, 3)))
    b = rng.standard_normal((20000, 5, 3)) * np.exp(rng.uniform(-20, 20, (20000, 5, 3)))
    assert np.array_equal(p1.dot3(a, b), np.einsum("ikj,ikj->ik", a, b))
    assert np.array_equal(p1.cross3(a, b), np.cross(a, b, axisa=2, axisb=2))


def test_row_sum_and_norm_orders_used_by_the_kernels():
    """Sums the HIP kernels spell out: 4 terms left to right (np.sum and
    np.linalg.norm over axis 1), 8 squares pairwise (np.linalg.norm)."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal((50000, 4)) * np.exp(rng.uniform(-15, 15, (50000, 4)))
    assert np.array_equal(np.sum(x, axis=1), ((x[:, 0] + x[:, 1]) + x[:, 2]) + x[:, 3])
    s = x * x
    assert np.array_equal(np.linalg.norm(x, axis=1),
                          np.sqrt(((s[:, 0] + s[:, 1]) + s[:, 2]) + s[:, 3]))
    y = rng.standard_normal((50000, 8)) * np.exp(rng.uniform(-15, 15, (50000, 8)))
    q = y * y
    pair = ((q[:, 0] + q[:, 1]) + (q[:, 2] + q[:, 3])) + ((q[:, 4] + q[:, 5]) + (q[:, 6] + q[:, 7]))
    assert np.array_equal(np.linalg.norm(y, axis=1), np.sqrt(pair))
    v = x[:, :3]
    assert np.array_equal(np.linalg.norm(v, axis=1),
                          np.sqrt((v[:, 0] ** 2 + v[:, 1] ** 2) + v[:, 2] ** 2))


# -------------------------------------------------------------- topology ----
def test_five_tets_tile_every_cell():
    m = _mesh()
    nd = m.nodes()
    el = m.elements
    v = nd[el[:, 1:]] - nd[el[:, :1]]
    vol = np.abs(np.linalg.det(v)) / 6
    cell_vol = np.prod(m.step)
    assert np.allclose(vol.reshape(-1, 5).sum(1), cell_vol, rtol=1e-12)
    # the centre tet is a third of the cell, corner tets a sixth
    assert np.allclose(vol.reshape(-1, 5)[:, 0], cell_vol / 3, rtol=1e-12)
    assert np.allclose(vol.reshape(-1, 5)[:, 1:], cell_vol / 6, rtol=1e-12)
    assert np.allclose(m.element_size(), vol, rtol=1e-12)


def test_neighbours_match_brute_force_face_sharing():
    m = _mesh((3, 4, 2))
    el = m.elements
    nb = m.neighbours()
    sets = [set(e) for e in el]
    # brute force: tets sharing exactly 3 nodes
    for t in range(m.ntetra):
        cand = {u for u in range(m.ntetra) if u != t and len(sets[t] & sets[u]) == 3}
        got = {int(u) for u in nb[t] if u >= 0}
        assert got == cand, t
    assert np.all((nb >= 0).sum(1) <= 4)


def test_face_table_pairs_and_faces():
    m = _mesh((3, 4, 2))
    pairs, faces = m.face_table()
    c = m.ncells
    P = 4 * np.prod(c) + 2 * ((c[0] - 1) * c[1] * c[2] + c[0] * (c[1] - 1) * c[2]
                              + c[0] * c[1] * (c[2] - 1))
    assert pairs.shape == (P, 2) and faces.shape == (P, 3)
    assert np.all(pairs[:, 0] > pairs[:, 1])
    key = pairs[:, 0] * m.ntetra + pairs[:, 1]
    assert np.all(np.diff(key) > 0)  # sorted by larger id, then smaller
    el = m.elements
    for (a, b), f in zip(pairs, faces):
        assert set(f) == set(el[a]) & set(el[b])


# ------------------------------------------------------------ A16 / A18 ----
def test_barycentric_closed_form_and_linear_reproduction():
    m = _mesh()
    rng = np.random.default_rng(3)
    pts = m.origin + rng.uniform(0.001, 0.999, (4000, 3)) * m.ncells * m.step
    vert, bc, tet, ins = m.element_for_location(pts)
    assert ins.all()
    assert np.all(bc >= 0)
    assert np.allclose(bc.sum(1), 1.0, atol=1e-13)
    assert np.allclose(np.einsum("nk,nkd->nd", bc, vert), pts, atol=1e-12)
    # the tet's nodes are the element's
    pos = m.node_position(np.stack(np.unravel_index(m.elements[tet], m.nnodes, order="F"), -1))
    assert np.allclose(pos, vert, atol=1e-12)


def test_last_tet_wins_on_shared_faces():
    """A point on the face shared by two tets of one cell belongs to the
    one with the larger local id (dict(zip) in the reference)."""
    m = p1.TetMesh((0, 0, 0), (2, 2, 2), (1.0, 1.0, 1.0))
    # a dyadic point (exact arithmetic) of the face shared by the centre tet
    # (0) and tet 4 of cell 0 (even cell: tet 0 = corners 7,1,2,4; tet 4 =
    # corners 3,1,2,7 -> face 1,2,7)
    face = 0.5 * np.array([1.0, 0, 0]) + 0.25 * np.array([0, 1.0, 0]) + 0.25 * np.ones(3)
    _, _, tet, ins = m.element_for_location(face[None, :])
    assert ins[0] and tet[0] == 4


def test_gradients_reproduce_linear_fields():
    m = _mesh()
    g = m.element_gradients()
    nd = m.node_position(np.stack(np.unravel_index(np.arange(m.n_nodes), m.nnodes, order="F"), -1))
    a = np.array([0.3, -1.2, 2.5])
    f = nd @ a + 0.7
    grad = np.einsum("tdk,tk->td", g, f[m.elements])
    assert np.allclose(grad, a[None, :], rtol=1e-12, atol=1e-12)
    assert np.allclose(g.sum(2), 0.0, atol=1e-12)


def test_evaluation_of_linear_field_and_outside():
    m = _mesh()
    nd = m.nodes()
    a = np.array([1.5, 0.25, -0.5])
    c = nd @ a - 3.0
    rng = np.random.default_rng(5)
    pts = m.origin + rng.uniform(0.01, 0.99, (3000, 3)) * m.ncells * m.step
    assert np.allclose(m.evaluate_value(pts, c), pts @ a - 3.0, atol=1e-12)
    assert np.allclose(m.evaluate_gradient(pts, c), a[None, :], atol=1e-10)
    out = np.array([m.origin, m.maximum + 1, [np.nan, 0, 0]])
    assert np.all(np.isnan(m.evaluate_value(out, c)))
    assert np.all(np.isnan(m.evaluate_gradient(out, c)))


# -------------------------------------------------------------------- A17 ----
def test_edge_jump_operator_invariants():
    m = _mesh((4, 4, 3))
    s = p1.assemble_p1(m, {}, regularisation=0.1)
    A, b = s.build_matrix()
    nrows = A.shape[0]
    pairs, _ = m.face_table()
    assert nrows == pairs.shape[0] and np.all(b == 0)
    N = (A.T @ A).tocsr()
    assert abs(N - N.T).max() < 1e-15
    nd = m.nodes()
    for f in (nd[:, 0], nd[:, 1], nd[:, 2], np.ones(m.n_nodes)):
        assert np.abs(N @ f).max() < 1e-13
    # rows: 8 unmerged entries of unit norm (weight 0.1 applied)
    a, idc = p1.edge_jump_rows(m)
    assert np.allclose(np.linalg.norm(A.toarray(), axis=1) > 0, True)
    nrm = np.linalg.norm(a, axis=1)
    Ad = sparse.coo_matrix((np.ravel(a / nrm[:, None] * 0.1),
                            (np.repeat(np.arange(nrows), 8), idc.ravel())),
                           shape=(nrows, m.n_nodes)).tocsr()
    assert abs(Ad - A).max() < 1e-15


def test_p1_solve_reproduces_a_linear_field():
    """Consistent values of a linear field: the regulariser vanishes on it,
    so the least-squares minimiser is the field itself."""
    from oracle import fd as ofd

    m = p1.TetMesh((0, 0, 0), (6, 5, 7), (1 / 6, 1 / 5, 1 / 7))
    rng = np.random.default_rng(11)
    pts = rng.uniform(0.02, 0.98, (300, 3))
    f = lambda p: p @ np.array([0.2, -0.4, 1.0]) + 0.1  # noqa: E731
    vals = np.column_stack([pts, f(pts), np.ones(300)])
    s = p1.assemble_p1(m, {"value": vals})
    c = ofd.solve_exact(s)
    assert np.allclose(c, f(m.nodes()), atol=1e-6)


def test_p1_ignores_point_weights_and_gradient_rows():
    m = p1.TetMesh((0, 0, 0), (3, 3, 3), (1 / 3,) * 3)
    rng = np.random.default_rng(2)
    v = np.column_stack([rng.uniform(0.05, 0.95, (20, 3)), rng.standard_normal(20),
                         rng.uniform(0.1, 5, 20)])
    s1 = p1.assemble_p1(m, {"value": v})
    v2 = v.copy()
    v2[:, 4] = 1.0
    s2 = p1.assemble_p1(m, {"value": v2, "gradient": np.ones((5, 7))})
    A1, b1 = s1.build_matrix()
    A2, b2 = s2.build_matrix()
    assert abs(A1 - A2).max() == 0 and np.array_equal(b1, b2)


@pytest.mark.parametrize("cells", [(4, 5, 6), (9, 10, 11), (12, 8, 7)])
def test_operator_table_of_the_library_matches_oracle(cells):
    """The library's host-built 21-tap table (no GPU work) applied per node
    equals the oracle's A^T A of the edge-jump rows."""
    from loopstructural_amd import _lib

    step = (0.1, 0.13, 0.07)
    m = p1.TetMesh((0, 0, 0), cells, step)
    s = p1.assemble_p1(m, {}, regularisation=0.1)
    A, _ = s.build_matrix()
    N = (A.T @ A).tocsr()
    g = _lib.make_grid((0, 0, 0), step, np.array(cells) + 1)
    tab = np.zeros(343 * 21)
    nv = np.zeros(3, dtype=np.int32)
    _lib.check(_lib.load().lsamd_p1_operator_table(C.byref(g), 0.1, tab.ctypes.data_as(C.c_void_p),
                                                   nv.ctypes.data_as(C.c_void_p)))
    offs = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
            (2, 0, 0), (-2, 0, 0), (0, 2, 0), (0, -2, 0), (0, 0, 2), (0, 0, -2)]
    offs += [((1 if b & 1 else -1), (1 if b & 2 else -1), (1 if b & 4 else -1)) for b in range(8)]

    def vmap(i, n):
        if n <= 7 or i < 2:
            return i
        nvv = 6 if n % 2 == 0 else 7
        return i - (n - nvv) if i >= n - 2 else 2 + (i & 1)

    nx, ny, nz = m.nnodes
    x = np.random.default_rng(0).standard_normal(m.n_nodes)
    y = np.zeros(m.n_nodes)
    for n in range(m.n_nodes):
        i, j, k = n % nx, n // nx % ny, n // nx // ny
        row = tab[((vmap(k, nz) * nv[1] + vmap(j, ny)) * nv[0] + vmap(i, nx)) * 21:][:21]
        for q, (dx, dy, dz) in enumerate(offs):
            ii, jj, kk = i + dx, j + dy, k + dz
            if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                y[n] += row[q] * x[ii + nx * jj + nx * ny * kk]
            else:
                assert row[q] == 0
    ref = N @ x
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()


# P1 rows built from element gradients are pinned through their first moments:
# sum_k dl_k/dx_d v_k = e_d for the four vertices v_k of a tet, so a row
# a = u . grad(l) (normalised) maps the node coordinates to u/|a| and sums to 0.
def _inside_points(m, n, seed):
    rng = np.random.default_rng(seed)
    return m.origin + rng.uniform(0.02, 0.98, (n, 3)) * m.ncells * m.step, rng


def test_normal_rows_are_axis_derivatives():
    """LI/_p1interpolator.py:73-96: three rows per point (row 3i + d), grad_d(l)
    normalised, b = n_d/|grad_d(l)|, weight npw (point weights ignored)."""
    m = _mesh()
    p, rng = _inside_points(m, 25, 31)
    nv = rng.standard_normal((25, 3))
    s = p1.assemble_p1(m, {"normal": np.column_stack([p, nv, rng.uniform(0.1, 9, 25)])},
                       weights={"npw": 4.0})
    fam = s.constraints["norm"]
    M = fam["matrix"]
    assert M.shape[0] == 75
    mom = M @ m.nodes()
    d = np.tile(np.arange(3), 25)
    e = np.eye(3)[d]
    scale = mom[np.arange(75), d]
    assert np.all(scale > 0)
    assert np.allclose(mom, e * scale[:, None], atol=1e-12 * scale.max())
    assert np.allclose(fam["b"], nv.reshape(-1) * scale, rtol=1e-12)
    assert np.allclose(M @ np.ones(m.n_nodes), 0.0, atol=1e-12 * scale.max())
    assert np.allclose(sparse.linalg.norm(M, axis=1), 1.0)
    assert np.all(fam["w"] == 4.0)


def test_tangent_rows_and_tet_volume_weights():
    """LI/_p1interpolator.py:248-296: one row t_hat . grad(l) per point, b = 0,
    weight tpw x the volume of the containing tet (a third of the cell for the
    central tet, local id 0, a sixth for the four corner tets)."""
    m = _mesh()
    p, rng = _inside_points(m, 40, 32)
    v = rng.standard_normal((40, 3)) * rng.uniform(0.2, 5.0, (40, 1))
    s = p1.assemble_p1(m, {"tangent": np.column_stack([p, v])}, weights={"tpw": 0.5})
    fam = s.constraints["undefined gradient orthogonal constraint"]
    mom = fam["matrix"] @ m.nodes()
    u = v / np.linalg.norm(v, axis=1)[:, None]
    assert np.allclose(mom / np.linalg.norm(mom, axis=1)[:, None], u, atol=1e-12)
    assert np.allclose(fam["matrix"] @ np.ones(m.n_nodes), 0.0, atol=1e-10)
    tet = m.element_for_location(p)[2]
    cell = np.prod(m.step)
    vol = np.where(tet % 5 == 0, cell / 3, cell / 6)
    assert np.allclose(fam["w"], 0.5 * vol, rtol=1e-12)
    assert np.all(fam["b"] == 0)


def test_scaled_edge_jump_weights_average_the_two_tets():
    """LI/_p1interpolator.py:150-154: with the weight scale on, a face pair's
    weight is cgw x the mean of the two tets' node-averaged scales.  For a scale
    linear in the node coordinates that is the scale at the midpoint of the
    two tet barycentres, taken here from the brute-force face pairs."""
    m = _mesh(cells=(3, 2, 2))
    nd = m.nodes()
    a = np.array([0.5, -1.5, 2.0])
    scale = nd @ a + 3.0
    s = p1.assemble_p1(m, {}, regularisation=0.2, scale=scale)
    w = s.constraints["edge jump"]["w"]
    el = m.elements
    sets = [frozenset(e) for e in el]
    pairs = [(i, j) for i in range(len(el)) for j in range(i) if len(sets[i] & sets[j]) == 3]
    assert len(pairs) == len(w)
    bc = nd[el].mean(axis=1)
    expect = sorted(0.2 * (0.5 * (bc[i] + bc[j]) @ a + 3.0) for i, j in pairs)
    assert np.allclose(np.sort(w), expect, rtol=1e-12)
    s1 = p1.assemble_p1(m, {}, regularisation=0.2, scale=np.full(m.n_nodes, 1.5))
    assert np.allclose(s1.constraints["edge jump"]["w"], 0.3, rtol=1e-14)
