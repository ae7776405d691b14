"""Pins the CPU oracle of the FD path (runs without a GPU).

Checks come from the reference's own tests (cited), closed forms and
invariants chosen so that a dropped term, a wrong sign or index, or a
transposed operand in the oracle fails at least one of them.
"""
import numpy as np
import pytest
from scipy import sparse
from scipy.sparse import linalg as spla

from oracle import fd as ofd


def grid10():
    return ofd.Grid((0, 0, 0), (10, 10, 10), (1, 1, 1))


def barycentres(g):
    ci = np.stack(np.meshgrid(*[np.arange(n) for n in g.ncells], indexing="ij"), -1)
    ci = ci.reshape(-1, 3, order="F")
    return g.origin + g.step * ci + 0.5 * g.step


# ---------------------------------------------- reference known answers ----
def test_local_coordinates_kat():
    """tests/unit/interpolator/test_discrete_supports.py:106-112."""
    lc = ofd.local_coords(grid10(), np.array([[1.2, 1.5, 1.7]]))
    assert np.allclose(lc, [[0.2, 0.5, 0.7]])


def test_evaluate_value_of_x_at_barycentres_is_exact():
    """test_discrete_supports.py:33-43: sum(bary_x - evaluate(x)) == 0."""
    g = grid10()
    b = barycentres(g)
    assert np.sum(b[:, 0] - ofd.evaluate_value(g, g.nodes()[:, 0], b)) == 0


@pytest.mark.parametrize("steps", [10, 20, 100])
def test_evaluate_gradient_of_y_is_exact(steps):
    """test_discrete_supports.py:45-50."""
    g = ofd.Grid((0, 0, 0), (steps,) * 3, (1, 1, 1))
    b = barycentres(g)
    v = ofd.evaluate_gradient(g, g.nodes()[:, 1], b)
    assert np.sum(v - np.array([0, 1, 0])) == 0


@pytest.mark.parametrize("seed", range(5))
def test_gradient_of_linear_field_random_step(seed):
    """test_discrete_supports.py:83-94."""
    rng = np.random.default_rng(seed)
    step = rng.uniform(0, 100)
    g = ofd.Grid((0, 0, 0), (10, 10, 10), (step, step, step))
    n = rng.random(3)
    n /= np.linalg.norm(n)
    dist = g.nodes() @ n
    v = ofd.evaluate_gradient(g, dist, rng.uniform(1, 8, size=(100, 3)))
    assert np.all(np.isclose(np.sum(v - n[None, :], axis=1), 0, atol=1e-3, rtol=1e-3))


def test_boundary_vertices_and_faces_finite_and_exact():
    """packages/loop_common/tests/test_structured_grid_boundary_eval.py:5-41."""
    g = grid10()
    nodes = g.nodes()
    vals = nodes[:, 0] + 2.0 * nodes[:, 1] + 3.0 * nodes[:, 2]
    o, m = g.origin, g.maximum
    corners = np.array([[o[0], o[1], o[2]], [m[0], o[1], o[2]], [o[0], m[1], o[2]],
                        [o[0], o[1], m[2]], [m[0], m[1], m[2]]])
    ev = ofd.evaluate_value(g, vals, corners)
    assert np.all(np.isfinite(ev))
    assert np.allclose(ev, corners @ [1.0, 2.0, 3.0])
    vals = nodes[:, 0] - 0.5 * nodes[:, 1] + 0.25 * nodes[:, 2]
    pts = np.array([[m[0], 2.5, 4.5], [m[0], 0.0, 0.0], [m[0], m[1], 7.25]])
    ev = ofd.evaluate_value(g, vals, pts)
    assert np.all(np.isfinite(ev))
    assert np.allclose(ev, pts @ [1.0, -0.5, 0.25])


def test_outside_handling():
    """test_discrete_supports.py:58-80 and the NaN/outside conventions of
    LI/_discrete_interpolator.py:1356-1362, _3d_structured_grid.py:318-322."""
    g = grid10()
    b = barycentres(g)
    ins = ofd.inside(g, b + 5)
    assert np.all(~ins == np.any((b + 5) > g.maximum, axis=1))
    idc, ins = ofd.cell_corners(g, b - 5)
    assert np.all(idc[ins] < g.n_nodes) and np.all(idc[~ins] == -1)
    c = np.arange(g.n_nodes, dtype=float)
    pts = np.array([[-1.0, 5, 5], [np.nan, 1, 1], [5, 5, 5]])
    v = ofd.evaluate_value(g, c, pts)
    assert np.isnan(v[0]) and v[1] == 0.0 and np.isfinite(v[2])
    gr = ofd.evaluate_gradient(g, c, pts)
    assert np.all(gr[0] == 0) and np.all(gr[1] == 0)


def test_points_on_max_face_use_last_cell():
    g = grid10()
    ci, ins = ofd.cell_index(g, np.array([[10.0, 10.0, 10.0], [0.0, 0.0, 0.0]]))
    assert ins.all() and np.array_equal(ci, [[9, 9, 9], [0, 0, 0]])
    lc = ofd.local_coords(g, np.array([[10.0, 10.0, 10.0]]))
    assert np.array_equal(lc, [[1.0, 1.0, 1.0]])


def test_mask_kats():
    """tests/unit/interpolator/test_operator.py."""
    for m in (ofd.DXX, ofd.DYY, ofd.DZZ, ofd.DXY, ofd.DXZ, ofd.DYZ):
        assert m.shape == (3, 3, 3) and np.isclose(m.sum(), 0.0)
    assert np.array_equal(ofd.DYY, ofd.DXX.swapaxes(1, 2))
    assert ofd.DXX[1, 1, 0] == 1 and ofd.DXX[1, 1, 1] == -2 and ofd.DXX[1, 1, 2] == 1


def test_border_row_counts():
    """packages/loop_interpolation/tests/test_discrete_interpolator.py:22-51."""
    g = ofd.Grid((0, 0, 0), (4, 4, 4), (1, 1, 1))
    w = {k: 0.0 for k in ("dxy", "dyz", "dxz", "dxx", "dyy", "dzz", "cpw", "gpw", "npw",
                          "tpw", "ipw")}
    w.update(dx=1.0, dy=1.0, dz=1.0)
    s = ofd.assemble_fd(g, {}, weights=w)
    for name in ["dx_lower", "dx_upper", "dy_lower", "dy_upper", "dz_lower", "dz_upper"]:
        assert s.constraints[name]["matrix"].shape[0] == g.nnodes[1] * g.nnodes[2]
    assert s.constraints["dxx"]["matrix"].shape[0] == 3 ** 3


# ------------------------------------------------- closed forms ------------
def test_trilinear_partition_of_unity_and_linear_reproduction():
    rng = np.random.default_rng(0)
    lc = rng.uniform(0, 1, (500, 3))
    w = ofd.trilinear(lc)
    assert np.allclose(w.sum(1), 1.0)
    corners = np.stack([ofd.CORNER_DX, ofd.CORNER_DY, ofd.CORNER_DZ], 1)
    assert np.allclose(w @ corners, lc)


def test_element_gradient_of_linear_field_is_exact():
    g = ofd.Grid((0.3, -1, 2), (5, 6, 7), (0.5, 0.25, 2.0))
    rng = np.random.default_rng(1)
    p = rng.uniform(g.origin, g.maximum, (200, 3))
    a = np.array([1.5, -0.7, 0.2])
    idc, _ = ofd.cell_corners(g, p)
    T = ofd.element_gradient(g, p)
    f = g.nodes() @ a
    assert np.allclose(np.einsum("ndk,nk->nd", T, f[idc]), a)


@pytest.mark.parametrize("fam,mono,expected", [
    # normalised second-derivative rows applied to monomials, h = 1:
    # Dxx x^2 = 2/sqrt6 ; Dxy xy = -2 ; Dxy xz = 0 (a transposed family fails)
    ("dxx", lambda P: P[:, 0] ** 2, 2 / np.sqrt(6)), ("dyy", lambda P: P[:, 1] ** 2, 2 / np.sqrt(6)),
    ("dzz", lambda P: P[:, 2] ** 2, 2 / np.sqrt(6)), ("dxx", lambda P: P[:, 1] ** 2, 0.0),
    ("dxy", lambda P: P[:, 0] * P[:, 1], -2.0), ("dxz", lambda P: P[:, 0] * P[:, 2], -2.0),
    ("dyz", lambda P: P[:, 1] * P[:, 2], -2.0), ("dxy", lambda P: P[:, 0] * P[:, 2], 0.0),
    ("dyz", lambda P: P[:, 0] * P[:, 1], 0.0),
])
def test_regulariser_rows_on_monomials(fam, mono, expected):
    g = ofd.Grid((0, 0, 0), (5, 4, 6), (1, 1, 1))
    w = ofd.default_weights()
    s = ofd.assemble_fd(g, {}, weights=w)
    M = s.constraints[fam]["matrix"]
    r = M @ mono(g.nodes())
    assert np.allclose(r, expected, atol=1e-12)
    # every row is unit-norm and sums to zero (annihilates constants)
    assert np.allclose(sparse.linalg.norm(M, axis=1), 1.0)
    assert np.allclose(M @ np.ones(g.n_nodes), 0.0)


def test_regulariser_weights_and_order():
    """Family order, cross weights /4, border weight dx independent of
    `regularisation` (LI/_finite_difference_interpolator.py:160-167,
    LC/supports/_3d_structured_grid.py:489-510)."""
    g = ofd.Grid((0, 0, 0), (4, 4, 4), (1, 1, 1))
    s = ofd.assemble_fd(g, {}, regularisation=0.3)
    assert list(s.constraints) == ["dxy", "dyz", "dxz", "dxx", "dyy", "dzz", "dx_lower",
                                   "dx_upper", "dy_lower", "dy_upper", "dz_lower", "dz_upper"]
    assert np.allclose(s.constraints["dxy"]["w"], 0.3 / 4)
    assert np.allclose(s.constraints["dxx"]["w"], 0.3)
    assert np.allclose(s.constraints["dx_lower"]["w"], 1.0)


def test_normal_equations_annihilate_linear_fields_without_data():
    g = ofd.Grid((0, 0, 0), (6, 5, 7), (0.2, 0.3, 0.1))
    s = ofd.assemble_fd(g, {}, regularisation=0.7)
    N, _ = ofd.normal_equations(s, add_ridge=False)
    lin = g.nodes() @ [0.3, -2.0, 1.0] + 4.0
    assert np.allclose(N @ lin, 0.0, atol=1e-10)
    assert abs(N - N.T).max() < 1e-14


def test_get_vectors_orthonormal_basis():
    rng = np.random.default_rng(3)
    n = rng.standard_normal((400, 3))
    n /= np.linalg.norm(n, axis=1)[:, None]
    nn = n.copy()
    s, d = ofd.get_vectors(nn)
    s, d = s.T, d.T
    assert np.allclose(np.sum(s * n, 1), 0, atol=1e-12)
    assert np.allclose(np.sum(d * n, 1), 0, atol=1e-12)
    assert np.allclose(np.sum(s * d, 1), 0, atol=1e-12)
    assert np.allclose(s[:, 2], 0.0)
    assert np.allclose(np.linalg.norm(s, axis=1), 1) and np.allclose(np.linalg.norm(d, axis=1), 1)
    s0, _ = ofd.get_vectors(np.array([[0.0, 0.0, 1.0]]))
    assert np.allclose(s0[:, 0], [0.0, -1.0, 0.0])


def test_value_and_normal_rows_closed_form():
    """Value rows are the normalised trilinear weights with b = v/|N|;
    normal rows are T_d/|T_d| with b = n_d/|T_d|; weight npw/3 * w_pt."""
    g = ofd.Grid((0, 0, 0), (4, 4, 4), (0.25, 0.25, 0.25))
    p = np.array([[0.3, 0.6, 0.1]])
    s = ofd.assemble_fd(g, {"value": np.array([[0.3, 0.6, 0.1, 2.0, 0.5]]),
                            "normal": np.array([[0.3, 0.6, 0.1, 0.0, 0.6, 0.8, 2.0]])})
    w = ofd.trilinear(ofd.local_coords(g, p))[0]
    row = s.constraints["value"]["matrix"].toarray()[0]
    idc, _ = ofd.cell_corners(g, p)
    assert np.allclose(row[idc[0]], w / np.linalg.norm(w))
    assert np.isclose(s.constraints["value"]["b"][0], 2.0 / np.linalg.norm(w))
    assert np.isclose(s.constraints["value"]["w"][0], 0.5)
    T = ofd.element_gradient(g, p)[0]
    for d in range(3):
        r = s.constraints[f"norm_{d}"]["matrix"].toarray()[0][idc[0]]
        assert np.allclose(r, T[d] / np.linalg.norm(T[d]))
        assert np.isclose(s.constraints[f"norm_{d}"]["w"][0], 10.0 / 3 * 2.0)


# Rows built from shape-function gradients are pinned through their first
# moments: for a trilinear element, sum_k dN_k/dx_d X_k = e_d exactly, so a row
# a = v.T (normalised) maps the node coordinates X to v / |v.T| and sums to 0.
# The direction v comes from an independent closed form below; a wrong sign,
# a swapped axis or a strike/dip mix-up changes that direction.
def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _closed_form_strike_dip(v):
    """Strike = (n_y, -n_x, 0)/|.| ((0,-1,0) for a vertical normal) and
    dip = strike x n normalised, n = v/|v| (LC/math/_maths.py:126-160,
    :276-332 reduce to these)."""
    n = _unit(v)
    r = np.hypot(n[:, 0], n[:, 1])
    s = np.zeros_like(n)
    s[:, 1] = -1.0
    m = r > 0
    s[m] = np.column_stack([n[m, 1], -n[m, 0], np.zeros(m.sum())]) / r[m, None]
    return s, _unit(np.cross(s, n))


def _grid_and_points(seed, n):
    g = ofd.Grid((0.5, -1.0, 0.2), (5, 4, 6), (0.3, 0.5, 0.2))
    rng = np.random.default_rng(seed)
    p = g.origin + rng.uniform(0.02, 0.98, (n, 3)) * g.ncells * g.step
    return g, p, rng


def test_gradient_rows_are_strike_and_dip_derivatives():
    """LI/_finite_difference_interpolator.py:342-401: two families
    ('gradient' = strike, then 'gradient_0' = dip), rows (v.T)/|v.T| with
    the input vector's length cancelled, b = 0, weight gpw."""
    g, p, rng = _grid_and_points(21, 40)
    v = rng.standard_normal((40, 3)) * rng.uniform(0.2, 5.0, (40, 1))
    v[0] = [0.0, 0.0, 2.5]
    v[1] = [0.0, 0.0, -0.3]
    v[2] = [1.0, 0.0, 0.0]
    s = ofd.assemble_fd(g, {"gradient": np.column_stack([p, v, np.ones(40)])},
                        weights={"gpw": 3.0})
    assert list(s.constraints)[:2] == ["gradient", "gradient_0"]
    strike, dip = _closed_form_strike_dip(v)
    T = ofd.element_gradient(g, p)
    X = g.nodes()
    for fam, vec in (("gradient", strike), ("gradient_0", dip)):
        M = s.constraints[fam]["matrix"]
        mom = M @ X
        assert np.allclose(_unit(mom), vec, atol=1e-12)
        # scale 1/|v.T| (T pinned independently), constants annihilated
        tn = np.linalg.norm(np.einsum("nd,ndk->nk", vec, T), axis=1)
        assert np.allclose(np.linalg.norm(mom, axis=1) * tn, 1.0, rtol=1e-12)
        assert np.allclose(M @ np.ones(g.n_nodes), 0.0, atol=1e-13)
        assert np.allclose(sparse.linalg.norm(M, axis=1), 1.0)
        assert np.all(s.constraints[fam]["b"] == 0)
        assert np.all(s.constraints[fam]["w"] == 3.0)
    # a linear field whose gradient is parallel to the normal satisfies both rows
    f = X @ v[5]
    assert np.allclose(s.constraints["gradient"]["matrix"][5] @ f, 0.0, atol=1e-12)
    assert np.allclose(s.constraints["gradient_0"]["matrix"][5] @ f, 0.0, atol=1e-12)


def test_tangent_rows_and_zero_vectors():
    """LI/_finite_difference_interpolator.py:478-539: one row t_hat.T/|t_hat.T|,
    b = 0, weight tpw; a zero vector gives an all-zero row (normalisation
    skips it, LI/_discrete_interpolator.py:361-366)."""
    g, p, rng = _grid_and_points(22, 30)
    v = rng.standard_normal((30, 3)) * rng.uniform(0.2, 5.0, (30, 1))
    v[3] = 0.0
    s = ofd.assemble_fd(g, {"tangent": np.column_stack([p, v, np.ones(30)])},
                        weights={"tpw": 0.7})
    M = s.constraints["gradient orthogonal"]["matrix"]
    mom = M @ g.nodes()
    ok = np.ones(30, dtype=bool)
    ok[3] = False
    assert np.allclose(_unit(mom[ok]), _unit(v[ok]), atol=1e-12)
    T = ofd.element_gradient(g, p)
    tn = np.linalg.norm(np.einsum("nd,ndk->nk", _unit(v[ok]), T[ok]), axis=1)
    assert np.allclose(np.linalg.norm(mom[ok], axis=1) * tn, 1.0, rtol=1e-12)
    assert M[3].nnz == 0 or np.all(M[3].toarray() == 0)
    assert np.allclose(M @ np.ones(g.n_nodes), 0.0, atol=1e-13)
    assert np.all(s.constraints["gradient orthogonal"]["w"] == 0.7)


@pytest.mark.parametrize("n0,n1", [(7, 3), (200, 201)])
def test_interface_rows_pair_differences(n0, n1):
    """LI/_finite_difference_interpolator.py:271-340: per id, rows
    [N_i, -N_j]/|.| over the pairs i<j of that id when it has at most 200
    points, else the star (0, j); b = 0, weight ipw; a NaN id and a lone
    point add nothing."""
    g, p, rng = _grid_and_points(23, n0 + n1 + 2)
    ids = np.concatenate([np.zeros(n0), np.ones(n1), [np.nan, 5.0]])
    s = ofd.assemble_fd(g, {"interface": np.column_stack([p, ids, np.ones(len(ids))])},
                        weights={"ipw": 2.0})
    names = [k for k in s.constraints if k.startswith("interface")]
    assert names == ["interface_0.0", "interface_1.0"]
    X = g.nodes()
    W = ofd.trilinear(ofd.local_coords(g, p))
    for name, sel, n in (("interface_0.0", ids == 0, n0), ("interface_1.0", ids == 1, n1)):
        if n > 200:
            i, j = np.zeros(n - 1, dtype=int), np.arange(1, n)
        else:
            i, j = np.triu_indices(n, k=1)
        M = s.constraints[name]["matrix"]
        assert M.shape[0] == len(i) == (n * (n - 1) // 2 if n <= 200 else n - 1)
        ps, ws = p[sel], W[sel]
        rn = np.sqrt(np.sum(ws[i] ** 2, 1) + np.sum(ws[j] ** 2, 1))
        # trilinear rows reproduce positions: row . X = (p_i - p_j)/|row|
        assert np.allclose(M @ X, (ps[i] - ps[j]) / rn[:, None], atol=1e-12)
        f = X @ np.array([0.3, -1.2, 2.0]) + 4.0
        assert np.allclose(M @ f, ((ps[i] - ps[j]) @ [0.3, -1.2, 2.0]) / rn, atol=1e-12)
        assert np.allclose(M @ np.ones(g.n_nodes), 0.0, atol=1e-13)
        assert np.all(s.constraints[name]["b"] == 0) and np.all(s.constraints[name]["w"] == 2.0)


# ---------------------------------------------------------------- solve ----
def test_exact_solution_is_the_dense_least_squares_minimiser():
    g = ofd.Grid((0, 0, 0), (3, 3, 3), (1 / 3,) * 3)
    rng = np.random.default_rng(5)
    pv = rng.uniform(0, 1, (20, 3))
    s = ofd.assemble_fd(g, {"value": np.column_stack([pv, rng.standard_normal(20),
                                                      np.ones(20)])}, regularisation=0.2)
    A, b = s.build_matrix()
    A, b, _ = ofd.preprocess(A, b, scaling=False)
    x_lstsq = np.linalg.lstsq(A.toarray(), b, rcond=None)[0]
    x = ofd.solve_exact(s)
    assert np.allclose(x, x_lstsq, rtol=1e-9, atol=1e-12)


def test_jacobi_pcg_equals_scipy_cg_on_scaled_system():
    """The GPU solver's iterate sequence: Jacobi-PCG on N == scipy cg on the
    column-scaled system (LI/_solver_strategy.py:48-71 + c = S y)."""
    g = ofd.Grid((0, 0, 0), (6, 6, 6), (1 / 6,) * 3)
    rng = np.random.default_rng(9)
    pv = rng.uniform(0, 1, (60, 3))
    f, gr = ofd.field_c1(pv)
    s = ofd.assemble_fd(g, {"value": np.column_stack([pv, f, np.ones(60)]),
                            "normal": np.column_stack([pv[:20], gr[:20], np.ones(20)])},
                        regularisation=0.1)
    N, rhs = ofd.normal_equations(s)
    for rtol in (1e-3, 1e-6, 1e-9):
        _, info, it_ref = ofd.solve_cg(s, rtol=rtol, return_info=True)
        _, it, ok = ofd.jacobi_pcg(N, rhs, rtol=rtol)
        assert ok and info == 0
        assert abs(it - it_ref) <= 1
    # identical iterates after k steps (before rounding differences grow)
    for k in (1, 3, 10):
        c_ref = ofd.solve_cg(s, rtol=1e-300, maxiter=k)
        c, it, _ = ofd.jacobi_pcg(N, rhs, rtol=1e-300, maxiter=k)
        assert it == k
        assert np.linalg.norm(c - c_ref) <= 1e-10 * np.linalg.norm(c_ref)


def test_lsmr_and_cg_reach_the_same_minimiser():
    g = ofd.Grid((0, 0, 0), (5, 5, 5), (0.2,) * 3)
    rng = np.random.default_rng(2)
    pv = rng.uniform(0, 1, (50, 3))
    s = ofd.assemble_fd(g, {"value": np.column_stack([pv, pv[:, 2], np.ones(50)])},
                        regularisation=0.1)
    cx = ofd.solve_exact(s)
    assert np.allclose(ofd.solve_lsmr(s, tol=1e-14, maxiter=100000), cx, atol=1e-6)
    assert np.allclose(ofd.solve_cg(s, rtol=1e-13), cx, atol=1e-8)


def test_regularisation_scale_against_brute_force_distances():
    """compute_weighting (LI/_finite_difference_interpolator.py:66-88, a scipy
    KDTree query) equals 1 + alpha exp(-d^2 / 2 sigma^2) with d the brute-force
    nearest distance; the scale adds it to ones for the inside normal points
    only, sigma defaulting to 10 x the node count on x (:436-444)."""
    g = ofd.Grid((0.0, 0.0, 0.0), (4, 5, 3), (0.25, 0.2, 0.3))
    rng = np.random.default_rng(24)
    pts = rng.uniform(-0.2, 1.1, (17, 3))
    X = g.nodes()
    d2 = ((X[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    for sigma in (0.15, 2.0):
        w = ofd.compute_weighting(X, pts, alpha=3.0, sigma=sigma)
        assert np.allclose(w, 1 + 3.0 * np.exp(-d2.min(1) / (2 * sigma ** 2)), rtol=1e-13)
    normals = np.column_stack([pts, np.ones((17, 3))])
    ins = ofd.inside(g, pts)
    assert 0 < ins.sum() < 17
    sigma = g.nnodes[0] * 10
    ref = 2 + 10.0 * np.exp(-d2[:, ins].min(1) / (2 * sigma ** 2))
    assert np.allclose(ofd.regularisation_scale(g, normals), ref, rtol=1e-13)
