"""Explicit system of the drop-in interpolators on the GPU against the oracle:
``build_matrix`` (LI/_discrete_interpolator.py:684-755), the family matrices
``constraints[name]['matrix']`` (:383-389) and ``solve_system(solver=callable)``
(LI/_solver_strategy.py:32-45 with c = S y, LI/_discrete_interpolator.py:1288-1313)."""
import numpy as np
import pytest
from scipy import sparse
from scipy.sparse import linalg as spla

from oracle import fd as ofd
from oracle import p1 as op1

pytestmark = pytest.mark.gpu


def _fd_case():
    import loopstructural_amd as L

    rng = np.random.default_rng(21)
    p = rng.uniform(0, 1, (90, 3))
    values = np.column_stack([p, p[:, 2] + 0.1 * np.sin(6 * p[:, 0]), rng.uniform(0.5, 2, 90)])
    pn = rng.uniform(0, 1, (20, 3))
    nrm = rng.standard_normal((20, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    normals = np.column_stack([pn, nrm, np.ones(20)])
    grads = np.column_stack([rng.uniform(0, 1, (15, 3)), rng.standard_normal((15, 3)),
                             np.ones(15)])
    tang = np.column_stack([rng.uniform(0, 1, (10, 3)), rng.standard_normal((10, 3)),
                            np.ones(10)])
    iface = np.column_stack([rng.uniform(0, 1, (12, 3)), rng.integers(0, 2, 12), np.ones(12)])
    data = {"value": values, "normal": normals, "gradient": grads, "tangent": tang,
            "interface": iface}
    cells = (7, 6, 8)
    step = (1 / 7, 1 / 6, 1 / 8)
    grid = L.StructuredGrid(origin=np.zeros(3), nsteps=np.array(cells), step_vector=np.array(step))
    it = L.FiniteDifferenceInterpolator(grid)
    it.set_value_constraints(values)
    it.set_normal_constraints(normals)
    it.set_gradient_constraints(grads.copy())
    it.set_tangent_constraints(tang.copy())
    it.set_interface_constraints(iface)
    kw = dict(regularisation=0.3, cpw=2.0, gpw=0.7, dx=0.5)
    it.setup_interpolator(**kw)
    w = ofd.default_weights()
    w.update({"cpw": 2.0, "gpw": 0.7, "dx": 0.5})
    og = ofd.Grid((0, 0, 0), cells, step)
    ref = ofd.assemble_fd(og, {k: v.copy() for k, v in data.items()}, weights=w,
                          regularisation=0.3)
    return it, ref


def test_fd_build_matrix_equals_oracle():
    it, ref = _fd_case()
    assert list(it.constraints) == list(ref.constraints)
    for name in ref.constraints:
        M = it.constraints[name]["matrix"]
        R = ref.constraints[name]["matrix"]
        assert sparse.issparse(M) and M.shape == R.shape, name
        assert (abs(M - R).max() if R.nnz else 0.0) <= 1e-13, name
        assert np.allclose(it.constraints[name]["b"], ref.constraints[name]["b"], rtol=0,
                           atol=1e-13), name
        assert np.array_equal(it.constraints[name]["w"], ref.constraints[name]["w"]), name
    A, b = it.build_matrix()
    A0, b0 = ref.build_matrix()
    assert A.shape == A0.shape and abs(A - A0).max() <= 1e-13
    assert np.allclose(b, b0, rtol=0, atol=1e-13)


def test_fd_callable_solver_reaches_minimiser():
    it, ref = _fd_case()
    seen = {}

    def normal_spsolve(A, b):
        seen["shape"] = A.shape
        seen["csr"] = sparse.isspmatrix_csr(A)
        return spla.spsolve((A.T @ A).tocsc(), A.T @ b)

    assert it.solve_system(solver=normal_spsolve)
    cx = ofd.solve_exact(ref)
    assert np.linalg.norm(it.c - cx) <= 1e-9 * np.linalg.norm(cx)
    A0, _ = ref.build_matrix()
    assert seen["shape"] == (A0.shape[0] + A0.shape[1], A0.shape[1]) and seen["csr"]
    t = it.get_last_solve_timing()
    assert t["solver"] == "callable" and t["up_to_date"]
    # without column scaling the callable sees the unscaled system, same minimiser
    it.apply_scaling_matrix = False
    assert it.solve_system(solver=normal_spsolve)
    assert np.linalg.norm(it.c - cx) <= 1e-8 * np.linalg.norm(cx)


def test_fd_weight_scaled_rows_in_build_matrix():
    import loopstructural_amd as L

    rng = np.random.default_rng(4)
    p = rng.uniform(0, 1, (40, 3))
    values = np.column_stack([p, p[:, 0], np.ones(40)])
    pn = rng.uniform(0, 1, (10, 3))
    normals = np.column_stack([pn, np.tile([0.0, 0.0, 1.0], (10, 1)), np.ones(10)])
    grid = L.StructuredGrid(origin=np.zeros(3), nsteps=np.array([6, 6, 6]),
                            step_vector=np.full(3, 1 / 6))
    it = L.FiniteDifferenceInterpolator(grid)
    it.set_value_constraints(values)
    it.set_normal_constraints(normals)
    it.setup_interpolator(regularisation=0.2, use_regularisation_weight_scale=True,
                          regularisation_weight_sigma=0.2)
    og = ofd.Grid((0, 0, 0), (6, 6, 6), (1 / 6,) * 3)
    scale = ofd.regularisation_scale(og, normals, sigma=0.2)
    ref = ofd.assemble_fd(og, {"value": values.copy(), "normal": normals.copy()},
                          regularisation=0.2, scale=scale)
    A, b = it.build_matrix()
    A0, b0 = ref.build_matrix()
    assert A.shape == A0.shape and abs(A - A0).max() <= 1e-12


def test_p1_build_matrix_and_callable():
    from loopstructural_amd import P1Interpolator, TetMesh

    m = op1.TetMesh((0, 0, 0), (5, 4, 6), (0.2, 0.25, 1 / 6))
    rng = np.random.default_rng(8)
    ext = m.ncells * m.step
    pv = rng.uniform(0, 1, (60, 3)) * ext
    data = {"value": np.column_stack([pv, pv[:, 2], np.ones(60)]),
            "normal": np.column_stack([rng.uniform(0, 1, (10, 3)) * ext,
                                       np.tile([0.0, 0.0, 1.0], (10, 1)), np.ones(10)]),
            "tangent": np.zeros((0, 7))}
    mesh = TetMesh(origin=m.origin, nsteps_cells=m.ncells, step_vector=m.step)
    it = P1Interpolator(mesh)
    it.set_value_constraints(data["value"])
    it.set_normal_constraints(data["normal"])
    it.setup_interpolator(regularisation=0.1)
    ref = op1.assemble_p1(m, {k: v.copy() for k, v in data.items()}, regularisation=0.1)
    A, b = it.build_matrix()
    A0, b0 = ref.build_matrix()
    assert A.shape == A0.shape
    # edge-jump rows may differ in sign from the oracle's (face node order);
    # the normal equations do not
    N, N0 = (A.T @ A).tocsr(), (A0.T @ A0).tocsr()
    assert abs(N - N0).max() <= 1e-12 * abs(N0).max()
    assert np.allclose(A.T @ b, A0.T @ b0, rtol=0, atol=1e-12)
    assert it.solve_system(solver=lambda A, b: spla.lsqr(A, b, atol=1e-15, btol=1e-15,
                                                         iter_lim=200000)[0])
    cx = ofd.solve_exact(ref)
    assert np.linalg.norm(it.c - cx) <= 1e-8 * np.linalg.norm(cx)
