"""The oracle's rectilinear-grid restatement pinned on the CPU by the
reference's own RectilinearGrid known-answer tests
(packages/loop_common/tests/test_rectilinear_grid.py) and by the uniform-grid
oracle on a uniform rectilinear grid."""
import numpy as np

from oracle import fd as ofd
from oracle import rect as orect

NONUNIFORM = (np.array([0.0, 0.2, 0.5, 1.0, 2.0, 4.0, 5.0]), np.array([0.0, 0.3, 0.7, 1.5, 3.0]),
              np.array([0.0, 0.25, 0.75, 2.0]))


def _barycentres(g):
    ci = np.stack(np.meshgrid(*[np.arange(n - 1) for n in g.nnodes], indexing="ij"), -1)
    ci = ci.reshape(-1, 3)
    return np.column_stack([0.5 * (a[ci[:, d]] + a[ci[:, d] + 1]) for d, a in enumerate(g.axis)]), ci


def test_cell_index_and_local_coordinates_kats():
    g = orect.RectGrid([0.0, 1.0, 3.0, 6.0], [0.0, 2.0, 5.0], [0.0, 1.5, 4.0])
    bc, ci = _barycentres(g)
    idx, ins = g.cell_index(bc)
    assert ins.all() and np.array_equal(idx, ci)
    g = orect.RectGrid([0.0, 1.0, 3.0], [0.0, 2.0], [0.0, 0.5, 1.5])
    assert np.allclose(g.local_coords(np.array([[0.0, 0.0, 0.0]])), 0.0)
    assert np.allclose(g.local_coords(np.array([[3.0, 2.0, 1.5]])), 1.0)
    g = orect.RectGrid([0.0, 2.0, 6.0], [0.0, 1.0], [0.0, 1.0])
    assert np.isclose(g.local_coords(np.array([[4.0, 0.5, 0.5]]))[0, 0], 0.5)
    g = orect.RectGrid(*NONUNIFORM)
    out = np.array([[-1.0, 1.0, 1.0], [6.0, 1.0, 1.0], [1.0, -1.0, 1.0], [1.0, 4.0, 1.0]])
    assert not g.inside(out).any()
    assert not g.inside(g.nodes()[:1]).any()   # strict: the origin is outside


def test_linear_fields_are_exact():
    g = orect.RectGrid(*NONUNIFORM)
    bc, _ = _barycentres(g)
    rng = np.random.default_rng(0)
    p = g.origin + rng.uniform(0.01, 0.99, (200, 3)) * (g.maximum - g.origin)
    for axis in range(3):
        c = g.nodes()[:, axis]
        assert np.allclose(g.evaluate_value(c, bc), bc[:, axis], atol=1e-10)
        grad = g.evaluate_gradient(c, p)
        e = np.zeros(3)
        e[axis] = 1.0
        assert np.allclose(grad, e[None, :], atol=1e-10)
    f = g.nodes() @ np.array([0.3, -1.2, 2.0])
    assert np.allclose(g.evaluate_value(f, p), p @ np.array([0.3, -1.2, 2.0]), atol=1e-10)


def test_scaled_rows_kats():
    g = orect.RectGrid(*NONUNIFORM)
    for axis in range(3):
        A, cols, rows = g.scaled_rows(axis)
        assert A.shape[1] == 3 and np.allclose(A.sum(axis=1), 0.0, atol=1e-12)
        assert np.all((cols >= 0) & (cols < g.n_nodes))
        # a quadratic along the axis has second derivative 2 at every centre
        x = g.nodes()[:, axis]
        assert np.allclose(np.sum(A * (x ** 2)[cols], axis=1), 2.0, atol=1e-9)
    for ax, cx in ((0, 1), (0, 2), (1, 2)):
        A, cols, rows = g.scaled_rows(ax, cx)
        assert A.shape[1] == 4 and np.allclose(A.sum(axis=1), 0.0, atol=1e-12)
        x = g.nodes()
        assert np.allclose(np.sum(A * (x[:, ax] * x[:, cx])[cols], axis=1), 1.0, atol=1e-9)
    h = 0.5
    g = orect.RectGrid([0.0, h, 2 * h, 3 * h], [0.0, h], [0.0, h])
    A, _, _ = g.scaled_rows(0)
    assert np.allclose(A, np.array([1 / h ** 2, -2 / h ** 2, 1 / h ** 2])[None, :], atol=1e-10)


def test_uniform_rectilinear_data_rows_equal_the_uniform_oracle():
    cells = (5, 4, 6)
    og = ofd.Grid((0, 0, 0), cells, (0.2, 0.25, 1 / 6))
    g = orect.RectGrid(*[np.linspace(0, 1, n + 1) for n in cells])
    rng = np.random.default_rng(2)
    p = rng.uniform(0.01, 0.99, (50, 3))
    vals = np.column_stack([p, p[:, 2], np.ones(50)])
    nrm = np.column_stack([rng.uniform(0.01, 0.99, (10, 3)), np.tile([0, 0, 1.0], (10, 1)),
                           np.ones(10)])
    s = orect.assemble(g, {"value": vals, "normal": nrm}, regularisation=0.1)
    r = ofd.assemble_fd(og, {"value": vals.copy(), "normal": nrm.copy()}, regularisation=0.1)
    for name in ("value", "norm_0", "norm_1", "norm_2", "dx_lower", "dz_upper"):
        D = s.constraints[name]["matrix"] - r.constraints[name]["matrix"]
        assert (abs(D).max() if D.nnz else 0.0) <= 1e-12, name
    # interior families: the same normalised rows; the rectilinear centres
    # also cover the faces along the other axes
    for name in ("dxx", "dxy"):
        M, R = s.constraints[name]["matrix"], r.constraints[name]["matrix"]
        assert M.shape[0] > R.shape[0]
        Md = {tuple(np.nonzero(M[i].toarray()[0])[0]): M[i].toarray()[0] for i in range(M.shape[0])}
        for i in range(R.shape[0]):
            row = R[i].toarray()[0]
            key = tuple(np.nonzero(row)[0])
            assert key in Md and np.allclose(np.abs(Md[key]), np.abs(row), atol=1e-12)
