"""Explicit materialisation of matrix-free families (CPU): the regulariser
rows the package generates from a mask equal the oracle's explicit rows
(LI/_finite_difference_interpolator.py:643-692, borders :636-641, :1052-1144),
and the preprocessing equals the oracle's (LI/_solver_pipeline.py:86-124)."""
import numpy as np
import pytest
from scipy import sparse

from loopstructural_amd import explicit
from oracle import fd as ofd

GRIDS = [(4, 4, 4), (5, 7, 6), (9, 3, 8), (2, 5, 5)]


@pytest.mark.parametrize("cells", GRIDS)
def test_stencil_families_equal_oracle_rows(cells):
    og = ofd.Grid((0, 0, 0), cells, (1.0, 1.0, 1.0))
    ref = ofd.assemble_fd(og, {}, regularisation=0.1)
    nn = og.nnodes
    masks = {"dxy": ofd.DXY, "dyz": ofd.DYZ, "dxz": ofd.DXZ, "dxx": ofd.DXX, "dyy": ofd.DYY,
             "dzz": ofd.DZZ}
    for name, mask in masks.items():
        M = explicit.stencil_matrix(nn, mask, explicit.interior_centres(nn))
        if name not in ref.constraints:
            assert M.shape[0] == 0
            continue
        R = ref.constraints[name]["matrix"]
        assert M.shape == R.shape
        assert abs(M - R).max() <= 1e-15, name
    for axis, (key, mask) in enumerate((("dx", ofd.DXX), ("dy", ofd.DYY), ("dz", ofd.DZZ))):
        for upper in (False, True):
            name = f"{key}_{'upper' if upper else 'lower'}"
            M = explicit.stencil_matrix(nn, mask, explicit.layer_centres(nn, axis, upper))
            if name not in ref.constraints:
                assert M.shape[0] == 0
                continue
            R = ref.constraints[name]["matrix"]
            assert M.shape == R.shape and abs(M - R).max() <= 1e-15, name


def test_centre_sets():
    nn = (5, 6, 7)
    ids = explicit.interior_centres(nn)
    i, j, k = ids % 5, (ids // 5) % 6, ids // 30
    assert ids.shape[0] == 3 * 4 * 5 and np.all(np.diff(ids) > 0)
    assert np.all((i >= 1) & (i <= 3) & (j >= 1) & (j <= 4) & (k >= 1) & (k <= 5))
    up = explicit.layer_centres(nn, 1, True)
    assert up.shape[0] == 5 * 7 and np.all((up // 5) % 6 == 4)


def test_stack_and_preprocess_equal_oracle():
    og = ofd.Grid((0, 0, 0), (5, 4, 6), (0.2, 0.25, 1 / 6))
    rng = np.random.default_rng(2)
    p = rng.uniform(0, 1, (40, 3))
    data = {"value": np.column_stack([p, p[:, 2], np.ones(40)])}
    ref = ofd.assemble_fd(og, data, regularisation=0.2)
    A0, b0 = ref.build_matrix()
    A, b = explicit.stack(((c["matrix"], c["b"], c["w"]) for c in ref.constraints.values()),
                          og.n_nodes)
    assert abs(A - A0).max() == 0 and np.array_equal(b, b0)
    Ar, br, Sr = ofd.preprocess(A0, b0)
    Ae, be, Se = explicit.preprocess(A, b, 1e-8)
    assert Ae.shape == Ar.shape and np.array_equal(be, br)
    assert np.allclose(Se.diagonal(), Sr.diagonal(), rtol=1e-14, atol=0)
    assert abs(Ae - Ar).max() <= 1e-14
    A2, b2, S2 = explicit.preprocess(A, b, 1e-8, add_ridge=False, scaling=False)
    assert S2 is None and A2.shape == A.shape and sparse.isspmatrix_csr(A2)
