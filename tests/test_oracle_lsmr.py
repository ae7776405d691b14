"""The GPU LSMR works on a rectangular operator whose regulariser block is
not the reference's row for row: rows of one mask centred at one node are
merged into one row of the summed squared weight.  LSMR's iterates depend on
A only through A^T A, A^T b and |b| (Golub-Kahan: beta_1 = |b|,
alpha_k v_k = A^T u_k - beta_k v_{k-1}, and every inner product of u's
reduces to one of v's through A^T A), so any A' with the same three gives the
same iterates in exact arithmetic.  Checked here with scipy's lsmr on the
oracle system against (1) the system with the regulariser block replaced by
its R factor and (2) the rows permuted: identical to roundoff over a short
fixed run (istop 7); at the default tolerances the stopping code matches and
the stopping iteration moves by at most one (rounding of a different row
order is amplified by the loss of orthogonality over hundreds of steps).  A
changed row weight moves the iterates far beyond roundoff."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import fd as ofd


def _system(cells=6, seed=3):
    from loopstructural_amd import synthetic

    values, normals = synthetic.constraints(60, 15, seed)
    og = ofd.Grid((0, 0, 0), (cells,) * 3, (1.0 / cells,) * 3)
    sysm = ofd.assemble_fd(og, {"value": values, "normal": normals}, regularisation=0.1)
    A, b = sysm.build_matrix()
    A, b, S = ofd.preprocess(A, b, ridge=1e-8)
    return sp.csr_matrix(A), np.asarray(b, dtype=float)


def _split(A, b):
    data = np.flatnonzero(b != 0)
    reg = np.flatnonzero(b == 0)
    return data, reg


def _lsmr(A, b, **kw):
    out = spla.lsmr(A, b, **kw)
    return out[0], out[1], out[2]


def test_lsmr_iterates_depend_on_normal_equations_only():
    A, b = _system()
    data, reg = _split(A, b)
    Ad = A[data]
    R = np.linalg.qr(A[reg].toarray(), mode="r")
    A2 = sp.vstack([Ad, sp.csr_matrix(R)]).tocsr()
    b2 = np.concatenate([b[data], np.zeros(R.shape[0])])
    assert np.allclose((A2.T @ A2).toarray(), (A.T @ A).toarray(), rtol=0, atol=1e-12)
    kw = {"atol": 0.0, "btol": 0.0, "conlim": 0.0, "maxiter": 25}
    x1, s1, i1 = _lsmr(A, b, **kw)
    x2, s2, i2 = _lsmr(A2, b2, **kw)
    assert (s1, i1) == (s2, i2) == (7, 25)
    assert np.linalg.norm(x1 - x2) <= 1e-10 * np.linalg.norm(x1)
    x1, s1, i1 = _lsmr(A, b)
    x2, s2, i2 = _lsmr(A2, b2)
    assert s1 == s2 and abs(i1 - i2) <= 1


def test_lsmr_row_order_and_weight_change_detected():
    A, b = _system(seed=4)
    rng = np.random.default_rng(0)
    p = rng.permutation(A.shape[0])
    kw = {"atol": 0.0, "btol": 0.0, "conlim": 0.0, "maxiter": 25}
    x1, _, _ = _lsmr(A, b, **kw)
    x2, _, _ = _lsmr(A[p], b[p], **kw)
    assert np.linalg.norm(x1 - x2) <= 1e-10 * np.linalg.norm(x1)
    # a regulariser row with a different weight changes A^T A: iterates move
    data, reg = _split(A, b)
    D = np.ones(A.shape[0])
    D[reg[len(reg) // 2]] = 2.0
    x3, _, _ = _lsmr(sp.diags(D) @ A, b, **kw)
    assert np.linalg.norm(x1 - x3) > 1e-6 * np.linalg.norm(x1)
