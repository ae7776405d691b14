"""GPU LSMR on P1 supports (solve_system('lsmr'), LI/_solver_strategy.py:108-136)
against scipy's lsmr on the oracle's explicit P1 system (oracle/p1.py), ridge
rows and column scaling from oracle/fd.py's preprocess.

The GPU runs LSMR on the data rows plus one merged block of edge-jump rows of
weight sqrt(sum w^2); that leaves A^T A, A^T b and |b| unchanged, so the
iterates agree with scipy's on the reference's stacked matrix in exact
arithmetic.  Bars as for the FD supports: a fixed run of 10 iterations
(istop 7) to 1e-12 relative; tight tolerances to the unique minimiser
(spsolve) to 1e-8.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import fd as ofd
from oracle import p1 as op1

pytestmark = pytest.mark.gpu

FIXED = {"atol": 0.0, "btol": 0.0, "conlim": 0.0, "maxiter": 10}
MESH = ((0.0, 0.0, 0.0), (10, 9, 11), (0.1, 0.111, 0.0909))


def _data(m, seed, nv=200, nn=40, nt=30):
    rng = np.random.default_rng(seed)
    ext = m.ncells * m.step
    pv = m.origin + rng.uniform(-0.02, 1.02, (nv, 3)) * ext
    vals = np.column_stack([pv, np.sin(pv[:, 0]) + pv[:, 2], rng.uniform(0.5, 2, nv)])
    pn = m.origin + rng.uniform(0.0, 1.0, (nn, 3)) * ext
    nrm = rng.standard_normal((nn, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    normals = np.column_stack([pn, nrm, np.ones(nn)])
    pt = m.origin + rng.uniform(0.0, 1.0, (nt, 3)) * ext
    tang = np.column_stack([pt, rng.standard_normal((nt, 3)), np.ones(nt)])
    return {"value": vals, "normal": normals, "tangent": tang}


def _interp(m, data, scaling=True, **kw):
    from loopstructural_amd import P1Interpolator, TetMesh

    mesh = TetMesh(origin=m.origin, nsteps_cells=m.ncells, step_vector=m.step)
    it = P1Interpolator(mesh)
    it.set_value_constraints(data["value"])
    it.set_normal_constraints(data["normal"])
    it.set_tangent_constraints(data["tangent"])
    it.setup_interpolator(**kw)
    it.apply_scaling_matrix = scaling
    return it


def _scipy(s, scaling=True, **kw):
    A, b = s.build_matrix()
    A, b, S = ofd.preprocess(A, b, ridge=1e-8, scaling=scaling)
    out = spla.lsmr(A, b, **kw)
    y = out[0]
    return (S @ y if S is not None else y), int(out[1]), int(out[2])


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("scaling", [True, False])
def test_p1_lsmr_fixed_iterations_match_scipy(scaling):
    m = op1.TetMesh(*MESH)
    data = _data(m, 31)
    it = _interp(m, data, scaling=scaling, regularisation=0.1)
    assert it.solve_system("lsmr", solver_kwargs=dict(FIXED))
    t = it.get_last_solve_timing()
    assert (t["istop"], t["iterations"]) == (7, 10)
    c, istop, itn = _scipy(op1.assemble_p1(m, data, regularisation=0.1), scaling=scaling,
                           **FIXED)
    assert (istop, itn) == (7, 10)
    assert _rel(it.c, c) <= 1e-12


def test_p1_lsmr_two_edge_families_merged():
    """A second edge-jump family: the reference stacks both row sets."""
    m = op1.TetMesh(*MESH)
    data = _data(m, 32)
    it = _interp(m, data, regularisation=0.1)
    it.minimise_edge_jumps(0.25, name="extra edge jump")
    assert it.solve_system("lsmr", solver_kwargs=dict(FIXED))
    s = op1.assemble_p1(m, data, regularisation=0.1)
    a, idc = op1.edge_jump_rows(m)
    s.add(a, np.zeros(a.shape[0]), idc, 0.25, "extra edge jump")
    c, istop, itn = _scipy(s, **FIXED)
    assert (istop, itn) == (7, 10)
    assert _rel(it.c, c) <= 1e-12


def test_p1_lsmr_weight_scaled_regulariser():
    m = op1.TetMesh((0, 0, 0), (8, 8, 8), (0.125, 0.125, 0.125))
    data = _data(m, 33, nt=0)
    data["tangent"] = np.zeros((0, 7))
    it = _interp(m, data, use_regularisation_weight_scale=True)
    assert it.solve_system("lsmr", solver_kwargs=dict(FIXED))
    # the scale: 1 + 1 + 10 exp(-d^2 / 2 sigma^2), sigma = nnodes[0] * 10
    nodes = m.nodes()
    pn = data["normal"][:, :3]
    _, _, ins = m.element_for_location(pn)[1:]
    d2 = ((nodes[:, None, :] - pn[None, ins, :]) ** 2).sum(-1).min(1)
    sigma = m.nnodes[0] * 10
    scale = 1.0 + (1 + 10 * np.exp(-d2 / (2 * sigma ** 2)))
    c, istop, itn = _scipy(op1.assemble_p1(m, data, scale=scale), **FIXED)
    assert (istop, itn) == (7, 10)
    assert _rel(it.c, c) <= 1e-12


def test_p1_lsmr_tight_reaches_minimiser():
    m = op1.TetMesh(*MESH)
    data = _data(m, 34)
    it = _interp(m, data, regularisation=0.1)
    assert it.solve_system("lsmr", solver_kwargs={"atol": 1e-13, "btol": 1e-13,
                                                  "maxiter": 200000})
    t = it.get_last_solve_timing()
    assert t["istop"] in (1, 2), t
    ref = ofd.solve_exact(op1.assemble_p1(m, data, regularisation=0.1))
    assert _rel(it.c, ref) <= 1e-8, t
