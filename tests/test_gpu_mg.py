"""Multigrid-preconditioned flexible CG on FD systems.

The preconditioner changes the iterates, not the minimiser: solutions are
checked against the exact least-squares minimiser (spsolve on the oracle's
normal equations) at 1e-8 relative L2 for rtol 1e-12, the final residual
against the same FP64 stopping test as the Jacobi form, and the iteration
count against the Jacobi form's.
"""
import numpy as np
import pytest
import torch

from oracle import fd as ofd

pytestmark = pytest.mark.gpu

W = {"dxx": 0.1, "dyy": 0.1, "dzz": 0.1, "dxy": 0.1, "dxz": 0.1, "dyz": 0.1,
     "dx": 1.0, "dy": 1.0, "dz": 1.0}


def _rows(g, n, seed, nv, nn):
    from loopstructural_amd import _lib, synthetic
    from loopstructural_amd import fd_system as fs

    values, normals = synthetic.constraints(nv, nn, seed, value_noise=0.02)
    blocks = []
    for pts, kind, wcol, ws in ((normals, _lib.ROWS_NORMAL, 6, 10.0 / 3),
                                (values, _lib.ROWS_VALUE, 4, 1.0)):
        p = torch.from_numpy(pts).cuda()
        _, ins, _, _ = fs.locate(g, p[:, :3].contiguous(), want_w=False)
        blocks.append(fs.build_rows(g, kind, p[ins].contiguous(), wcol, ws, n))
    return fs.concat_rows(blocks, n), values, normals


def _system(cells, seed, nv=300, nn=60):
    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    cells = np.array(cells)
    step = 1.0 / cells
    nn_ = cells + 1
    g = _lib.make_grid((0, 0, 0), step, nn_)
    n = int(np.prod(nn_))
    rows, values, normals = _rows(g, n, seed, nv, nn)
    s = fs.FdSystem(g, fs.reg_struct(W, nn_, 1e-8))
    s.set_rows(rows)
    og = ofd.Grid((0, 0, 0), cells, step)
    sysm = ofd.assemble_fd(og, {"value": values.copy(), "normal": normals.copy()},
                           regularisation=0.1)
    return s, sysm


@pytest.mark.parametrize("cells", [(3, 4, 2), (12, 12, 12), (17, 9, 24), (31, 31, 31)])
def test_multigrid_solve_matches_minimiser(cells):
    s, sysm = _system(cells, 5)
    ref = ofd.solve_exact(sysm)
    c, info = s.solve(rtol=1e-12, multigrid=True, check_every=8)
    assert info["converged"] and info["residual"] < info["threshold"]
    err = np.linalg.norm(c.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err <= 1e-8, (cells, err, info)
    cj, ij = s.solve(rtol=1e-12, check_every=64)
    if min(cells) >= 12:
        assert info["iterations"] * 4 < ij["iterations"], (info, ij)


def test_multigrid_rebuilds_after_new_rows_and_rhs():
    """The hierarchy outlives set_rows: its row-dependent parts are rebuilt
    (twice here, with rows of another size in between) and the solution is
    the minimiser of the new rows, equal to a fresh system's."""
    s, sysm = _system((20, 20, 20), 6)
    c1, _ = s.solve(rtol=1e-12, multigrid=True)
    rhs = torch.from_numpy(np.random.default_rng(3).standard_normal(s.n)).cuda()
    c2, info = s.solve(rtol=1e-12, multigrid=True, rhs=rhs)
    N, _ = ofd.normal_equations(sysm)
    res = N @ c2.cpu().numpy() - rhs.cpu().numpy()
    assert np.linalg.norm(res) <= 1e-8 * np.linalg.norm(rhs.cpu().numpy())
    s2, sysm2 = _system((20, 20, 20), 7)
    c3, _ = s2.solve(rtol=1e-12, multigrid=True)
    ref = ofd.solve_exact(sysm2)
    assert np.linalg.norm(c3.cpu().numpy() - ref) <= 1e-8 * np.linalg.norm(ref)
    # new rows on the first system: seed 8 (other sizes), then seed 7 again
    g, n = s.grid, s.n
    s.set_rows(_rows(g, n, 8, 500, 20)[0])
    s.solve(rtol=1e-12, multigrid=True)
    s.set_rows(_rows(g, n, 7, 300, 60)[0])
    c4, _ = s.solve(rtol=1e-12, multigrid=True)
    assert np.linalg.norm(c4.cpu().numpy() - ref) <= 1e-8 * np.linalg.norm(ref)


def test_interpolator_multigrid_option():
    from loopstructural_amd import FiniteDifferenceInterpolator, StructuredGrid, synthetic

    values, normals = synthetic.constraints(200, 50, 9)
    grid = StructuredGrid(origin=(0, 0, 0), nsteps=(24, 24, 24), step_vector=(1 / 24,) * 3)
    it = FiniteDifferenceInterpolator(grid)
    it.set_value_constraints(values)
    it.set_normal_constraints(normals)
    it.setup_interpolator(regularisation=0.1)
    it.solve_system("cg", solver_kwargs={"rtol": 1e-12, "preconditioner": "mg"})
    og = ofd.Grid((0, 0, 0), (24, 24, 24), (1 / 24,) * 3)
    ref = ofd.solve_exact(ofd.assemble_fd(og, {"value": values, "normal": normals},
                                          regularisation=0.1))
    assert np.linalg.norm(it.c - ref) <= 1e-8 * np.linalg.norm(ref)
    assert it.get_last_solve_timing()["iterations"] < 400


def test_schwarz_smoother_against_point_jacobi(monkeypatch):
    """Both smoothers converge to the same solution (the minimiser itself is
    checked at this size above); the Schwarz blocks (data cells and boundary
    layers) cut the iteration count of point Jacobi by at least 40% (a CPU
    prototype of the same hierarchy: 212 -> 73 iterations at 31^3)."""
    its, sol = {}, {}
    for smoother in ("jacobi", "schwarz"):
        monkeypatch.setenv("LSAMD_MG_SMOOTHER", smoother)
        s, _ = _system((31, 31, 31), 11, nv=1000, nn=200)
        c, info = s.solve(rtol=1e-12, multigrid=True, check_every=8)
        assert info["converged"], (smoother, info)
        its[smoother] = info["iterations"]
        sol[smoother] = c.cpu().numpy()
    diff = np.linalg.norm(sol["schwarz"] - sol["jacobi"]) / np.linalg.norm(sol["jacobi"])
    assert diff <= 1e-8, diff
    assert its["schwarz"] <= 0.6 * its["jacobi"], its
