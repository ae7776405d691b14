"""Multigrid solves of several right-hand sides per launch sequence
(realisations that share the constraint locations: one matrix, perturbed
observed values and normal directions).

Each lane is checked against the single-system multigrid solve of the same
right-hand side (both to rtol 1e-12: agreement 1e-9 relative) and against the
exact minimiser of its realisation (spsolve on the oracle's normal equations,
1e-8).  Groups are 4, 8, 16, 32 or 64 lanes wide (32 once at least 24 right-hand
sides are left, 64 once at least 48, else the smallest width that holds
them): 3 -> 4, 8 -> 8, 11 -> 16, 21 -> 16 + 8, 27 -> 32 (5 idle lanes),
37 -> 32 + 8, 50 -> 64 (14 idle), which covers every width, padded groups,
and two widths in one call (a work per width is
kept on the handle); a 128^3 case takes the fine-level and transfer kernels of
large grids.  The right-hand side of a realisation from its row values
is checked against the system's own rhs after set_rows with those rows.  A
failure while the lane work is allocated leaves no half-built work behind: the
next call builds it afresh and solves correctly.
"""
import numpy as np
import pytest
import torch

from oracle import fd as ofd

pytestmark = pytest.mark.gpu

W = {"dxx": 0.1, "dyy": 0.1, "dzz": 0.1, "dxy": 0.1, "dxz": 0.1, "dyz": 0.1,
     "dx": 1.0, "dy": 1.0, "dz": 1.0}


def _rows(fs, _lib, g, n, values, normals):
    blocks = []
    for pts, kind, wcol, ws in ((normals, _lib.ROWS_NORMAL, 6, 10.0 / 3),
                                (values, _lib.ROWS_VALUE, 4, 1.0)):
        p = torch.from_numpy(pts).cuda()
        _, ins, _, _ = fs.locate(g, p[:, :3].contiguous(), want_w=False)
        blocks.append(fs.build_rows(g, kind, p[ins].contiguous(), wcol, ws, n))
    return fs.concat_rows(blocks, n)


def _setup(cells, nreal, seed=31):
    from loopstructural_amd import _lib, synthetic
    from loopstructural_amd import fd_system as fs

    nn = cells + 1
    g = _lib.make_grid((0, 0, 0), (1.0 / cells,) * 3, (nn,) * 3)
    n = nn ** 3
    values, normals = synthetic.constraints(400, 80, seed)
    reals = [synthetic.perturb(values, normals, 100 + r) for r in range(nreal)]
    s = fs.FdSystem(g, fs.reg_struct(W, (nn,) * 3, 1e-8))
    s.set_rows(_rows(fs, _lib, g, n, *reals[0]))
    rhs = torch.stack([s.rhs_for(_rows(fs, _lib, g, n, v, m).b) for v, m in reals])
    return s, rhs, reals, g, n


def test_rhs_for_matches_set_rows():
    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    s, rhs, reals, g, n = _setup(15, 2)
    s2 = fs.FdSystem(g, fs.reg_struct(W, (16,) * 3, 1e-8))
    s2.set_rows(_rows(fs, _lib, g, n, *reals[1]))
    ref = s2.rhs()
    assert float(torch.linalg.vector_norm(rhs[1] - ref) / torch.linalg.vector_norm(ref)) <= 1e-14


@pytest.mark.parametrize("nreal", [3, 8, 11, 21, 27, 37, 50])
def test_multi_rhs_lanes_match_single_solves(nreal):
    s, rhs, reals, g, n = _setup(23, nreal)
    c, infos = s.solve_multi(rhs, rtol=1e-12)
    assert all(i["converged"] for i in infos), infos
    for r in range(nreal):
        c1, i1 = s.solve(rtol=1e-12, rhs=rhs[r], multigrid=True, check_every=8)
        d = float(torch.linalg.vector_norm(c[r] - c1) / torch.linalg.vector_norm(c1))
        assert d <= 1e-9, (r, d)
        # the multi-rhs V-cycle smooths the coarse levels with degree 6 (4 in
        # the single-system one): never more than one chunk of 8 beyond it
        assert infos[r]["iterations"] <= i1["iterations"] + 8, (infos[r], i1)
    og = ofd.Grid((0, 0, 0), (23,) * 3, (1 / 23,) * 3)
    for r in (0, nreal - 1):
        v, m = reals[r]
        ref = ofd.solve_exact(ofd.assemble_fd(og, {"value": v.copy(), "normal": m.copy()},
                                              regularisation=0.1))
        assert np.linalg.norm(c[r].cpu().numpy() - ref) <= 1e-8 * np.linalg.norm(ref)


def test_multi_rhs_at_128_cubed():
    """At 128^3 nodes the fine level has a covered list (from 2^17 nodes): its
    covered nodes are smoothed from the interleaved residual over the list and
    every node by the whole-line pass, and level 1 (65^3) is reached by the
    z-streamed restriction and prolongation.  Four realisations against the
    single-system multigrid solve of each (1e-9); the multi-rhs solve takes
    at most one captured chunk of 8 iterations more (its outer iteration
    keeps direction and residual in FP32, the residual replaced every 8
    iterations; its coarse levels smooth with degree 6 against 4)."""
    s, rhs, reals, g, n = _setup(127, 4)
    print("multi-rhs at 128^3: solving 4 lanes", flush=True)
    c, infos = s.solve_multi(rhs, rtol=1e-12)
    assert all(i["converged"] for i in infos), infos
    for r in range(4):
        c1, i1 = s.solve(rtol=1e-12, rhs=rhs[r], multigrid=True, check_every=8)
        d = float(torch.linalg.vector_norm(c[r] - c1) / torch.linalg.vector_norm(c1))
        print(f"lane {r}: {infos[r]['iterations']} vs {i1['iterations']} iterations, "
              f"rel. difference {d:.2e}", flush=True)
        assert d <= 1e-9, (r, d)
        assert infos[r]["iterations"] <= i1["iterations"] + 8, (infos[r], i1)


def test_multi_rhs_after_new_rows_and_zero_rhs():
    s, rhs, reals, g, n = _setup(19, 4)
    rhs[2] = 0.0
    c, infos = s.solve_multi(rhs, rtol=1e-12)
    assert infos[2]["iterations"] == 0 and float(c[2].abs().max()) == 0.0
    # new rows (other realisation set) on the same handle: graph rebuilt
    from loopstructural_amd import _lib, synthetic
    from loopstructural_amd import fd_system as fs

    values, normals = synthetic.constraints(300, 50, 77)
    s.set_rows(_rows(fs, _lib, g, n, values, normals))
    c2, infos2 = s.solve_multi(s.rhs()[None, :].repeat(2, 1), rtol=1e-12)
    c1, _ = s.solve(rtol=1e-12, multigrid=True)
    for r in range(2):
        assert float(torch.linalg.vector_norm(c2[r] - c1) / torch.linalg.vector_norm(c1)) <= 1e-9


def test_multi_rhs_work_rebuilt_after_a_failed_allocation(monkeypatch):
    from loopstructural_amd import _lib

    s, rhs, reals, g, n = _setup(15, 5)
    monkeypatch.setenv("LSAMD_FAULT_INJECT", "multi_rhs_work")
    with pytest.raises(_lib.LsamdError) as e:
        s.solve_multi(rhs, rtol=1e-12)
    assert e.value.code == _lib.LSAMD_ERR_INJECTED
    monkeypatch.delenv("LSAMD_FAULT_INJECT")
    c, infos = s.solve_multi(rhs, rtol=1e-12)
    assert all(i["converged"] for i in infos), infos
    for r in range(5):
        c1, _ = s.solve(rtol=1e-12, rhs=rhs[r], multigrid=True, check_every=8)
        assert float(torch.linalg.vector_norm(c[r] - c1) / torch.linalg.vector_norm(c1)) <= 1e-9
