"""Batched ensemble solver against single-system solves and the CPU oracle.

Tolerances: operator, rhs and diagonal 1e-13 relative to the largest entry
(FP64); FP64 solutions 1e-8 relative L2 against the exact minimiser at rtol
1e-12; mixed FP32/FP64 solutions 1e-6 at rtol 1e-10.
"""
import numpy as np
import pytest
import torch

from oracle import fd as ofd

pytestmark = pytest.mark.gpu

W = {"dxx": 0.1, "dyy": 0.1, "dzz": 0.1, "dxy": 0.1, "dxz": 0.1, "dyz": 0.1,
     "dx": 1.0, "dy": 1.0, "dz": 1.0}
SHAPES = [(7, 7, 7), (9, 8, 12), (16, 11, 7), (13, 13, 13), (8, 20, 10)]


def _case(shape, seed):
    from loopstructural_amd import _lib, synthetic

    cells = np.array(shape) - 1
    step = 1.0 / cells
    g = _lib.make_grid((0, 0, 0), step, shape)
    og = ofd.Grid((0, 0, 0), cells, step)
    rng = np.random.default_rng(seed)
    nv, nn = int(rng.integers(40, 200)), int(rng.integers(10, 40))
    values, normals = synthetic.constraints(nv, nn, seed, value_noise=0.02)
    return g, og, values, normals


def _rows(g, values, normals, n):
    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    blocks = []
    for pts, kind, wcol, ws in ((normals, _lib.ROWS_NORMAL, 6, 10.0 / 3),
                                (values, _lib.ROWS_VALUE, 4, 1.0)):
        p = torch.from_numpy(pts).cuda()
        _, ins, _, _ = fs.locate(g, p[:, :3].contiguous(), want_w=False)
        blocks.append(fs.build_rows(g, kind, p[ins].contiguous(), wcol, ws, n))
    return fs.concat_rows(blocks, n)


def _batch(shapes=SHAPES):
    from loopstructural_amd import fd_system as fs

    cases = [_case(s, 100 + i) for i, s in enumerate(shapes)]
    reg = fs.reg_struct(W, (7, 7, 7), 1e-8)
    bs = fs.BatchSystem([c[0] for c in cases], reg)
    rows = [_rows(c[0], c[2], c[3], int(np.prod(shapes[i]))) for i, c in enumerate(cases)]
    bs.set_rows(rows)
    singles = []
    for i, c in enumerate(cases):
        s = fs.FdSystem(c[0], fs.reg_struct(W, shapes[i], 1e-8))
        s.set_rows(rows[i])
        singles.append(s)
    return bs, singles, cases


def test_batch_operator_equals_single_systems():
    bs, singles, _ = _batch()
    x = torch.from_numpy(np.random.default_rng(0).standard_normal(bs.n)).cuda()
    y = bs.apply(x)
    for s, (ys, xs) in enumerate(zip(bs.split(y), bs.split(x))):
        ref = singles[s].apply(xs.contiguous())
        assert float((ys - ref).abs().max()) <= 1e-13 * float(ref.abs().max())
    for got, name in ((bs.rhs(), "rhs"), (bs.diag(), "diag")):
        for s, part in enumerate(bs.split(got)):
            ref = getattr(singles[s], name)()
            assert float((part - ref).abs().max()) <= 1e-13 * float(ref.abs().max()), name
    y32 = bs.apply(x.float()).double()
    assert float((y32 - y).abs().max()) <= 1e-5 * float(y.abs().max())


@pytest.mark.parametrize("tile,zc", [(64, 16), (32, 16), (64, 32), (32, 32)])
def test_batch_stencil_forms(monkeypatch, tile, zc):
    """Each stencil form of a batch (64 x 8 or 32 x 16 tiles, z-chunk 16 or
    32), pinned before the solve that fixes it, against the single systems'
    operators (1e-13) and its FP32 apply against FP64 (1e-5); the shapes
    include z-extents of several chunks and partial tiles."""
    monkeypatch.setenv("LSAMD_BATCH_TILE", str(tile))
    monkeypatch.setenv("LSAMD_BATCH_ZC", str(zc))
    shapes = SHAPES + [(21, 18, 45), (36, 20, 71), (70, 9, 40)]
    bs, singles, _ = _batch(shapes)
    bs.solve(rtol=1e-5, maxiter=1, precision="fp32", check_every=1)  # fixes the form
    x = torch.from_numpy(np.random.default_rng(1).standard_normal(bs.n)).cuda()
    y = bs.apply(x)
    for s, (ys, xs) in enumerate(zip(bs.split(y), bs.split(x))):
        ref = singles[s].apply(xs.contiguous())
        assert float((ys - ref).abs().max()) <= 1e-13 * float(ref.abs().max()), shapes[s]
    y32 = bs.apply(x.float()).double()
    assert float((y32 - y).abs().max()) <= 1e-5 * float(y.abs().max())


def test_batch_fp64_solve_matches_minimiser_and_single_iterations():
    bs, singles, cases = _batch()
    c, infos = bs.solve(rtol=1e-12, precision="fp64", check_every=32)
    for s, part in enumerate(bs.split(c)):
        g, og, values, normals = cases[s]
        sysm = ofd.assemble_fd(og, {"value": values.copy(), "normal": normals.copy()},
                               regularisation=0.1)
        ref = ofd.solve_exact(sysm)
        err = np.linalg.norm(part.cpu().numpy() - ref) / np.linalg.norm(ref)
        assert err <= 1e-8, (s, err)
        assert infos[s]["converged"]
        # same algorithm; the dot products are reduced over a different block
        # partition, so the iterates differ in rounding only
        _, si = singles[s].solve(rtol=1e-12, precision="fp64", check_every=32)
        assert abs(si["iterations"] - infos[s]["iterations"]) <= max(2, 0.01 * si["iterations"]), \
            (s, si, infos[s])


def test_batch_mixed_precision_solve():
    bs, _, cases = _batch()
    c, infos = bs.solve(rtol=1e-10, precision="fp32", inner_rtol=1e-3, check_every=32)
    for s, part in enumerate(bs.split(c)):
        g, og, values, normals = cases[s]
        sysm = ofd.assemble_fd(og, {"value": values.copy(), "normal": normals.copy()},
                               regularisation=0.1)
        ref = ofd.solve_exact(sysm)
        err = np.linalg.norm(part.cpu().numpy() - ref) / np.linalg.norm(ref)
        assert err <= 1e-6, (s, err)
        assert infos[s]["converged"] and infos[s]["outer"] >= 1


def test_batch_rejects_rows_of_another_system_and_small_grids():
    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    bs, _, cases = _batch()
    n0 = int(np.prod(SHAPES[0]))
    rows = [_rows(c[0], c[2], c[3], int(np.prod(SHAPES[i]))) for i, c in enumerate(cases)]
    bad = fs.FdRows(rows[0].row_ptr, rows[0].cols + n0, rows[0].coef, rows[0].b, rows[0].w,
                    rows[0].norm, n0)  # system 0 rows pointing into system 1
    with pytest.raises(_lib.LsamdError):
        bs.set_rows([bad] + rows[1:])
    g_small = _lib.make_grid((0, 0, 0), (0.2,) * 3, (6, 8, 8))
    with pytest.raises(_lib.LsamdError):
        fs.BatchSystem([g_small], fs.reg_struct(W, (6, 8, 8), 1e-8))
