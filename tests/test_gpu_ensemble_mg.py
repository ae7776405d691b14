"""EnsembleSolver: multigrid solves of an uneven ensemble, shape groups on
four host threads with a stream each, handles reused across the systems of
one shape.  Every system reaches the oracle's direct minimiser (spsolve on
the normal equations) to 1e-8, and a second call on the same solver (handles,
hierarchies and graphs reused) returns the same bits."""
import numpy as np
import pytest
import torch

from oracle import fd as ofd

pytestmark = pytest.mark.gpu

W = {"dxx": 0.1, "dyy": 0.1, "dzz": 0.1, "dxy": 0.1, "dxz": 0.1, "dyz": 0.1,
     "dx": 1.0, "dy": 1.0, "dz": 1.0}


def _rows(g, values, normals, n):
    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    blocks = []
    for pts, kind, wcol, ws in ((normals, _lib.ROWS_NORMAL, 6, 10.0 / 3),
                                (values, _lib.ROWS_VALUE, 4, 1.0)):
        p = torch.as_tensor(pts).cuda()
        _, ins, _, _ = fs.locate(g, p[:, :3].contiguous(), want_w=False)
        blocks.append(fs.build_rows(g, kind, p[ins].contiguous(), wcol, ws, n))
    return fs.concat_rows(blocks, n)


def test_ensemble_multigrid_matches_minimiser():
    from loopstructural_amd import _lib, synthetic
    from loopstructural_amd import fd_system as fs

    ens = synthetic.ensemble(14, 2024, nodes=(14, 18), n_values=(100, 400),
                             n_normals=(20, 80))
    grids, blocks = [], []
    for nn, v, m in ens:
        g = _lib.make_grid((0, 0, 0), (1.0 / (nn - 1),) * 3, (nn,) * 3)
        grids.append(g)
        blocks.append(_rows(g, v, m, nn ** 3))
    es = fs.EnsembleSolver(grids, fs.reg_struct(W, (16, 16, 16), 1e-8), workers=4)
    assert len(es.shapes) < len(ens)  # some shapes hold several systems
    print(f"ensemble: {len(ens)} systems, {len(es.shapes)} shapes", flush=True)
    c, infos = es.solve(blocks, rtol=1e-12)
    assert all(i["converged"] for i in infos)
    for s, (nn, v, m) in enumerate(ens):
        og = ofd.Grid((0, 0, 0), (nn - 1,) * 3, (1.0 / (nn - 1),) * 3)
        ref = ofd.solve_exact(ofd.assemble_fd(og, {"value": v.copy(), "normal": m.copy()},
                                              regularisation=0.1))
        err = float(np.linalg.norm(c[s].cpu().numpy() - ref) / np.linalg.norm(ref))
        assert err <= 1e-8, (s, nn, err)
    c2, _ = es.solve(blocks, rtol=1e-12)
    for a, b in zip(c, c2):
        assert torch.equal(a, b)
    es.close()
