"""Uneven ensembles by multigrid.  EnsembleSolver: shape groups on host
threads with a stream each, handles reused across the systems of one shape.
BatchedMultigrid: one hierarchy per system, all systems advanced by one
launch sequence per iteration.  Every system reaches the oracle's direct
minimiser (spsolve on the normal equations) to 1e-8, and a second call with
the same rows returns the same bits."""
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


@pytest.mark.parametrize("nodes,count", [((14, 18), 12), ((9, 24), 12), ((22, 24), 6)])
def test_batched_multigrid_matches_minimiser(nodes, count):
    """Every system of the batched multigrid solve (one launch sequence per
    iteration for all systems; hierarchies of two to four levels) reaches the
    oracle's direct minimiser to 1e-8 (sizes kept to <= 24^3 nodes: the
    direct solve grows fast with the grid); a second solve of the same rows
    returns the same bits.  Ensembles of 22-24 nodes per axis take the
    plane-streaming fine-level operator (every axis >= 16 nodes, z-chunks of
    16), the others the node-by-node one."""
    from loopstructural_amd import _lib, synthetic
    from loopstructural_amd import fd_system as fs

    ens = synthetic.ensemble(count, 77 + nodes[1], nodes=nodes, n_values=(100, 400),
                             n_normals=(20, 80))
    grids, blocks = [], []
    for nn, v, m in ens:
        g = _lib.make_grid((0, 0, 0), (1.0 / (nn - 1),) * 3, (nn,) * 3)
        grids.append(g)
        blocks.append(_rows(g, v, m, nn ** 3))
    print(f"batched multigrid: {len(ens)} systems, nodes {[e[0] for e in ens]}", flush=True)
    bm = fs.BatchedMultigrid(grids, fs.reg_struct(W, (16, 16, 16), 1e-8), workers=4)
    bm.prepare(blocks)
    c, infos = bm.solve(blocks, rtol=1e-12, maxiter=3000)
    print(f"batched multigrid: iterations {[i['iterations'] for i in infos]}", flush=True)
    assert all(i["converged"] for i in infos)
    for s, (nn, v, m) in enumerate(ens):
        print(f"batched multigrid: oracle solve of system {s} ({nn}^3)", flush=True)
        og = ofd.Grid((0, 0, 0), (nn - 1,) * 3, (1.0 / (nn - 1),) * 3)
        ref = ofd.solve_exact(ofd.assemble_fd(og, {"value": v.copy(), "normal": m.copy()},
                                              regularisation=0.1))
        err = float(np.linalg.norm(c[s].cpu().numpy() - ref) / np.linalg.norm(ref))
        assert err <= 1e-8, (s, nn, err)
    c2, infos2 = bm.solve(blocks, rtol=1e-12, maxiter=3000)
    for a, b in zip(c, c2):
        assert torch.equal(a, b)
    assert [i["iterations"] for i in infos] == [i["iterations"] for i in infos2]
    bm.close()
