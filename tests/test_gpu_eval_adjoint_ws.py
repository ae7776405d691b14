"""The adjoint of evaluation with caller storage (lsamd_fd_evaluate_adjoint_ws,
used by the autograd adjoint) gives the allocating entry point's output bit
for bit, rejects a workspace smaller than its size query, and both equal the
oracle's E^T g (E the explicit trilinear evaluation matrix) to 1e-14."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import fd as ofd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nn,n", [((9, 12, 7), 5000), ((40, 33, 21), 200000), ((5, 5, 5), 0)])
def test_workspace_adjoint_matches_allocating(nn, n):
    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    rng = np.random.default_rng(n + sum(nn))
    origin, step = np.array([0.1, -0.2, 0.3]), np.array([0.05, 0.07, 0.11])
    grid = _lib.make_grid(tuple(origin), tuple(step), nn)
    ext = step * (np.array(nn) - 1)
    p = origin + rng.uniform(-0.05, 1.05, (n, 3)) * ext
    pts = torch.tensor(p, device="cuda")
    gv = torch.tensor(rng.standard_normal(n), device="cuda")
    m = int(np.prod(nn))
    ref = torch.full((m,), np.nan, dtype=torch.float64, device="cuda")
    lib = _lib.load()
    assert lib.lsamd_fd_evaluate_adjoint(C.byref(grid), _lib.ptr(pts), n, _lib.ptr(gv),
                                         _lib.ptr(ref), _lib.stream()) == 0
    out = fs.evaluate_adjoint(grid, pts, gv, m)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    if n:
        nb = int(lib.lsamd_fd_evaluate_adjoint_ws_bytes(C.byref(grid), n))
        small = torch.empty(nb - 256, dtype=torch.uint8, device="cuda")
        rc = lib.lsamd_fd_evaluate_adjoint_ws(C.byref(grid), _lib.ptr(pts), n, _lib.ptr(gv),
                                              _lib.ptr(out), _lib.ptr(small), nb - 256,
                                              _lib.stream())
        assert rc != 0 and b"workspace smaller" in lib.lsamd_last_error()
    if 0 < n <= 5000:
        og = ofd.Grid(tuple(origin), tuple(v - 1 for v in nn), tuple(step))
        idc, ins = ofd.cell_corners(og, p)
        w = ofd.trilinear(ofd.local_coords(og, p[ins]))
        want = np.zeros(m)
        np.add.at(want, idc[ins].ravel(), (w * gv.cpu().numpy()[ins][:, None]).ravel())
        got = out.cpu().numpy()
        assert np.max(np.abs(got - want)) <= 1e-14 * max(1.0, np.max(np.abs(want)))
