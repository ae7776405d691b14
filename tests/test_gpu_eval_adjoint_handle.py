"""The adjoint of evaluation on a system handle (lsamd_fd_system_evaluate_adjoint,
used by the autograd adjoint, sort buffers cached on the handle) gives the
allocating entry point's output bit for bit on calls of growing and shrinking
size, and both equal the oracle's E^T g (E the explicit trilinear evaluation
matrix) to 1e-14."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import fd as ofd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nn", [(9, 12, 7), (40, 33, 21)])
def test_handle_adjoint_matches_allocating(nn):
    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    origin, step = np.array([0.1, -0.2, 0.3]), np.array([0.05, 0.07, 0.11])
    grid = _lib.make_grid(tuple(origin), tuple(step), nn)
    w = {k: 0.1 for k in ("dxx", "dyy", "dzz", "dxy", "dxz", "dyz")}
    system = fs.FdSystem(grid, fs.reg_struct(w, nn, 1e-8))
    ext = step * (np.array(nn) - 1)
    m = int(np.prod(nn))
    lib = _lib.load()
    for n in (5000, 200000, 0, 777):
        print(f"handle adjoint: grid {nn}, {n} points", flush=True)
        rng = np.random.default_rng(n + sum(nn))
        p = origin + rng.uniform(-0.05, 1.05, (n, 3)) * ext
        pts = torch.tensor(p, device="cuda").reshape(n, 3)
        gv = torch.tensor(rng.standard_normal(n), device="cuda")
        ref = torch.full((m,), np.nan, dtype=torch.float64, device="cuda")
        assert lib.lsamd_fd_evaluate_adjoint(C.byref(grid), _lib.ptr(pts), n, _lib.ptr(gv),
                                             _lib.ptr(ref), _lib.stream()) == 0
        out = system.evaluate_adjoint(pts, gv)
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
        if 0 < n <= 5000:
            og = ofd.Grid(tuple(origin), tuple(v - 1 for v in nn), tuple(step))
            idc, ins = ofd.cell_corners(og, p)
            wts = ofd.trilinear(ofd.local_coords(og, p[ins]))
            want = np.zeros(m)
            np.add.at(want, idc[ins].ravel(), (wts * gv.cpu().numpy()[ins][:, None]).ravel())
            got = out.cpu().numpy()
            assert np.max(np.abs(got - want)) <= 1e-14 * max(1.0, np.max(np.abs(want)))
    system.close()
