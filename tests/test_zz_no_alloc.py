"""The allocation-free adjoint evaluation (lsamd_fd_evaluate_adjoint_noalloc)
through the C ABI against the allocating entry point it mirrors: bit-identical
outputs on two calls in a row, the second with fewer points on another grid
(so the scratch holds the first call's larger contents), and a clear error
above the scratch capacity."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _call(name, grid, pts, gv, out):
    from loopstructural_amd import _lib

    return getattr(_lib.load(), name)(C.byref(grid), _lib.ptr(pts), pts.shape[0], _lib.ptr(gv),
                                      _lib.ptr(out), _lib.stream())


@pytest.mark.parametrize("cases", [((21, 17, 13), 40000, (9, 12, 7), 777),
                                   ((5, 5, 5), 65536, (33, 31, 29), 1)])
def test_noalloc_matches_allocating_entry_point(cases):
    from loopstructural_amd import _lib

    g1, n1, g2, n2 = cases
    rng = np.random.default_rng(sum(g1) + n1)
    for nn, n in ((g1, n1), (g2, n2)):
        print(f"noalloc case: grid {nn}, {n} points", flush=True)
        grid = _lib.make_grid((0.1, -0.2, 0.3), (0.05, 0.07, 0.11), nn)
        ext = np.array([0.05 * (nn[0] - 1), 0.07 * (nn[1] - 1), 0.11 * (nn[2] - 1)])
        p = np.array([0.1, -0.2, 0.3]) + rng.uniform(-0.05, 1.05, (n, 3)) * ext
        p[1::97, 1] = np.nan
        pts = torch.tensor(p, device="cuda")
        gv = torch.tensor(rng.standard_normal(n), device="cuda")
        m = int(np.prod(nn))
        ref = torch.full((m,), np.nan, dtype=torch.float64, device="cuda")
        out = torch.full((m,), np.nan, dtype=torch.float64, device="cuda")
        assert _call("lsamd_fd_evaluate_adjoint", grid, pts, gv, ref) == 0
        assert _call("lsamd_fd_evaluate_adjoint_noalloc", grid, pts, gv, out) == 0
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
        assert float(ref.abs().sum()) > 0


def test_noalloc_rejects_too_many_points():
    from loopstructural_amd import _lib

    grid = _lib.make_grid((0, 0, 0), (0.1, 0.1, 0.1), (4, 4, 4))
    n = _lib.NOALLOC_MAX_POINTS + 1
    pts = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    gv = torch.zeros(n, dtype=torch.float64, device="cuda")
    out = torch.zeros(64, dtype=torch.float64, device="cuda")
    rc = _call("lsamd_fd_evaluate_adjoint_noalloc", grid, pts, gv, out)
    assert rc != 0
    assert b"scratch capacity" in _lib.load().lsamd_last_error()
