"""Every form of the regular FD stencil gives the same bits: z-chunks 8, 16
and 32 with register-staged planes on 64 x 8 and 32 x 16 tiles and with
planes copied straight into LDS (LSAMD_STENCIL_ZC pins the form per handle),
on grids with partial xy tiles,
short and long z extents, in FP32 and FP64; and the FP64 result equals the
oracle's explicit normal operator to 1e-13."""
import os

import numpy as np
import pytest
import torch

from oracle import fd as ofd

pytestmark = pytest.mark.gpu

W = {"dxx": 0.1, "dyy": 0.1, "dzz": 0.1, "dxy": 0.1, "dxz": 0.1, "dyz": 0.1,
     "dx": 1.0, "dy": 1.0, "dz": 1.0}
FORMS = (8, 16, 32, 1008, 1016, 1032, 2008, 2016, 2032)


def _apply(nn, form, x):
    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    os.environ["LSAMD_STENCIL_ZC"] = str(form)
    try:
        g = _lib.make_grid((0, 0, 0), (0.1, 0.1, 0.1), nn)
        s = fs.FdSystem(g, fs.reg_struct(W, nn, 1e-8))
        y = s.apply(x)
        torch.cuda.synchronize()
        info = s.stencil_info("fp32" if x.dtype == torch.float32 else "fp64")
        s.close()
    finally:
        del os.environ["LSAMD_STENCIL_ZC"]
    assert info["zc"] == form
    return y


@pytest.mark.parametrize("nn", [(9, 11, 13), (64, 64, 40), (70, 37, 45), (130, 20, 57)])
def test_stencil_forms_bit_identical(nn):
    n = int(np.prod(nn))
    gen = torch.Generator(device="cuda").manual_seed(sum(nn))
    x64 = torch.rand(n, dtype=torch.float64, device="cuda", generator=gen) - 0.5
    for x in (x64, x64.float()):
        print(f"stencil forms: grid {nn}, {x.dtype}", flush=True)
        ref = _apply(nn, 16, x)
        for form in FORMS:
            if form == 16:
                continue
            assert torch.equal(_apply(nn, form, x), ref), (nn, form, x.dtype)
    if n <= 40000:
        og = ofd.Grid((0, 0, 0), tuple(v - 1 for v in nn), (0.1, 0.1, 0.1))
        N, _ = ofd.normal_equations(ofd.assemble_fd(og, {}, W))
        want = N @ x64.cpu().numpy()
        got = _apply(nn, 1016, x64).cpu().numpy()
        assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
