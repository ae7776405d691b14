"""Adjoint gradients of a held-out misfit through the FD solve.

Checked against central finite differences (h = 1e-4) of the GPU solve and,
independently of the adjoint formula, against central differences of the CPU
oracle's exact least-squares solution (spsolve).  Tolerance: 1e-4 relative to
the largest gradient component.
"""
import numpy as np
import pytest
import torch

from oracle import fd as ofd

pytestmark = pytest.mark.gpu

CELLS = 14


def _setup(precision="fp64", rtol=1e-13):
    from loopstructural_amd import synthetic
    from loopstructural_amd.adjoint import FDProblem

    values, normals = synthetic.constraints(150, 30, 77, value_noise=0.02)
    rng = np.random.default_rng(78)
    held = rng.uniform(0.02, 0.98, (250, 3))
    f_held, _ = synthetic.field1(held)
    prob = FDProblem((0, 0, 0), (CELLS,) * 3, (1.0 / CELLS,) * 3, values[:, :3], normals[:, :3],
                     regularisation=0.1, precision=precision, rtol=rtol)
    dev = prob.dev
    params = [torch.tensor(values[:, 3], device=dev, requires_grad=True),
              torch.tensor(rng.uniform(0.5, 2.0, 150), device=dev, requires_grad=True),
              torch.tensor(normals[:, 3:6], device=dev, requires_grad=True),
              torch.tensor(rng.uniform(0.5, 2.0, 30), device=dev, requires_grad=True)]
    target = torch.tensor(f_held, device=dev)

    def loss_fn(ps):
        c = prob.solve(*ps)
        return torch.mean((prob.evaluate(c, held) - target) ** 2)

    return prob, params, loss_fn, values, normals, held, f_held


def _pick(params, k, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        i = int(rng.integers(0, len(params)))
        j = int(rng.integers(0, params[i].numel()))
        out.append((i, j))
    return out


def test_adjoint_matches_central_differences_fp64():
    prob, params, loss_fn, *_ = _setup()
    loss = loss_fn(params)
    loss.backward()
    grads = [p.grad.detach().clone().reshape(-1) for p in params]
    gmax = max(float(g.abs().max()) for g in grads)
    h = 1e-4
    for i, j in _pick(params, 24, 1):
        with torch.no_grad():
            ps = [p.detach().clone() for p in params]
            ps[i].view(-1)[j] += h
            lp = float(loss_fn(ps))
            ps[i].view(-1)[j] -= 2 * h
            lm = float(loss_fn(ps))
        fd = (lp - lm) / (2 * h)
        assert abs(fd - float(grads[i][j])) <= 1e-4 * gmax, (i, j, fd, float(grads[i][j]))


def test_adjoint_matches_oracle_differences():
    """Gradient w.r.t. values and weights vs central differences of the CPU
    oracle's exact solve (no shared code with the adjoint path)."""
    prob, params, loss_fn, values, normals, held, f_held = _setup()
    loss_fn(params).backward()
    og = ofd.Grid((0, 0, 0), (CELLS,) * 3, (1.0 / CELLS,) * 3)

    def oracle_loss(v, vw, n, nw):
        vals = np.column_stack([values[:, :3], v, vw])
        nrm = np.column_stack([normals[:, :3], n, nw])
        s = ofd.assemble_fd(og, {"value": vals, "normal": nrm}, regularisation=0.1)
        c = ofd.solve_exact(s)
        return np.mean((ofd.evaluate_value(og, c, held) - f_held) ** 2)

    base = [p.detach().cpu().numpy().copy() for p in params]
    grads = [p.grad.detach().cpu().numpy().reshape(-1) for p in params]
    gmax = max(float(np.abs(g).max()) for g in grads)
    h = 1e-4
    for i, j in _pick(params, 6, 2):
        ps = [b.copy() for b in base]
        ps[i].reshape(-1)[j] += h
        lp = oracle_loss(*ps)
        ps[i].reshape(-1)[j] -= 2 * h
        lm = oracle_loss(*ps)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grads[i][j]) <= 1e-4 * gmax, (i, j, fd, grads[i][j])


@pytest.mark.parametrize("form", ["fp32", "mg"])
def test_mixed_precision_and_multigrid_adjoint_close_to_fp64(form):
    _, p64, loss64, *_ = _setup("fp64", 1e-13)
    loss64(p64).backward()
    _, p2, loss2, *_ = _setup(form, 1e-11)
    l2 = loss2(p2)
    l2.backward()
    for a, b in zip(p64, p2):
        ga, gb = a.grad.reshape(-1), b.grad.reshape(-1)
        assert float((ga - gb).abs().max()) <= 1e-4 * float(ga.abs().max())


def test_evaluate_adjoint_is_transpose():
    """<E c, g> == <c, E^T g> for random c, g (E^T by the sort-based kernel)."""
    from loopstructural_amd.adjoint import FDEvaluateFunction
    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    g = _lib.make_grid((0, 0, 0), (0.1, 0.1, 0.1), (11, 11, 11))
    w = {k: 0.1 for k in ("dxx", "dyy", "dzz", "dxy", "dxz", "dyz")}
    system = fs.FdSystem(g, fs.reg_struct(w, (11, 11, 11), 1e-8))
    rng = np.random.default_rng(4)
    pts = torch.tensor(rng.uniform(-0.05, 1.05, (3000, 3)), device="cuda")
    c = torch.tensor(rng.standard_normal(1331), device="cuda", requires_grad=True)
    v = FDEvaluateFunction.apply(c, system, pts)
    gv = torch.tensor(rng.standard_normal(3000), device="cuda")
    gv[torch.isnan(v)] = 0.0
    v = torch.nan_to_num(v)
    (v * gv).sum().backward()
    lhs = float((v * gv).sum())
    rhs = float((c.detach() * c.grad).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    system.close()
