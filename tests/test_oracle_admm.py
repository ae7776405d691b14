"""The oracle's ADMM restatement (LI/loopsolver/admm_solver.py:59-463) pinned
on the CPU: the brute-force active-set minimiser against scipy's SLSQP on a
dense problem, then ADMM with exact steps converging to that minimiser on a
small FD system with binding value inequalities and pair inequalities, and
to the unconstrained minimiser when no bound binds."""
import numpy as np
import pytest
from scipy import optimize, sparse

from oracle import admm as oadmm
from oracle import fd as ofd


def test_active_set_minimiser_matches_slsqp():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((12, 5))
    b = rng.standard_normal(12)
    Q = rng.standard_normal((3, 5))
    y_free = np.linalg.lstsq(A, b, rcond=None)[0]
    q = Q @ y_free
    bounds = np.column_stack([q - 0.3, q - 0.1])   # the free minimiser violates every row
    bounds[2] = [q[2] - 1.0, q[2] + 1.0]            # one row inactive
    y = oadmm.constrained_minimiser(sparse.csr_matrix(A), b, sparse.csr_matrix(Q), bounds)
    cons = [{"type": "ineq", "fun": lambda x, i=i: Q[i] @ x - bounds[i, 0]} for i in range(3)]
    cons += [{"type": "ineq", "fun": lambda x, i=i: bounds[i, 1] - Q[i] @ x} for i in range(3)]
    ref = optimize.minimize(lambda x: np.sum((A @ x - b) ** 2), y_free, method="SLSQP",
                            constraints=cons, options={"ftol": 1e-14, "maxiter": 500})
    assert np.allclose(y, ref.x, atol=1e-6)
    assert np.all(Q @ y >= bounds[:, 0] - 1e-9) and np.all(Q @ y <= bounds[:, 1] + 1e-9)


def _fd_problem(seed=3):
    og = ofd.Grid((0, 0, 0), (4, 4, 4), (0.25, 0.25, 0.25))
    rng = np.random.default_rng(seed)
    p = rng.uniform(0, 1, (30, 3))
    vals = np.column_stack([p, p[:, 2], np.ones(30)])
    sysm = ofd.assemble_fd(og, {"value": vals}, regularisation=0.1)
    A, b = sysm.build_matrix()
    A, b, S = ofd.preprocess(A, b)
    return og, A, b, S, rng


def test_admm_exact_steps_reach_constrained_minimiser():
    og, A, b, S, rng = _fd_problem()
    pi = rng.uniform(0.1, 0.9, (4, 3))
    # bounds on y = S^-1 c around values the free solution does not meet
    y_free = oadmm.constrained_minimiser(A, b, sparse.csr_matrix((0, og.n_nodes)),
                                         np.zeros((0, 2)))
    a, idc, _ = oadmm.value_inequality_rows(og, np.column_stack([pi, np.zeros((4, 2))]))
    Q, _ = oadmm.stack_inequalities([(a, idc, np.zeros((4, 2)))], og.n_nodes)
    q = Q @ y_free
    bounds = np.column_stack([q + 0.05 * np.abs(q) + 0.01, q + np.abs(q) + 1.0])
    bounds[3] = [q[3] - 1.0, q[3] + 1.0]
    y_star = oadmm.constrained_minimiser(A, b, Q, bounds)
    y = oadmm.admm_solve(A, b, Q, bounds, np.zeros(og.n_nodes), admm_weight=2.0, nmajor=3000,
                         inner="exact", admm_abs_tol=1e-13, admm_rel_tol=1e-13)
    assert np.linalg.norm(y - y_star) <= 1e-6 * np.linalg.norm(y_star)
    # feasible, some bound binds, and the free minimiser is not feasible
    qs = Q @ y_star
    assert np.all(qs >= bounds[:, 0] - 1e-9) and np.all(qs <= bounds[:, 1] + 1e-9)
    assert np.any(np.isclose(qs, bounds[:, 0], atol=1e-9, rtol=0))
    assert np.any(q < bounds[:, 0])
    # the reference's default (lsmr steps, admm_weight 0.01, 200 majors) moves towards it
    y_def = oadmm.admm_solve(A, b, Q, bounds, np.zeros(og.n_nodes))
    assert np.linalg.norm(y_def - y_star) < np.linalg.norm(y_free - y_star)


def test_admm_with_slack_bounds_is_the_free_minimiser():
    og, A, b, S, rng = _fd_problem(5)
    pi = rng.uniform(0.1, 0.9, (3, 3))
    a, idc, _ = oadmm.value_inequality_rows(og, np.column_stack([pi, np.zeros((3, 2))]))
    Q, _ = oadmm.stack_inequalities([(a, idc, np.zeros((3, 2)))], og.n_nodes)
    bounds = np.tile([-1e3, 1e3], (3, 1))
    y = oadmm.admm_solve(A, b, Q, bounds, np.zeros(og.n_nodes), admm_weight=0.3, nmajor=1500,
                         inner="exact", admm_abs_tol=0.0, admm_rel_tol=0.0)
    y_free = oadmm.constrained_minimiser(A, b, sparse.csr_matrix((0, og.n_nodes)),
                                         np.zeros((0, 2)))
    assert np.linalg.norm(y - y_free) <= 1e-8 * np.linalg.norm(y_free)


def test_pair_rows_order_and_values():
    og = ofd.Grid((0, 0, 0), (4, 4, 4), (0.25, 0.25, 0.25))
    rng = np.random.default_rng(9)
    pts = np.column_stack([rng.uniform(0, 1, (7, 3)), [0, 1, 0, 2, 1, 0, 2]])
    fams = oadmm.pair_inequality_rows(og, pts)
    assert [f[0] for f in fams] == ["inequality_pairs_0.0_1.0", "inequality_pairs_0.0_2.0",
                                    "inequality_pairs_1.0_2.0"]
    # on a linear field each row is f(upper) - f(lower)
    f = og.nodes() @ np.array([0.3, -0.2, 1.0])
    name, a, idc, bnd = fams[0]
    up = pts[pts[:, 3] == 0][:, :3] @ np.array([0.3, -0.2, 1.0])
    lo = pts[pts[:, 3] == 1][:, :3] @ np.array([0.3, -0.2, 1.0])
    want = (up[None, :] - lo[:, None]).ravel()  # upper index fastest
    assert np.allclose(np.sum(a * f[idc], axis=1), want, atol=1e-12)
    assert np.all(bnd[:, 0] == -np.inf) and np.all(bnd[:, 1] == 1.0)


@pytest.mark.parametrize("bad", ["sign", "dual"])
def test_planted_projection_mistakes_are_caught(bad, monkeypatch):
    """A projection with a flipped dual update, or a dual that is never
    accumulated, does not reach the constrained minimiser."""
    og, A, b, S, rng = _fd_problem()
    pi = rng.uniform(0.1, 0.9, (3, 3))
    a, idc, _ = oadmm.value_inequality_rows(og, np.column_stack([pi, np.zeros((3, 2))]))
    Q, _ = oadmm.stack_inequalities([(a, idc, np.zeros((3, 2)))], og.n_nodes)
    y_free = oadmm.constrained_minimiser(A, b, sparse.csr_matrix((0, og.n_nodes)),
                                         np.zeros((0, 2)))
    q = Q @ y_free
    bounds = np.column_stack([q + 0.2 * np.abs(q) + 0.05, q + np.abs(q) + 1.0])
    y_star = oadmm.constrained_minimiser(A, b, Q, bounds)
    good = oadmm._project

    def wrong(z, u, xmin, xmax, x):
        z2, u2 = good(z, u, xmin, xmax, x)
        return (z2, u - x + z2) if bad == "sign" else (z2, u)

    monkeypatch.setattr(oadmm, "_project", wrong)
    y = oadmm.admm_solve(A, b, Q, bounds, np.zeros(og.n_nodes), admm_weight=2.0, nmajor=400,
                         inner="exact", admm_abs_tol=0.0, admm_rel_tol=0.0)
    assert np.linalg.norm(y - y_star) > 1e-3 * np.linalg.norm(y_star)


class _CpuOps:
    """The x-update contract of loopstructural_amd.admm in node space c = S y,
    solved exactly with scipy: rows A (ridge included, unscaled) and Q S^-1."""

    def __init__(self, A, b, Q, S):
        import torch

        self.torch = torch
        self.A, self.b = sparse.csr_matrix(A), b
        self.S = S.diagonal()
        self.Qp = sparse.csr_matrix(Q) @ sparse.diags(1.0 / self.S)
        self.w = None

    def q(self, c):
        return self.torch.as_tensor(self.Qp @ c.numpy())

    def set_weights(self, w):
        self.w = w.numpy()

    def step(self, v, c):
        from scipy.sparse import linalg as spla

        W = sparse.diags(self.w)
        N = (self.A.T @ self.A + self.Qp.T @ W @ W @ self.Qp).tocsc()
        rhs = self.A.T @ self.b + self.Qp.T @ (self.w ** 2 * v.numpy())
        return self.torch.as_tensor(spla.spsolve(N, rhs))

    def y_norm(self, x):
        return float(np.linalg.norm(x.numpy() / self.S))


@pytest.mark.parametrize("kw", [
    dict(admm_weight=2.0, nmajor=40, admm_abs_tol=0.0, admm_rel_tol=0.0),
    dict(admm_weight=0.5, admm_weight_final=5.0, nmajor=30, admm_abs_tol=0.0, admm_rel_tol=0.0),
    dict(admm_weight=0.05, nmajor=60, adaptive_rho=True, admm_abs_tol=1e-9, admm_rel_tol=1e-9),
    dict(admm_weight=1.0, nmajor=40, active_set=True, admm_abs_tol=0.0, admm_rel_tol=0.0),
    dict(admm_weight=0.01, nmajor=200),
    dict(admm_weight=0.01, nmajor=200, inner_kwargs=dict(atol=1e-12, btol=1e-12,
                                                         maxiter=100000)),
])
def test_package_admm_loop_matches_oracle(kw):
    """loopstructural_amd.admm's loop in node space, with exact steps, takes
    the oracle's iterates (y = S^-1 c).  Against the reference's inexact LSMR
    steps (btol 1e-6) the results differ by that inexactness (about 1.5e-4
    here); with the oracle's LSMR steps tightened to 1e-12 they agree to
    about 2e-11."""
    kw = dict(kw)
    inner_kwargs = kw.pop("inner_kwargs", None)
    import torch

    from loopstructural_amd import admm

    og = ofd.Grid((0, 0, 0), (4, 4, 4), (0.25, 0.25, 0.25))
    rng = np.random.default_rng(3)
    p = rng.uniform(0, 1, (30, 3))
    sysm = ofd.assemble_fd(og, {"value": np.column_stack([p, p[:, 2], np.ones(30)])},
                           regularisation=0.1)
    A0, b0 = sysm.build_matrix()
    Ar, br, _ = ofd.preprocess(A0, b0, scaling=False)   # ridge rows, unscaled
    As, bs, S = ofd.preprocess(A0, b0)
    pi = rng.uniform(0.1, 0.9, (4, 3))
    a, idc, _ = oadmm.value_inequality_rows(og, np.column_stack([pi, np.zeros((4, 2))]))
    Q, _ = oadmm.stack_inequalities([(a, idc, np.zeros((4, 2)))], og.n_nodes)
    q = Q @ (ofd.solve_exact(sysm) / S.diagonal())
    bounds = np.column_stack([q + 0.1 * np.abs(q) + 0.02, q + 5.0])
    ops = _CpuOps(Ar, br, Q, S)
    c = admm.admm_solve(ops, torch.as_tensor(bounds), torch.zeros(og.n_nodes, dtype=torch.float64),
                        **kw)
    inner = "lsmr" if kw["admm_weight"] == 0.01 else "exact"
    y = oadmm.admm_solve(As, bs, Q, bounds, np.zeros(og.n_nodes), inner=inner,
                         inner_kwargs=inner_kwargs, **kw)
    tol = 1e-8 if inner == "exact" else (1e-9 if inner_kwargs else 1e-3)
    assert np.linalg.norm(c.numpy() - S @ y) <= tol * np.linalg.norm(S @ y)
