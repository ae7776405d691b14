"""P1 (5-tet) path on the GPU against the CPU oracle (oracle/p1.py).

Bit-exact: tet choice, barycentric coordinates, inside flags, elements,
neighbours, face pairs, value rows and evaluated values.  Floating-point
tolerances: element gradients (adjugate vs LU inverse) 1e-13 relative to the
largest entry; operator 1e-12 (FP64) and 1e-5 (FP32) relative to max |N x|;
solutions 1e-6 relative L2 against the exact least-squares minimiser.
"""
import numpy as np
import pytest
import torch

from oracle import fd as ofd
from oracle import p1 as op1

pytestmark = pytest.mark.gpu

MESHES = [((0.0, 0.0, 0.0), (6, 5, 7), (1 / 6, 1 / 5, 1 / 7)),
          ((10.0, -3.0, 2.5), (9, 8, 10), (0.37, 0.5, 0.25)),
          ((0.0, 0.0, 0.0), (2, 1, 3), (1.0, 1.0, 1.0))]


def _grid(m):
    from loopstructural_amd import _lib

    return _lib.make_grid(m.origin, m.step, m.nnodes)


def _points(m, n, seed):
    """Interior, outside, boundary, node, face and NaN points."""
    rng = np.random.default_rng(seed)
    ext = m.ncells * m.step
    p = m.origin + rng.uniform(-0.05, 1.05, (n, 3)) * ext
    nodes = m.nodes()[rng.integers(0, m.n_nodes, 50)]
    idx = rng.integers(0, m.nnodes, (60, 3)).astype(float)
    half = m.origin + (idx + rng.integers(0, 3, (60, 3)) * 0.5) * m.step  # edges/faces/centres
    quarter = m.origin + (idx + rng.integers(0, 5, (60, 3)) * 0.25) * m.step
    extra = np.array([m.origin, m.maximum, [np.nan, 0.5, 0.5], m.origin + 0.5 * ext])
    return np.ascontiguousarray(np.vstack([p, nodes, half, quarter, extra]))


@pytest.mark.parametrize("mi", range(len(MESHES)))
def test_locate_bit_exact(mi):
    from loopstructural_amd import fd_system as fs

    m = op1.TetMesh(*MESHES[mi])
    pts = _points(m, 3000, mi)
    vert, bc, tet, ins = m.element_for_location(pts)
    _, g_ref, _, _ = m.element_gradient_for_location(pts)
    t, nodes, gi, gbc, gg = fs.p1_locate(_grid(m), torch.from_numpy(pts), True, True)
    gi = gi.cpu().numpy()
    assert np.array_equal(gi, ins)
    assert np.array_equal(t.cpu().numpy(), tet)
    gbc = gbc.cpu().numpy()
    assert np.array_equal(gbc[ins], bc[ins])
    assert np.all(np.isnan(gbc[~ins]))
    nd = nodes.cpu().numpy()
    assert np.array_equal(nd[ins], m.elements[tet[ins]])
    gg = gg.cpu().numpy()
    scale = np.abs(g_ref[ins]).max()
    assert np.abs(gg[ins] - g_ref[ins]).max() <= 1e-13 * scale
    assert np.all(np.isnan(gg[~ins]))


@pytest.mark.parametrize("mi", range(len(MESHES)))
def test_topology_bit_exact(mi):
    from loopstructural_amd import fd_system as fs

    m = op1.TetMesh(*MESHES[mi])
    el, nb = fs.p1_elements(_grid(m), want_neighbours=True)
    assert np.array_equal(el.cpu().numpy(), m.elements)
    assert np.array_equal(nb.cpu().numpy(), m.neighbours())
    pairs, faces = fs.p1_pairs(_grid(m), want_faces=True)
    rp, rf = m.face_table()
    assert np.array_equal(pairs.cpu().numpy(), rp)
    assert np.array_equal(np.sort(faces.cpu().numpy(), 1), np.sort(rf, 1))
    g, size = fs.p1_tet_gradients(_grid(m), torch.arange(m.ntetra), want_size=True)
    ref = m.element_gradients()
    assert np.abs(g.cpu().numpy() - ref).max() <= 1e-13 * np.abs(ref).max()
    es = m.element_size()
    assert np.allclose(size.cpu().numpy(), es, rtol=1e-13)


@pytest.mark.parametrize("mi", range(len(MESHES)))
def test_edge_rows_match_oracle_up_to_row_sign(mi):
    from loopstructural_amd import fd_system as fs

    m = op1.TetMesh(*MESHES[mi])
    rows = fs.p1_edge_rows(_grid(m), 0.1, m.n_nodes)
    a, idc = op1.edge_jump_rows(m)
    an = a / np.linalg.norm(a, axis=1)[:, None]
    cols = rows.cols.cpu().numpy().reshape(-1, 8)
    coef = rows.coef.cpu().numpy().reshape(-1, 8)
    assert np.array_equal(cols, idc)
    sgn = np.sign(np.sum(coef * an, axis=1))
    assert np.all(sgn != 0)
    assert np.abs(coef - sgn[:, None] * an).max() <= 1e-13
    assert np.all(rows.w.cpu().numpy() == 0.1)
    # weight-scaled pairs (use_regularisation_weight_scale)
    scale = np.random.default_rng(1).uniform(1, 11, m.n_nodes)
    rs = fs.p1_edge_rows(_grid(m), 0.1, m.n_nodes, torch.from_numpy(scale))
    rel, _ = m.face_table()
    el = m.elements
    ref_w = 0.5 * (scale[el[rel[:, 0]]].mean(axis=1) + scale[el[rel[:, 1]]].mean(axis=1)) * 0.1
    assert np.array_equal(rs.w.cpu().numpy(), ref_w)


def _data(m, seed, nv=120, nn=30, nt=20):
    rng = np.random.default_rng(seed)
    ext = m.ncells * m.step
    pv = m.origin + rng.uniform(-0.02, 1.02, (nv, 3)) * ext
    vals = np.column_stack([pv, np.sin(pv[:, 0]) + pv[:, 2], rng.uniform(0.5, 2, nv)])
    pn = m.origin + rng.uniform(0.0, 1.0, (nn, 3)) * ext
    nrm = rng.standard_normal((nn, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    normals = np.column_stack([pn, nrm, np.ones(nn)])
    pt = m.origin + rng.uniform(0.0, 1.0, (nt, 3)) * ext
    tang = np.column_stack([pt, rng.standard_normal((nt, 3)), np.ones(nt)])
    return {"value": vals, "normal": normals, "tangent": tang}


def _family(s, name):
    c = s.constraints[name]
    return c["matrix"].tocsr(), c["b"], c["w"]


@pytest.mark.parametrize("mi", [0, 1])
def test_rows_match_oracle(mi):
    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    m = op1.TetMesh(*MESHES[mi])
    data = _data(m, 10 + mi)
    s = op1.assemble_p1(m, data, weights={"npw": 2.0, "cpw": 3.0, "tpw": 0.5})
    g = _grid(m)

    def gpu(kind, arr, ncols, w):
        pts = torch.from_numpy(np.ascontiguousarray(arr[:, :ncols]))
        _, _, ins, _, _ = fs.p1_locate(g, pts[:, :3].contiguous(), False)
        pin = pts.to(ins.device)[ins].contiguous()
        return fs.p1_build_rows(g, kind, pin, w, m.n_nodes)

    # value rows: bit-exact
    A, b, w = _family(s, "value")
    blk = gpu(_lib.ROWS_VALUE, data["value"], 4, 3.0)
    G = blk.to_scipy()
    assert np.array_equal(G.toarray(), A.toarray())
    assert np.array_equal(blk.b.cpu().numpy(), b)
    assert np.array_equal(blk.w.cpu().numpy(), w)
    # normal rows: point-major, gradient entries to a few ulp
    A, b, w = _family(s, "norm")
    blk = gpu(_lib.ROWS_NORMAL, data["normal"], 6, 2.0)
    assert np.abs(blk.to_scipy().toarray() - A.toarray()).max() <= 1e-13
    assert np.abs(blk.b.cpu().numpy() - b).max() <= 1e-13
    assert np.array_equal(blk.w.cpu().numpy(), w)
    # gradient orthogonal rows: vector normalised, weight * tet volume
    A, b, w = _family(s, "undefined gradient orthogonal constraint")
    tv = data["tangent"][:, :6].copy()
    nr = np.linalg.norm(tv[:, 3:6], axis=1)
    tv[:, 3:6] /= nr[:, None]
    blk = gpu(_lib.ROWS_TANGENT, tv, 6, 0.5)
    assert np.abs(blk.to_scipy().toarray() - A.toarray()).max() <= 1e-13
    assert np.allclose(blk.w.cpu().numpy(), w, rtol=1e-13)


@pytest.mark.parametrize("mi", range(len(MESHES)))
def test_evaluate(mi):
    from loopstructural_amd import fd_system as fs

    m = op1.TetMesh(*MESHES[mi])
    pts = _points(m, 2000, 20 + mi)
    c = np.random.default_rng(mi).standard_normal(m.n_nodes)
    v, gr = fs.p1_evaluate(_grid(m), torch.from_numpy(c), torch.from_numpy(pts), True, True)
    rv = m.evaluate_value(pts, c)
    rg = m.evaluate_gradient(pts, c)
    v = v.cpu().numpy()
    gr = gr.cpu().numpy()
    assert np.array_equal(np.isnan(v), np.isnan(rv))
    ok = ~np.isnan(rv)
    assert np.array_equal(v[ok], rv[ok])
    assert np.array_equal(np.isnan(gr), np.isnan(rg))
    ok = ~np.isnan(rg[:, 0])
    assert np.abs(gr[ok] - rg[ok]).max() <= 1e-12 * np.abs(rg[ok]).max()


@pytest.mark.parametrize("cells", [(6, 5, 7), (12, 9, 16), (3, 2, 2)])
def test_operator_matches_explicit_normal_equations(cells):
    from loopstructural_amd import fd_system as fs

    m = op1.TetMesh((0, 0, 0), cells, (0.3, 0.2, 0.25))
    data = _data(m, 5)
    s = op1.assemble_p1(m, data, regularisation=0.2)
    N, rhs = ofd.normal_equations(s)
    system = fs.P1System(_grid(m), 0.2, 1e-8)
    blocks = []
    for kind, key, ncols, w in ((1, "normal", 6, 1.0), (0, "value", 4, 1.0)):
        pts = torch.from_numpy(np.ascontiguousarray(data[key][:, :ncols]))
        _, _, ins, _, _ = fs.p1_locate(_grid(m), pts[:, :3].contiguous(), False)
        pin = pts.to(ins.device)[ins].contiguous()
        blocks.append(fs.p1_build_rows(_grid(m), kind, pin, w, m.n_nodes))
    tv = data["tangent"][:, :6].copy()
    tv[:, 3:6] /= np.linalg.norm(tv[:, 3:6], axis=1)[:, None]
    pts = torch.from_numpy(tv)
    _, _, ins, _, _ = fs.p1_locate(_grid(m), pts[:, :3].contiguous(), False)
    blocks.append(fs.p1_build_rows(_grid(m), 3, pts.to(ins.device)[ins].contiguous(), 1.0,
                                   m.n_nodes))
    system.set_rows(fs.concat_rows(blocks, m.n_nodes))
    x = np.random.default_rng(2).standard_normal(m.n_nodes)
    ref = N @ x
    y = system.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    y32 = system.apply(torch.from_numpy(x).float().cuda()).double().cpu().numpy()
    assert np.abs(y32 - ref).max() <= 1e-5 * np.abs(ref).max()
    assert np.abs(system.rhs().cpu().numpy() - rhs).max() <= 1e-13 * np.abs(rhs).max()
    assert np.allclose(system.diag().cpu().numpy(), N.diagonal(), rtol=1e-12, atol=1e-20)


def _interp(m, data, **kw):
    from loopstructural_amd import P1Interpolator, TetMesh

    mesh = TetMesh(origin=m.origin, nsteps_cells=m.ncells, step_vector=m.step)
    it = P1Interpolator(mesh)
    it.set_value_constraints(data["value"])
    it.set_normal_constraints(data["normal"])
    it.set_tangent_constraints(data["tangent"])
    it.setup_interpolator(**kw)
    return it


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_interpolator_solve_matches_minimiser(precision):
    m = op1.TetMesh((0, 0, 0), (10, 9, 11), (0.1, 0.111, 0.0909))
    data = _data(m, 7, nv=300, nn=60, nt=30)
    it = _interp(m, data, regularisation=0.1)
    it.solve_system("cg", solver_kwargs={"rtol": 1e-12, "precision": precision})
    ref = ofd.solve_exact(op1.assemble_p1(m, data, regularisation=0.1))
    assert np.linalg.norm(it.c - ref) <= 1e-6 * np.linalg.norm(ref)
    pts = _points(m, 500, 3)
    ev = it.evaluate_value(pts)
    rv = m.evaluate_value(pts, it.c)
    rv[np.any(np.isnan(pts), axis=1)] = 0.0
    assert np.array_equal(np.isnan(ev), np.isnan(rv))


def test_interpolator_weight_scaled_regulariser():
    m = op1.TetMesh((0, 0, 0), (8, 8, 8), (0.125, 0.125, 0.125))
    data = _data(m, 9, nv=200, nn=40, nt=0)
    data["tangent"] = np.zeros((0, 7))
    it = _interp(m, data, use_regularisation_weight_scale=True)
    it.solve_system("cg", solver_kwargs={"rtol": 1e-12})
    # the scale: 1 + 1 + 10 exp(-d^2 / 2 sigma^2), sigma = nsteps[0] * 10
    nodes = m.nodes()
    pn = data["normal"][:, :3]
    _, _, ins = m.element_for_location(pn)[1:]
    d2 = ((nodes[:, None, :] - pn[None, ins, :]) ** 2).sum(-1).min(1)
    sigma = m.nnodes[0] * 10
    scale = 1.0 + (1 + 10 * np.exp(-d2 / (2 * sigma ** 2)))
    s = op1.assemble_p1(m, data, scale=scale)
    ref = ofd.solve_exact(s)
    assert np.linalg.norm(it.c - ref) <= 1e-6 * np.linalg.norm(ref)


def test_horizontal_layers_reproduced():
    """tests/integration/test_interpolator.py:109-118: layers z = const are
    fitted to atol 1e-3 at the data points."""
    from loopstructural_amd import P1Interpolator, TetMesh

    rng = np.random.default_rng(0)
    pts = rng.uniform(0.05, 0.95, (400, 3))
    vals = np.column_stack([pts, pts[:, 2], np.ones(400)])
    it = P1Interpolator(TetMesh(origin=(0, 0, 0), nsteps_cells=(12, 12, 12),
                                step_vector=(1 / 12,) * 3))
    it.set_value_constraints(vals)
    it.setup_interpolator()
    it.solve_system("cg", solver_kwargs={"rtol": 1e-10})
    assert np.allclose(it.evaluate_value(pts), pts[:, 2], atol=1e-3)
    g = it.evaluate_gradient(pts)
    assert np.allclose(g, [0, 0, 1], atol=1e-3)
