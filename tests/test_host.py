"""Host-side logic and the C ABI surface (no GPU needed)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import loopstructural_amd as L
from loopstructural_amd import _lib, constraints as cons, fd_system as fs
from loopstructural_amd.interpolators import _mask_kind

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    syms = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            syms |= set(re.findall(r"\b(lsamd_\w+)\s*\(", txt))
    return syms


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = _declared_symbols()
    assert len(syms) >= 15
    for s in sorted(syms):
        assert hasattr(lib, s), s
    assert set(_lib.SIGNATURES) >= syms - {"lsamd_version", "lsamd_last_error"}
    assert lib.lsamd_version() >= 1


def test_fault_injection_reports_error_without_touching_the_gpu(monkeypatch):
    monkeypatch.setenv("LSAMD_FAULT_INJECT", "locate,solve")
    g = _lib.make_grid((0, 0, 0), (1, 1, 1), (4, 4, 4))
    with pytest.raises(_lib.LsamdError) as e:
        _lib.call("lsamd_fd_locate", ctypes.byref(g), None, 1, None, None, None, None, None)
    assert e.value.code == _lib.LSAMD_ERR_INJECTED
    assert "locate" in str(e.value)


def test_argument_errors_are_reported():
    g = _lib.make_grid((0, 0, 0), (1, 1, 1), (1, 4, 4))  # 1 node on x: invalid
    with pytest.raises(_lib.LsamdError) as e:
        _lib.call("lsamd_fd_locate", ctypes.byref(g), None, 1, None, None, None, None, None)
    assert e.value.code == _lib.LSAMD_ERR_ARG
    with pytest.raises(_lib.LsamdError):
        _lib.call("lsamd_rows_normalise", 4, 200, None, None, None, None)


@pytest.mark.skipif(__import__("torch").cuda.is_available(), reason="checks the no-GPU path")
def test_compute_fails_loudly_without_gpu():
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.device()
    grid = L.StructuredGrid(nsteps=[4, 4, 4])
    interp = L.FiniteDifferenceInterpolator(grid)
    interp.set_value_constraints(np.array([[1.0, 1.0, 1.0, 0.0]]))
    with pytest.raises(RuntimeError):
        interp.setup_interpolator()


def test_value_validation_semantics():
    """LI/constraints.py:128-188, LI/_validation.py:25-85."""
    a = np.array([[0, 0, 0, 1.0, np.nan], [np.nan, 0, 0, 1, 1], [1, 1, 1, 2, 3]])
    v = cons.value_array(a)
    assert v.shape == (2, 5)
    assert v[0, 4] == 1.0 and v[1, 4] == 3.0
    v4 = cons.value_array(a[:, :4][[0, 2]])
    assert np.all(v4[:, 4] == 1.0)
    with pytest.raises(cons.ShapeError):
        cons.value_array(np.zeros((3, 7)))
    with pytest.raises(cons.DtypeError):
        cons.value_array([["a", "b", "c", "d"]])


def test_gradient_validation_drops_zero_vectors():
    a = np.array([[0, 0, 0, 0, 0, 0, 1.0], [0, 0, 0, 0, 0, 1, 2], [0, 0, 0, 1e-20, 0, 0, 1]])
    g = cons.gradient_array(a, is_normal=True)
    assert g.shape == (1, 7) and g[0, 6] == 2.0
    with pytest.raises(cons.ShapeError):
        cons.gradient_array(np.zeros((2, 5)))


def test_interpolator_set_constraints_errors_are_validation_errors():
    grid = L.StructuredGrid(nsteps=[4, 4, 4])
    interp = L.FiniteDifferenceInterpolator(grid)
    with pytest.raises(L.ValidationError):
        interp.set_value_constraints(np.zeros((3, 2)))
    interp.set_normal_constraints(np.array([[0, 0, 0, 0, 0, 1.0]]))
    assert interp.get_norm_constraints().shape == (1, 7) and interp.n_n == 1
    assert not interp.up_to_date


def test_structured_grid_geometry():
    """LC/supports/_3d_base_structured.py:24-245 and the reference's
    test_create_support_origin_nsteps (tests/unit/interpolator/test_discrete_supports.py)."""
    g = L.StructuredGrid(origin=np.zeros(3), nsteps_cells=np.array([10, 10, 10]),
                         step_vector=np.array([0.1, 0.1, 0.1]))
    assert np.sum(g.maximum - np.ones(3)) == 0
    assert np.array_equal(g.nsteps, [11, 11, 11]) and g.n_nodes == 1331 and g.n_elements == 1000
    d = L.StructuredGrid()
    assert np.sum(d.maximum - np.ones(3) * 10) == 0
    nodes = g.nodes
    assert nodes.shape == (1331, 3) and np.allclose(nodes[1], [0.1, 0, 0])
    assert np.allclose(nodes[11], [0, 0.1, 0])
    with pytest.raises(Exception):
        L.StructuredGrid(nsteps=[0, 2, 2])


def test_regulariser_struct_and_mask_mapping():
    w = {"dxx": 0.1, "dyy": 0.2, "dzz": 0.3, "dxy": 0.4, "dxz": 0.5, "dyz": 0.6,
         "dx": 1.0, "dy": 0.0, "dz": 2.0}
    r = fs.reg_struct(w, (5, 5, 2), 1e-8)
    assert np.isclose(r.w2_xx, 0.01) and np.isclose(r.w2_xy, (0.4 / 4) ** 2)
    assert r.w2_bx == 1.0 and r.w2_by == 0.0 and r.w2_bz == 0.0  # dy=0; nz<3
    assert np.isclose(r.ridge2, 1e-16)
    assert _mask_kind(L.Operator.Dxx_mask) == "xx"
    assert _mask_kind(L.Operator.Dyz_mask) == "yz"
    assert _mask_kind(3.0 * L.Operator.Dxz_mask) == "xz"
    assert _mask_kind(L.Operator.Lapacian) is None


def test_interface_pair_modes():
    grid = L.StructuredGrid(nsteps=[4, 4, 4])
    interp = L.FiniteDifferenceInterpolator(grid)
    assert interp._interface_pair_indices(4).shape == (6, 2)
    interp.interface_pair_mode = "chain"
    assert np.array_equal(interp._interface_pair_indices(3), [[0, 1], [1, 2]])
    interp.interface_pair_mode = "auto"
    interp.interface_pair_threshold = 3
    assert np.array_equal(interp._interface_pair_indices(5)[:, 0], np.zeros(4))


def test_to_dict_roundtrip_fields():
    grid = L.StructuredGrid(nsteps=[3, 3, 3])
    interp = L.FiniteDifferenceInterpolator(grid)
    interp.set_value_constraints(np.array([[1.0, 1, 1, 2.0, 1.0]]))
    d = interp.to_dict()
    assert d["type"] == "FINITE_DIFFERENCE" and len(d["c"]) == 64
    assert d["support"]["nsteps"] == [3, 3, 3]
    assert d["data"]["value"] == [[1.0, 1.0, 1.0, 2.0, 1.0]]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stdout


def test_factory_creates_and_rebuilds_interpolators():
    """InterpolatorFactory (LI/_interpolator_factory.py:40-140): type strings,
    support-given creation, refusals, and a to_dict -> from_dict round trip."""
    from loopstructural_amd import InterpolatorFactory, InterpolatorType

    grid = L.StructuredGrid(origin=[1.0, 2.0, 3.0], nsteps=[4, 5, 6], step_vector=[0.5, 0.5, 0.25])
    it = InterpolatorFactory.create_interpolator("FDI", support=grid, solver="lsmr")
    assert isinstance(it, L.FiniteDifferenceInterpolator) and it.solver == "lsmr"
    mesh = L.TetMesh(origin=[0, 0, 0], nsteps_cells=[3, 4, 5], step_vector=[1, 1, 1])
    for name in ("PLI", "P1", InterpolatorType.PIECEWISE_LINEAR, "PIECEWISE_LINEAR"):
        assert isinstance(InterpolatorFactory.create_interpolator(name, support=mesh),
                          L.P1Interpolator)
    with pytest.raises(ValueError):
        InterpolatorFactory.create_interpolator(None, support=grid)
    with pytest.raises(ValueError):
        InterpolatorFactory.create_interpolator("FDI")
    with pytest.raises(NotImplementedError):
        InterpolatorFactory.create_interpolator("P2", support=mesh)
    flat = _Box([0, 0, 0], [1, 1, 1])
    flat.dimensions = 2
    with pytest.raises(NotImplementedError):
        InterpolatorFactory.create_interpolator("FDI", boundingbox=flat, nelements=100)

    vals = np.array([[1.5, 2.5, 3.5, 2.0, 1.0], [2.0, 3.0, 4.0, -1.0, 0.5]])
    nrms = np.array([[1.5, 2.5, 3.5, 0.0, 0.0, 1.0, 1.0]])
    for interp in (L.FiniteDifferenceInterpolator(grid), L.P1Interpolator(
            L.TetMesh(origin=[1.0, 2.0, 3.0], nsteps_cells=[4, 5, 6],
                      step_vector=[0.5, 0.5, 0.25]))):
        interp.set_value_constraints(vals)
        interp.set_normal_constraints(nrms)
        interp.c = np.arange(interp.dof, dtype=float)
        d = interp.to_dict()
        back = InterpolatorFactory.from_dict(d)
        assert type(back) is type(interp)
        assert np.array_equal(back.support.nsteps, interp.support.nsteps)
        assert np.allclose(back.support.origin, interp.support.origin)
        assert np.allclose(back.support.step_vector, interp.support.step_vector)
        assert np.array_equal(back.data["value"], interp.data["value"])
        assert np.array_equal(back.data["normal"], interp.data["normal"])
        assert np.array_equal(back.c, interp.c)
        assert back.up_to_date == interp.up_to_date


class _Box:
    """Minimal stand-in for a caller's bounding box: translation-only local
    frame (project subtracts the origin), a settable element count that
    picks cubic cells, and a buffered copy."""

    def __init__(self, origin, maximum):
        self.origin = np.asarray(origin, float)
        self.maximum = np.asarray(maximum, float)
        self.nsteps = np.array([2, 2, 2])
        self.dimensions = 3

    @property
    def nelements(self):
        return int(np.prod(self.nsteps))

    @nelements.setter
    def nelements(self, n):
        h = (np.prod(self.maximum - self.origin) / n) ** (1 / 3)
        self.nsteps = np.ceil((self.maximum - self.origin) / h).astype(int)

    def project(self, pts):
        return np.asarray(pts, float) - self.origin

    def with_buffer(self, buffer=0.2):
        pad = (self.maximum - self.origin) * buffer
        return _Box(self.origin - pad, self.maximum + pad)


@pytest.mark.parametrize("kind, cls", [("FDI", "StructuredGrid"), ("PLI", "TetMesh")])
def test_factory_builds_support_from_bounding_box(kind, cls):
    from loopstructural_amd import InterpolatorFactory

    box = _Box([10.0, -5.0, 2.0], [14.0, -3.0, 3.0])
    it = InterpolatorFactory.create_interpolator(kind, boundingbox=box, nelements=4000)
    sup = it.support
    assert type(sup).__name__ == cls and it.bounding_box is box
    # the support spans the box in its local frame, with the box's cell counts
    assert np.allclose(sup.origin, 0.0)
    assert np.allclose(sup.maximum, box.maximum - box.origin)
    cells = np.asarray(sup.nsteps) - 1
    assert np.array_equal(cells, box.nsteps)
    assert 4000 <= np.prod(cells) <= 4000 * 1.6
    # element_volume sets the count; buffer enlarges the box, not the caller's
    it2 = InterpolatorFactory.create_interpolator(kind, boundingbox=box, element_volume=0.01,
                                                  buffer=0.1)
    assert np.allclose(it2.support.maximum - it2.support.origin, 1.2 * (box.maximum - box.origin))
    assert it2.bounding_box is box
