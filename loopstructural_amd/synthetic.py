"""Synthetic constraint sets of the benchmark workloads (unit box).

Fields: f1 = z + 0.1 sin(2 pi x) cos(2 pi y) (values and unit normals of grad
f1) and f2 = z + 0.2 sin(3 pi x) sin(2 pi y).  Seeds use
``numpy.random.default_rng``.  Arrays follow the reference's constraint
layouts: values (N,5) [x,y,z,v,w], normals (N,7) [x,y,z,nx,ny,nz,w].
"""
from __future__ import annotations

import numpy as np


def field1(p):
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    f = z + 0.1 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    g = np.stack([0.2 * np.pi * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y),
                  -0.2 * np.pi * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y),
                  np.ones_like(x)], axis=1)
    return f, g


def field2(p):
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    f = z + 0.2 * np.sin(3 * np.pi * x) * np.sin(2 * np.pi * y)
    g = np.stack([0.6 * np.pi * np.cos(3 * np.pi * x) * np.sin(2 * np.pi * y),
                  0.4 * np.pi * np.sin(3 * np.pi * x) * np.cos(2 * np.pi * y),
                  np.ones_like(x)], axis=1)
    return f, g


def field3(p):
    """f = z + 0.1 sin(2 pi x) (the ensemble workload)."""
    x, z = p[:, 0], p[:, 2]
    f = z + 0.1 * np.sin(2 * np.pi * x)
    g = np.stack([0.2 * np.pi * np.cos(2 * np.pi * x), np.zeros_like(x), np.ones_like(x)],
                 axis=1)
    return f, g


def constraints(n_values, n_normals, seed, field=field1, value_noise=0.0, normal_tilt_deg=0.0):
    """(values (N,5), normals (M,7)) sampled uniformly in [0,1]^3."""
    rng = np.random.default_rng(seed)
    pv = rng.uniform(0.0, 1.0, (n_values, 3))
    f, _ = field(pv)
    if value_noise:
        f = f + rng.normal(0.0, value_noise, n_values)
    pn = rng.uniform(0.0, 1.0, (n_normals, 3))
    _, g = field(pn)
    n = g / np.linalg.norm(g, axis=1)[:, None]
    if normal_tilt_deg:
        # rotate each normal by the given angle about a random axis perpendicular to it
        a = rng.standard_normal((n_normals, 3))
        a -= np.sum(a * n, axis=1)[:, None] * n
        a /= np.linalg.norm(a, axis=1)[:, None]
        t = np.deg2rad(normal_tilt_deg)
        n = n * np.cos(t) + np.cross(a, n) * np.sin(t)
    values = np.column_stack([pv, f, np.ones(n_values)])
    normals = np.column_stack([pn, n, np.ones(n_normals)])
    return values, normals


def grid_points(n):
    """n^3 evaluation points from linspace(0,1,n), 'ij' meshgrid, flattened."""
    t = np.linspace(0.0, 1.0, n)
    xx, yy, zz = np.meshgrid(t, t, t, indexing="ij")
    return np.column_stack([xx.ravel(), yy.ravel(), zz.ravel()])


def fd_workload(cells, n_values, n_normals, seed):
    """Grid spec and constraints of a unit-box FD solve with ``cells`` per axis."""
    values, normals = constraints(n_values, n_normals, seed)
    return {"origin": np.zeros(3), "nsteps": np.array([cells] * 3),
            "step_vector": np.full(3, 1.0 / cells), "values": values, "normals": normals}


def ensemble(nsys, seed, nodes=(32, 96), n_values=(500, 5000), n_normals=(100, 1000)):
    """Uncertainty ensemble: per realisation a cubic grid with U{nodes} nodes
    per axis, U{n_values} values (noise sigma 0.02) and U{n_normals} normals
    (rotated 5 degrees about a random perpendicular axis) of field3."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(nsys):
        nn = int(rng.integers(nodes[0], nodes[1] + 1))
        nv = int(rng.integers(n_values[0], n_values[1] + 1))
        nm = int(rng.integers(n_normals[0], n_normals[1] + 1))
        v, n = constraints(nv, nm, int(rng.integers(0, 2**31)), field=field3, value_noise=0.02,
                           normal_tilt_deg=5.0)
        out.append((nn, v, n))
    return out


def perturb(values, normals, seed, value_noise=0.02, normal_tilt_deg=5.0):
    """One realisation of perturbed constraints (value noise, normal tilt)."""
    rng = np.random.default_rng(seed)
    v = values.copy()
    v[:, 3] += rng.normal(0.0, value_noise, v.shape[0])
    n = normals.copy()
    u = n[:, 3:6]
    a = rng.standard_normal(u.shape)
    a -= np.sum(a * u, axis=1)[:, None] * u
    a /= np.linalg.norm(a, axis=1)[:, None]
    t = np.deg2rad(normal_tilt_deg)
    n[:, 3:6] = u * np.cos(t) + np.cross(a, u) * np.sin(t)
    return v, n
