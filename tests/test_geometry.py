"""Rotation and anisotropy conventions, pinned on the CPU.

The reference's SRF index-map tests (``tests/test_srf.py:85-225``) assert
that a field under a rotation or anisotropy equals the unrotated field at
mapped grid indices.  Both fields come from the same modes, so the maps hold
exactly when the isometrized coordinates of the two points coincide.  Here
every one of those maps is checked on ``isometrize`` itself (the positions
the generator receives, ``covmodel/base.py:572-582``): a wrong sign, a
transposed matrix or a swapped rotation plane moves the mapped point by
O(1).  The GPU versions of the same tests run the whole SRF.
"""

import numpy as np
import pytest

from gstools_amd import Gaussian
from gstools_amd.geometry import (generate_grid, matrix_anisometrize, matrix_derotate, matrix_isometrize,
                                  matrix_rotate)

X_GRID = np.linspace(0.0, 12.0, 48)
Y_GRID = np.linspace(0.0, 10.0, 46)
Z_GRID = np.linspace(0.0, 10.0, 40)
C_GRID = np.linspace(-6.0, 6.0, 8)


def _iso_structured(model, axes):
    """Isometrized coordinates, shaped (d, *grid)."""
    shape = tuple(len(a) for a in axes)
    return model.isometrize(generate_grid(axes)).reshape((model.dim,) + shape)


def _same(a, b):
    assert np.allclose(a, b, rtol=0, atol=1e-12), (a, b)


def test_anisotropy_2d_index_map():
    m = Gaussian(dim=2, var=1.5, len_scale=4.0)
    iso = _iso_structured(m, (X_GRID, Y_GRID))
    m.anis = 0.5
    ani = _iso_structured(m, (X_GRID, Y_GRID))
    for a, b in (((0, 0), (0, 0)), ((0, 4), (0, 2)), ((0, 10), (0, 5))):
        _same(iso[(slice(None),) + a], ani[(slice(None),) + b])


def test_anisotropy_3d_index_map():
    m = Gaussian(dim=2, var=1.5, len_scale=4.0)
    m.dim = 3
    iso = _iso_structured(m, (X_GRID, Y_GRID, Z_GRID))
    m.anis = (0.5, 4.0)
    ani = _iso_structured(m, (X_GRID, Y_GRID, Z_GRID))
    for a, b in (((0, 0, 0), (0, 0, 0)), ((0, 4, 0), (0, 2, 0)), ((0, 10, 0), (0, 5, 0)),
                 ((0, 0, 1), (0, 0, 4)), ((0, 0, 3), (0, 0, 12))):
        _same(iso[(slice(None),) + a], ani[(slice(None),) + b])


def test_rotation_unstructured_2d_index_map():
    m = Gaussian(dim=2, var=1.5, len_scale=4.0)
    x_u, y_u = np.meshgrid(C_GRID, C_GRID)  # 'xy' indexing, as the reference test
    pos = np.array([x_u.ravel(), y_u.ravel()])
    m.anis = 0.25
    f = m.isometrize(pos).reshape(2, 8, 8)
    m.angles = -np.pi / 2.0
    r = m.isometrize(pos).reshape(2, 8, 8)
    _same(f[:, 0, 0], r[:, -1, 0])
    _same(f[:, 1, 2], r[:, -3, 1])


def test_rotation_structured_2d_index_map():
    m = Gaussian(dim=2, var=1.5, len_scale=4.0, anis=0.25)
    f = _iso_structured(m, (C_GRID, C_GRID))
    m.angles = -np.pi / 2.0
    r = _iso_structured(m, (C_GRID, C_GRID))
    _same(f[:, 0, 0], r[:, 0, -1])
    _same(f[:, 1, 2], r[:, 2, 6])


def test_rotation_unstructured_3d_index_map():
    m = Gaussian(dim=3, var=1.5, len_scale=4.0, anis=(0.25, 0.5))
    x_u, y_u, z_u = np.meshgrid(C_GRID, C_GRID, C_GRID)
    pos = np.array([x_u.ravel(), y_u.ravel(), z_u.ravel()])
    f = m.isometrize(pos).reshape(3, 8, 8, 8)
    m.angles = (-np.pi / 2.0, -np.pi / 2.0)
    r = m.isometrize(pos).reshape(3, 8, 8, 8)
    _same(f[:, 0, 0, 0], r[:, -1, -1, 0])
    _same(f[:, 1, 2, 0], r[:, -3, -1, 1])
    _same(f[:, 0, 0, 1], r[:, -1, -2, 0])


def test_rotation_structured_3d_index_map():
    m = Gaussian(dim=2, var=1.5, len_scale=4.0)
    m.dim = 3
    m.anis = 0.25
    assert np.array_equal(m.anis, [1.0, 0.25])  # padded with leading ones (geometric.py:94-123)
    grid = (C_GRID, C_GRID, C_GRID)
    f = _iso_structured(m, grid)
    m.angles = -np.pi / 2.0
    r = _iso_structured(m, grid)
    _same(f[:, 0, 0, 0], r[:, 0, 7, 0])
    _same(f[:, 0, 0, 1], r[:, 0, 7, 1])
    m.angles = (0, -np.pi / 2.0)
    r = _iso_structured(m, grid)
    _same(f[:, 0, 0, 0], r[:, 7, 0, 0])
    _same(f[:, 0, 1, 0], r[:, 7, 1, 0])
    _same(f[:, 1, 1, 0], r[:, 7, 1, 1])


@pytest.mark.parametrize("dim", [2, 3])
def test_rotation_matrices_closed_forms(dim):
    """Givens planes in the reference's order with alternating signs
    (``geometric.py:142-234``): 2D is the plain rotation by +angle; 3D is
    R = G12(c) G02(-b) G01(a); derotate is its inverse and transpose."""
    rs = np.random.RandomState(dim)
    ang = rs.uniform(-3, 3, size=dim * (dim - 1) // 2)
    rot = matrix_rotate(dim, ang)
    if dim == 2:
        c, s = np.cos(ang[0]), np.sin(ang[0])
        ref = np.array([[c, -s], [s, c]])
    else:
        def giv(i, j, t):
            g = np.eye(3)
            g[i, i] = g[j, j] = np.cos(t)
            g[i, j], g[j, i] = -np.sin(t), np.sin(t)
            return g
        ref = giv(1, 2, ang[2]) @ giv(0, 2, -ang[1]) @ giv(0, 1, ang[0])
    assert np.allclose(rot, ref, atol=1e-15)
    assert np.allclose(matrix_derotate(dim, ang) @ rot, np.eye(dim), atol=1e-14)
    assert np.allclose(matrix_derotate(dim, ang), rot.T, atol=1e-15)
    anis = rs.uniform(0.2, 3.0, size=dim - 1)
    assert np.allclose(matrix_isometrize(dim, ang, anis) @ matrix_anisometrize(dim, ang, anis), np.eye(dim),
                       atol=1e-13)


def test_folded_wavenumbers_equal_isometrized_phases():
    """k'.x with k' = M^T k equals k.(M x): the folding the structured path
    uses for rotated / anisotropic grids (phases agree to rounding)."""
    m = Gaussian(dim=3, var=2.0, len_scale=[50, 20, 5], angles=[0.5, 0.2, 0.1])
    mat = matrix_isometrize(3, m.angles, m.anis)
    rs = np.random.RandomState(0)
    k = rs.normal(size=(3, 64)) * 0.3
    x = generate_grid((np.arange(0.0, 1000.0, 97.0), np.arange(0.0, 800.0, 89.0), np.arange(0.0, 600.0, 53.0)))
    direct = k.T @ (mat @ x)
    folded = (mat.T @ k).T @ x
    assert np.allclose(folded, direct, rtol=0, atol=1e-12 * np.max(np.abs(direct)))
