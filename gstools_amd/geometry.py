"""Rotation / anisotropy matrices and grids (``src/gstools/tools/geometric.py:63-356``).

``iso_pos = diag(1, 1/anis) . R^-1(angles) . pos`` is what ``Field.pre_pos``
hands to the generator (``field/base.py:254-297`` ->
``covmodel/base.py:572-582``).  The structured summation path never builds
the grid: it folds the isometrize matrix into the wavenumbers instead
(``k' = M^T k``), so these helpers are needed on the host only.
"""

import numpy as np

__all__ = [
    "set_angles", "set_anis", "no_of_angles", "rotation_planes", "givens_rotation",
    "matrix_rotate", "matrix_derotate", "matrix_isotropify", "matrix_isometrize",
    "generate_grid", "latlon2pos", "chordal_to_great_circle",
    "EARTH_RADIUS", "KM_SCALE", "DEGREE_SCALE", "RADIAN_SCALE",
]

# Geographic scales (``src/gstools/tools/__init__.py:112-121``): the earth's
# radius in km, and the factors turning great-circle angles (radians) into km,
# degrees or radians.
EARTH_RADIUS = 6371.0
KM_SCALE = 6371.0
DEGREE_SCALE = 57.29577951308232
RADIAN_SCALE = 1.0


def no_of_angles(dim):
    return (dim * (dim - 1)) // 2


def set_angles(dim, angles):
    """Pad / cut the angles to ``dim*(dim-1)/2`` entries (``geometric.py:63-91``)."""
    out = np.atleast_1d(np.asarray(angles, dtype=np.double))[: no_of_angles(dim)]
    return np.pad(out, (0, no_of_angles(dim) - len(out)), "constant", constant_values=0.0)


def set_anis(dim, anis):
    """Pad the anisotropy ratios to ``dim - 1`` entries with leading ones (``geometric.py:94-123``)."""
    out = np.atleast_1d(np.asarray(anis, dtype=np.double))[: dim - 1]
    if len(out) < dim - 1:
        out = np.pad(out, (dim - len(out) - 1, 0), "constant", constant_values=1.0)
    return out


def rotation_planes(dim):
    return [(i, j) for j in range(1, dim) for i in range(j)]


def givens_rotation(dim, plane, angle):
    res = np.eye(dim, dtype=np.double)
    res[plane[0], plane[0]] = np.cos(angle)
    res[plane[1], plane[1]] = np.cos(angle)
    res[plane[0], plane[1]] = -np.sin(angle)
    res[plane[1], plane[0]] = np.sin(angle)
    return res


def matrix_rotate(dim, angles):
    """Tait-Bryan rotation: Givens rotations with alternating signs (``geometric.py:184-207``)."""
    angles = set_angles(dim, angles)
    res = np.eye(dim, dtype=np.double)
    for i, (angle, plane) in enumerate(zip(angles, rotation_planes(dim))):
        res = np.matmul(givens_rotation(dim, plane, (-1) ** i * angle), res)
    return res


def matrix_derotate(dim, angles):
    """Inverse of :func:`matrix_rotate` (``geometric.py:210-234``)."""
    angles = -set_angles(dim, angles)
    res = np.eye(dim, dtype=np.double)
    for i, (angle, plane) in enumerate(zip(angles, rotation_planes(dim))):
        res = np.matmul(res, givens_rotation(dim, plane, (-1) ** i * angle))
    return res


def matrix_isotropify(dim, anis):
    return np.diag(np.concatenate(([1.0], 1.0 / set_anis(dim, anis))))


def matrix_isometrize(dim, angles, anis):
    """``diag(1, 1/anis) . derotate(angles)`` (``geometric.py:275-294``)."""
    return np.matmul(matrix_isotropify(dim, anis), matrix_derotate(dim, angles))


def matrix_anisotropify(dim, anis):
    return np.diag(np.concatenate(([1.0], set_anis(dim, anis))))


def matrix_anisometrize(dim, angles, anis):
    """``rotate(angles) . diag(1, anis)``, the inverse of :func:`matrix_isometrize` (``geometric.py:297-316``)."""
    return np.matmul(matrix_rotate(dim, angles), matrix_anisotropify(dim, anis))


def generate_grid(pos):
    """Meshgrid ('ij') of a structured position tuple -> (d, P) (``geometric.py:340-356``)."""
    return np.asarray(np.meshgrid(*pos, indexing="ij"), dtype=np.double).reshape((len(pos), -1))


def latlon2pos(latlon, radius=1.0, dtype=np.double):
    """(lat, lon) in degrees -> 3D points on a sphere of ``radius`` (``geometric.py:628-667``)."""
    latlon = np.asarray(latlon, dtype=dtype).reshape((2, -1))
    lat, lon = np.deg2rad(latlon[:2])
    return np.array((radius * np.cos(lat) * np.cos(lon), radius * np.cos(lat) * np.sin(lon),
                     radius * np.sin(lat) * np.ones_like(lon)), dtype=dtype)


def chordal_to_great_circle(dist, radius=1.0):
    """Chordal distance -> great-circle distance on a sphere, clipped to [0, 2 radius] (``geometric.py:712-733``)."""
    diameter = 2 * radius
    return diameter * np.arcsin(np.maximum(np.minimum(np.divide(dist, diameter), 1), 0))
