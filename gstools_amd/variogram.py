"""Empirical variograms whose pair binning runs on the GPU.

Drop-ins for the external variogram kernels GSTools imports
(``src/gstools/variogram/variogram.py:12-15``, dispatched at ``:44-113``):
``unstructured``, ``directional``, ``structured`` and ``ma_structured`` with
the same arguments and float64 results, computed by ``gsx_vario_*`` (FP64
distances with the reference's rounding, exact counts, sums reduced in a
fixed order).

:func:`vario_estimate`, :func:`vario_estimate_axis` and
:func:`standard_bins` restate the front end (``variogram.py:151-470``,
``:473-565``, ``binning.py:25-104``): stacked fields, masks and ``no_data``,
structured and unstructured positions, sampling, directions from vectors or
angles with tolerance and bandwidth, Matheron / Cressie estimators, counts,
lat-lon data (great-circle distances, ``geo_scale``), and mean / trend
removal.  Not restated: normalizers (``normalizer`` or ``fit_normalizer``
other than the defaults raise ``NotImplementedError``).
"""

import numpy as np
import torch

from . import _native
from .geometry import RADIAN_SCALE, chordal_to_great_circle, generate_grid, latlon2pos
from .krige import eval_func
from .summate import _as_f64, _device, _ptr, _stream, host_array

__all__ = [
    "unstructured", "directional", "structured", "ma_structured",
    "vario_estimate", "vario_estimate_axis", "standard_bins",
]

AXIS = ["x", "y", "z"]
AXIS_DIR = {"x": 0, "y": 1, "z": 2}


def _estimator_code(estimator_type):
    if estimator_type not in ("m", "c"):
        raise ValueError(f"Unknown estimator type: {estimator_type}")
    return 0 if estimator_type == "m" else 1


def _ws(nbytes, dev):
    if nbytes < 0:
        _native.check(int(nbytes), "variogram")
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=dev)


# ------------------------------------------------------------- kernels ----
def _tensor(x, name, dev, ndim):
    """float64 device tensor of ``ndim`` dimensions (leading axes added) from numpy or torch input."""
    t = _as_f64(x if isinstance(x, torch.Tensor) else np.asarray(x, dtype=np.double), name, dev)
    while t.dim() < ndim:
        t = t.unsqueeze(0)
    return t.contiguous()


def _pairs(f, bin_edges, pos, estimator_type, direction=None, angles_tol=np.pi / 8, bandwidth=-1.0,
           separate_dirs=False, haversine=False):
    dev = _device()
    ft = _tensor(f, "f", dev, 2)
    pt = _tensor(pos, "pos", dev, 2)
    et = _tensor(bin_edges, "bin_edges", dev, 1)
    nf, npnt = ft.shape
    dim = pt.shape[0]
    if pt.shape[1] != npnt:
        raise ValueError(f"pos has {pt.shape[1]} points, the fields {npnt}")
    if not 1 <= dim <= 3:
        raise ValueError("variogram: only 1 to 3 dimensions are supported")
    nb = max(et.numel() - 1, 0)
    est = _estimator_code(estimator_type)
    lib = _native.lib()
    if haversine:
        if dim != 2 or direction is not None:
            raise ValueError("variogram: great-circle distances need (lat, lon) positions and no direction")
        gamma = torch.zeros(nb, dtype=torch.float64, device=dev)
        counts = torch.zeros(nb, dtype=torch.int64, device=dev)
        ws = _ws(lib.gsx_vario_workspace(0, npnt, 2, nb, 1), dev)
        _native.check(lib.gsx_vario_unstructured_haversine_f64(nf, npnt, nb, _ptr(ft), _ptr(et), _ptr(pt), est,
                                                               _ptr(gamma), _ptr(counts), _ptr(ws), ws.numel(),
                                                               _stream()), "unstructured (haversine)")
        return gamma, counts
    if direction is None:
        gamma = torch.zeros(nb, dtype=torch.float64, device=dev)
        counts = torch.zeros(nb, dtype=torch.int64, device=dev)
        ws = _ws(lib.gsx_vario_workspace(0, npnt, dim, nb, 1), dev)
        _native.check(lib.gsx_vario_unstructured_f64(nf, npnt, dim, nb, _ptr(ft), _ptr(et), _ptr(pt), est, _ptr(gamma),
                                                     _ptr(counts), _ptr(ws), ws.numel(), _stream()), "unstructured")
        return gamma, counts
    dt = _tensor(direction, "direction", dev, 2)
    nd = dt.shape[0]
    if dt.shape[1] != dim:
        raise ValueError(f"Can't interpret directions: {direction}")
    gamma = torch.zeros((nd, nb), dtype=torch.float64, device=dev)
    counts = torch.zeros((nd, nb), dtype=torch.int64, device=dev)
    ws = _ws(lib.gsx_vario_workspace(1, npnt, dim, nb, nd), dev)
    _native.check(lib.gsx_vario_directional_f64(nf, npnt, dim, nb, nd, _ptr(ft), _ptr(et), _ptr(pt), _ptr(dt),
                                                float(angles_tol), float(bandwidth), int(bool(separate_dirs)), est,
                                                _ptr(gamma), _ptr(counts), _ptr(ws), ws.numel(), _stream()),
                  "directional")
    return gamma, counts


def unstructured(f, bin_edges, pos, estimator_type="m", distance_type="e", num_threads=None):
    """Drop-in for the backend ``unstructured`` -> (gamma (B,), counts (B,)).

    ``distance_type`` "e": Euclidean distances; "h": ``pos`` is (lat, lon) in
    degrees and a pair's distance its great-circle angle in radians.
    """
    if distance_type not in ("e", "h"):
        raise ValueError(f"Unknown distance type: {distance_type}")
    gamma, counts = _pairs(f, bin_edges, pos, estimator_type, haversine=distance_type == "h")
    return host_array(gamma), host_array(counts)


def directional(f, bin_edges, pos, direction, angles_tol=np.pi / 8.0, bandwidth=-1.0, separate_dirs=False,
                estimator_type="m", num_threads=None):
    """Drop-in for the backend ``directional`` -> (gamma (D, B), counts (D, B))."""
    gamma, counts = _pairs(f, bin_edges, pos, estimator_type, direction, angles_tol, bandwidth, separate_dirs)
    return host_array(gamma), host_array(counts)


def _lags(f, mask, estimator_type):
    dev = _device()
    ft = _tensor(f, "f", dev, 1)
    if ft.dim() == 1:
        ft = ft.reshape(-1, 1)
    n0, n1 = ft.shape
    mt = None
    if mask is not None:
        m = np.ascontiguousarray(np.asarray(mask, dtype=bool).reshape(n0, n1), dtype=np.uint8)
        mt = torch.from_numpy(m).to(dev)
    gamma = torch.zeros(n0, dtype=torch.float64, device=dev)
    counts = torch.zeros(n0, dtype=torch.int64, device=dev)
    lib = _native.lib()
    ws = _ws(lib.gsx_vario_workspace(2, n0, n1, 0, 1), dev)
    _native.check(lib.gsx_vario_structured_f64(n0, n1, _ptr(ft), _ptr(mt), _estimator_code(estimator_type),
                                               _ptr(gamma), _ptr(counts), _ptr(ws), ws.numel(), _stream()),
                  "structured")
    return gamma, counts


def structured(f, estimator_type="m", num_threads=None):
    """Drop-in for the backend ``structured``: lags along the first axis -> gamma (n0,)."""
    return host_array(_lags(f, None, estimator_type)[0])


def ma_structured(f, mask, estimator_type="m", num_threads=None):
    """Drop-in for the backend ``ma_structured`` (masked values skipped) -> gamma (n0,)."""
    return host_array(_lags(f, mask, estimator_type)[0])


# ------------------------------------------------------------ helpers -----
def _struct_pos_dim(pos, dim):
    if dim == 1:
        return (np.asarray(pos, dtype=np.double).reshape(-1),)
    if len(pos) != dim:
        raise ValueError("Formatting: position tuple doesn't match dimension.")
    return tuple(np.asarray(p, dtype=np.double).reshape(-1) for p in pos)


def _struct_pos_shape(pos, shape):
    """Axes and dimension of a structured position tuple for a stacked field (``geometric.py:428-515``)."""
    full, stacked = int(np.prod(shape)), int(np.prod(shape[1:]))
    try:
        check = np.array(pos, dtype=np.double, ndmin=2)
    except ValueError:
        dim = len(pos)
        axes = _struct_pos_dim(pos, dim)
        size = int(np.prod([len(a) for a in axes]))
        if size not in (stacked, full):
            raise ValueError("Formatting: position tuple doesn't match dimension.") from None
        return axes, dim
    if check.size in (full, stacked):
        return _struct_pos_dim(check, 1), 1
    size = int(np.prod([p.size for p in check]))
    if size in (full, stacked):
        dim = len(check)
        return _struct_pos_dim(pos, dim), dim
    raise ValueError("Formatting: position tuple doesn't match dimension.")


def _unstruct_pos_shape(pos, shape):
    """(dim, n) points of an unstructured position tuple for a stacked field (``geometric.py:518-584``)."""
    full, stacked = int(np.prod(shape)), int(np.prod(shape[1:]))
    pre_len = len(np.atleast_1d(pos))
    pos = np.atleast_2d(np.asarray(pos, dtype=np.double))
    dim = len(pos) if pre_len == len(pos) else 1
    pnt_cnt = pos[0].size
    ok = (dim == 1 and pos.size in (full, stacked)) or (dim > 1 and pnt_cnt in (full, stacked))
    if not ok:
        raise ValueError("Formatting: position tuple doesn't match dimension.")
    return pos.reshape((dim, -1)), dim


def ang2dir(angles, dtype=np.double, dim=None):
    """Direction vectors from n-D spherical angles (``geometric.py:587-625``)."""
    pre_dim = np.asanyarray(angles).ndim
    angles = np.atleast_2d(np.asarray(angles, dtype=dtype))
    if len(angles.shape) > 2:
        raise ValueError(f"Can't interpret angles array {angles}")
    dim = angles.shape[1] + 1 if dim is None else dim
    if dim == 2 and angles.shape[0] == 1 and pre_dim < 2:
        angles = angles.T
    if dim != angles.shape[1] + 1 or dim == 1:
        raise ValueError(f"Wrong dim. ({dim}) for angles {angles}")
    vec = np.empty((angles.shape[0], dim), dtype=dtype)
    vec[:, 0] = np.prod(np.sin(angles), axis=1)
    for i in range(1, dim):
        vec[:, i] = np.prod(np.sin(angles[:, i:]), axis=1)
        vec[:, i] *= np.cos(angles[:, (i - 1)])
    if dim in [2, 3]:
        vec[:, [0, 1]] = vec[:, [1, 0]]
    return vec


def _separate_dirs_test(direction, angles_tol):
    if direction is None or direction.shape[0] < 2:
        return False
    sep = True
    for i in range(direction.shape[0] - 1):
        for j in range(i + 1, direction.shape[0]):
            s_prod = np.minimum(np.abs(np.dot(direction[i], direction[j])), 1)
            sep &= np.arccos(s_prod) >= 2 * angles_tol
    return bool(sep)


def standard_bins(pos=None, dim=2, latlon=False, mesh_type="unstructured", bin_no=None, max_dist=None,
                  geo_scale=RADIAN_SCALE):
    """Sturges-rule bins up to a third of the bounding-box diagonal (``binning.py:29-104``);
    lat-lon points are taken on a sphere of radius ``geo_scale`` and the diagonal turned
    into a great-circle distance."""
    dim = 2 if latlon else int(dim)
    if bin_no is None or max_dist is None:
        if pos is None:
            raise ValueError("standard_bins: no pos tuple given.")
        if mesh_type != "unstructured":
            pos = generate_grid(_struct_pos_dim(pos, dim))
        else:
            pos = np.asarray(pos, dtype=np.double).reshape(dim, -1)
        pos = latlon2pos(pos, radius=geo_scale) if latlon else pos
        pnt_cnt = len(pos[0])
        box = np.asarray([[np.min(axis), np.max(axis)] for axis in pos])
        diam = np.linalg.norm(box[:, 0] - box[:, 1])
        diam = chordal_to_great_circle(diam, geo_scale) if latlon else diam
        bin_no = int(np.ceil(2 * np.log2(pnt_cnt) + 1)) if bin_no is None else int(bin_no)
        max_dist = diam / 3 if max_dist is None else float(max_dist)
    return np.linspace(0, max_dist, num=bin_no + 1, dtype=np.double)


# ----------------------------------------------------------- front end ----
def vario_estimate(pos, field, bin_edges=None, sampling_size=None, sampling_seed=None, estimator="matheron",
                   latlon=False, direction=None, angles=None, angles_tol=np.pi / 8, bandwidth=None, no_data=np.nan,
                   mask=np.ma.nomask, mesh_type="unstructured", return_counts=False, mean=None, normalizer=None,
                   trend=None, fit_normalizer=False, geo_scale=RADIAN_SCALE, **std_bins):
    """Empirical variogram (``variogram.py:151-470``); see the module notes for what is not restated."""
    if normalizer is not None or fit_normalizer:
        raise NotImplementedError("vario_estimate: normalizers are not part of this build")
    if bin_edges is not None:
        bin_edges = np.atleast_1d(np.asarray(bin_edges, dtype=np.double))
        bin_centers = (bin_edges[:-1] + bin_edges[1:]) / 2.0
    field = np.ma.array(field, ndmin=2, dtype=np.double, copy=True)
    masked = np.ma.is_masked(field) or np.any(mask)
    if masked and np.all(mask):
        bin_centers = np.empty(0) if bin_edges is None else bin_centers
        estimates = np.zeros_like(bin_centers)
        if return_counts:
            return bin_centers, estimates, np.zeros_like(estimates, dtype=int)
        return bin_centers, estimates
    if not masked:
        field = field.filled()
    if mesh_type != "unstructured":
        axes, dim = _struct_pos_shape(pos, field.shape)
        pos = generate_grid(axes)
    else:
        pos, dim = _unstruct_pos_shape(pos, field.shape)
    if latlon and dim != 2:
        raise ValueError("Variogram: given field needs to be 2D for lat-lon.")
    pnt_cnt = len(pos[0])
    field = field.reshape((-1, pnt_cnt))
    if masked:
        if np.size(mask) > 1:
            select = np.invert(np.logical_or(np.reshape(mask, pnt_cnt), np.all(field.mask, axis=0)))
        else:
            select = np.invert(np.all(field.mask, axis=0))
        pos = pos[:, select]
        field.fill_value = np.nan
        field = field[:, select].filled()
        pnt_cnt = len(pos[0])
    if not np.isnan(no_data):
        field[np.isclose(field, float(no_data))] = np.nan
    dir_no = 0
    if direction is not None and dim > 1:
        direction = np.atleast_2d(np.asarray(direction, dtype=np.double))
        if len(direction.shape) > 2 or direction.shape[1] != dim:
            raise ValueError(f"Can't interpret directions: {direction}")
        dir_no = direction.shape[0]
    if angles is not None and direction is None and dim > 1:
        direction = ang2dir(angles=angles, dtype=np.double, dim=dim)
        dir_no = direction.shape[0]
    if dir_no > 0:
        if latlon:
            raise ValueError("Directional variogram not allowed for lat-lon.")
        norms = np.linalg.norm(direction, axis=1)
        if np.any(np.isclose(norms, 0)):
            raise ValueError(f"Zero length directions: {direction}")
        direction = np.divide(direction, norms[:, np.newaxis])
        bandwidth = float(bandwidth) if bandwidth is not None else -1.0
        angles_tol = float(angles_tol)
    if sampling_size is not None and sampling_size < pnt_cnt:
        idx = np.random.RandomState(sampling_seed).choice(np.arange(pnt_cnt), sampling_size, replace=False)
        field = field[:, idx]
        pos = pos[:, idx]
    if bin_edges is None:
        bin_edges = standard_bins(pos, dim, latlon, geo_scale=geo_scale, **std_bins)
        bin_centers = (bin_edges[:-1] + bin_edges[1:]) / 2.0
    if latlon:
        # the kernel takes radians (out of place: a caller's edge array is left as it was)
        bin_edges = bin_edges / geo_scale
    # trend, then mean (a normalizer would act between them: variogram.py:433-440,
    # normalizer/tools.py:162-186)
    for i in range(field.shape[0]):
        field[i] -= eval_func(trend, pos, dim, broadcast=True)
        field[i] -= eval_func(mean, pos, dim, broadcast=True)
    code = "c" if estimator.lower() == "cressie" else "m"
    if estimator.lower() not in ("matheron", "cressie"):
        raise ValueError(f"Unknown variogram estimator function: {estimator}")
    if dir_no == 0:
        estimates, counts = unstructured(field, bin_edges, pos, estimator_type=code,
                                         distance_type="h" if latlon else "e")
    else:
        estimates, counts = directional(field, bin_edges, pos, direction, angles_tol, bandwidth,
                                        separate_dirs=_separate_dirs_test(direction, angles_tol),
                                        estimator_type=code)
        if dir_no == 1:
            estimates, counts = estimates[0], counts[0]
    est_out = (estimates, counts)
    return (bin_centers,) + est_out[: 2 if return_counts else 1]


def vario_estimate_axis(field, direction="x", estimator="matheron", no_data=np.nan):
    """Variogram along an array axis (``variogram.py:473-565``)."""
    missing_mask = np.isnan(field) if np.isnan(no_data) else np.isclose(field, no_data)
    missing = np.any(missing_mask)
    masked = np.ma.is_masked(field) or missing
    if masked:
        field = np.ma.array(field, ndmin=1, dtype=np.double)
        if missing:
            field.mask = np.logical_or(field.mask, missing_mask)
        mask = np.ma.getmaskarray(field)
    else:
        field = np.atleast_1d(np.asarray(field, dtype=np.double))
    axis = AXIS_DIR[direction] if direction in AXIS else int(direction)
    field = field.swapaxes(0, axis)
    field = field.reshape((field.shape[0], -1))
    code = "c" if estimator.lower() == "cressie" else "m"
    if masked:
        mask = mask.swapaxes(0, axis).reshape((mask.shape[0], -1))
        return ma_structured(np.ma.getdata(field), mask, code)
    return structured(field, code)


vario_estimate_unstructured = vario_estimate
vario_estimate_structured = vario_estimate_axis
