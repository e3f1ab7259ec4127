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


def directions_separated(direction, angles_tol):
    """Whether every two directions are at least 2 * angles_tol apart, so
    that a pair can match at most one of them (``variogram.py:568-593``)."""
    if direction is None or len(direction) < 2:
        return False
    cosines = np.minimum(np.abs(direction @ direction.T), 1.0)
    upper = np.triu_indices(len(direction), k=1)
    return bool(np.all(np.arccos(cosines[upper]) >= 2 * angles_tol))


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
def _field_and_points(field, mask, pos, mesh_type):
    """Stacked fields (F, n) with masked entries as NaN, and their (dim, n)
    positions; points masked by ``mask`` or in every field are dropped."""
    data = np.ma.array(field, ndmin=2, dtype=np.double, copy=True)
    mask_arr = np.asarray(mask, dtype=bool) if mask is not np.ma.nomask else np.zeros((), dtype=bool)
    if mesh_type == "unstructured":
        pts, dim = _unstruct_pos_shape(pos, data.shape)
    else:
        axes, dim = _struct_pos_shape(pos, data.shape)
        pts = generate_grid(axes)
    n = pts.shape[1]
    data = data.reshape((-1, n))
    if not (np.ma.is_masked(data) or mask_arr.any()):
        return np.ma.getdata(data).copy(), pts, dim
    drop = np.all(np.ma.getmaskarray(data), axis=0)
    if mask_arr.size > 1:
        drop |= mask_arr.reshape(n)
    values = np.ma.getdata(data)[:, ~drop].copy()
    values[np.ma.getmaskarray(data)[:, ~drop]] = np.nan
    return values, pts[:, ~drop], dim


def _unit_directions(direction, angles, dim):
    """(D, dim) unit direction vectors from vectors or angles, or None."""
    if dim < 2 or (direction is None and angles is None):
        return None
    if direction is None:
        vecs = ang2dir(angles=angles, dtype=np.double, dim=dim)
    else:
        vecs = np.atleast_2d(np.asarray(direction, dtype=np.double))
        if vecs.ndim != 2 or vecs.shape[1] != dim:
            raise ValueError(f"vario_estimate: directions must be vectors of length {dim}, got {direction!r}")
    lengths = np.linalg.norm(vecs, axis=1)
    if np.any(np.isclose(lengths, 0)):
        raise ValueError(f"vario_estimate: directions of zero length: {vecs}")
    return vecs / lengths[:, None]


def vario_estimate(pos, field, bin_edges=None, sampling_size=None, sampling_seed=None, estimator="matheron",
                   latlon=False, direction=None, angles=None, angles_tol=np.pi / 8, bandwidth=None, no_data=np.nan,
                   mask=np.ma.nomask, mesh_type="unstructured", return_counts=False, mean=None, normalizer=None,
                   trend=None, fit_normalizer=False, geo_scale=RADIAN_SCALE, **std_bins):
    """Empirical variogram of one or more stacked fields (``variogram.py:151-470``).

    Steps: drop masked points, mark masked / ``no_data`` values NaN (a pair
    with a NaN value counts in no bin), take directions from vectors or
    angles, subsample ``sampling_size`` points (``RandomState(sampling_seed)``,
    without replacement), choose standard bins if none are given, remove
    trend then mean, and bin all pairs on the GPU (Matheron or Cressie; for
    lat-lon points great-circle distances in units of ``geo_scale``).
    Returns bin centers, estimates and, with ``return_counts``, pair counts;
    with several directions the estimates have one row per direction.
    """
    if normalizer is not None or fit_normalizer:
        raise NotImplementedError("vario_estimate: normalizers are not part of this build")
    kind = str(estimator).lower()
    if kind not in ("matheron", "cressie"):
        raise ValueError(f"vario_estimate: estimator must be 'matheron' or 'cressie', got {estimator!r}")
    edges = None if bin_edges is None else np.atleast_1d(np.asarray(bin_edges, dtype=np.double))
    fully_masked = mask is not np.ma.nomask and np.size(mask) > 0 and np.all(mask)
    if fully_masked or (np.ma.is_masked(field) and np.all(np.ma.getmaskarray(field))):
        centers = np.empty(0) if edges is None else (edges[:-1] + edges[1:]) / 2.0
        zeros = np.zeros_like(centers)
        return (centers, zeros, np.zeros_like(zeros, dtype=int)) if return_counts else (centers, zeros)
    values, pts, dim = _field_and_points(field, mask, pos, mesh_type)
    if latlon and dim != 2:
        raise ValueError("vario_estimate: lat-lon fields need two coordinates (latitude, longitude)")
    if not np.isnan(no_data):
        values[np.isclose(values, float(no_data))] = np.nan
    dirs = _unit_directions(direction, angles, dim)
    if dirs is not None and latlon:
        raise ValueError("vario_estimate: directional variograms are not defined for lat-lon data")
    n = pts.shape[1]
    if sampling_size is not None and sampling_size < n:
        keep = np.random.RandomState(sampling_seed).choice(n, sampling_size, replace=False)
        values, pts = values[:, keep], pts[:, keep]
    if edges is None:
        edges = standard_bins(pts, dim, latlon, geo_scale=geo_scale, **std_bins)
    centers = (edges[:-1] + edges[1:]) / 2.0
    # trend first, then mean (a normalizer would act in between:
    # variogram.py:433-440, normalizer/tools.py:162-186)
    values -= eval_func(trend, pts, dim, broadcast=True)
    values -= eval_func(mean, pts, dim, broadcast=True)
    code = "c" if kind == "cressie" else "m"
    if dirs is None:
        # the kernel takes great-circle edges in radians (divided out of place)
        kernel_edges = edges / geo_scale if latlon else edges
        est, cnt = unstructured(values, kernel_edges, pts, estimator_type=code, distance_type="h" if latlon else "e")
    else:
        est, cnt = directional(values, edges, pts, dirs, float(angles_tol),
                               -1.0 if bandwidth is None else float(bandwidth),
                               separate_dirs=directions_separated(dirs, angles_tol), estimator_type=code)
        if len(dirs) == 1:
            est, cnt = est[0], cnt[0]
    return (centers, est, cnt) if return_counts else (centers, est)


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
