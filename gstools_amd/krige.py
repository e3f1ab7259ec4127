"""Kriging whose per-point sums run on the GPU.

Drop-ins for the external kriging kernels GSTools imports
(``src/gstools/krige/base.py:16-19``, dispatched at ``:42-61``):
``calc_field_krige(krig_mat, krig_vecs, cond, num_threads)`` and
``calc_field_krige_and_variance(...)`` take the same float64 numpy arrays and
return the same fresh arrays; the sums run in ``gsx_krige_f64`` (FP64 GEMM
with a fused epilogue on the matrix cores for the variance form, an HBM-bound
weighted column sum for the field alone).

:class:`Krige` restates the reference's kriging front end
(``krige/base.py:64-753``, methods ``krige/methods.py``): the kriging matrix
with measurement errors, unbiasedness and drift rows, its (pseudo-)inverse,
the right-hand sides per chunk of target points, constant or callable mean,
trend, ``exact`` / ``cond_err``, ``only_mean`` and ``get_mean``.  For the
Gaussian, Exponential, Stable and Spherical models the covariance rows of the
right-hand sides are computed on the GPU (``gsx_krige_cov_f64``); other
models are evaluated on the host exactly as the reference does.  The
subclasses :class:`Simple`, :class:`Ordinary`, :class:`Universal`,
:class:`ExtDrift` and :class:`Detrended` set the same defaults as the
reference.  Not restated: data normalizers and variogram fitting
(``normalizer`` other than ``None``, ``fit_normalizer`` or ``fit_variogram``
raise ``NotImplementedError``).
"""

from itertools import combinations_with_replacement

import numpy as np
import scipy.linalg as spl
import torch
from scipy.spatial.distance import cdist

from . import _native
from .geometry import generate_grid
from .summate import _as_f64, _device, _ptr, _stream, host_array

__all__ = [
    "calc_field_krige", "calc_field_krige_and_variance", "krige_sum_device", "krige_cov_device",
    "Krige", "Simple", "Ordinary", "Universal", "ExtDrift", "Detrended",
]

P_INV = {"pinv": spl.pinv, "pinvh": spl.pinvh}


# ------------------------------------------------------------------ sums ----
def krige_sum_device(krig_mat, krig_vecs, cond, variance=True):
    """Kriging sums on the GPU -> ``field`` or ``(field, error)`` float64 tensors (P,)."""
    dev = _device()
    m = _as_f64(krig_mat, "krig_mat", dev)
    v = _as_f64(krig_vecs, "krig_vecs", dev)
    c = _as_f64(cond, "cond", dev).reshape(-1)
    if v.dim() != 2:
        raise ValueError(f"krig_vecs: expected shape (K, P), got {tuple(v.shape)}")
    k, p = v.shape
    if tuple(m.shape) != (k, k) or c.numel() != k:
        raise ValueError(f"krig_mat must be ({k}, {k}) and cond ({k},), got {tuple(m.shape)} and {c.numel()}")
    field = torch.empty(p, dtype=torch.float64, device=dev)
    error = torch.empty(p, dtype=torch.float64, device=dev) if variance else None
    lib = _native.lib()
    ws_bytes = int(lib.gsx_krige_workspace(k, p, int(variance)))
    if ws_bytes < 0:
        _native.check(ws_bytes, "krige")
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    _native.check(lib.gsx_krige_f64(k, p, _ptr(m), _ptr(v), _ptr(c), _ptr(field), _ptr(error), _ptr(ws), ws_bytes,
                                    _stream()), "krige")
    return (field, error) if variance else field


def _check_numpy(**arrays):
    for name, arr in arrays.items():
        if isinstance(arr, np.ndarray) and arr.dtype != np.float64:
            raise ValueError(f"Buffer dtype mismatch for {name}: expected 'double' but got '{arr.dtype}'")


def calc_field_krige(krig_mat, krig_vecs, cond, num_threads=None):
    """Drop-in for the backend ``calc_field_krige`` -> field (P,) float64."""
    _check_numpy(krig_mat=krig_mat, krig_vecs=krig_vecs, cond=cond)
    return host_array(krige_sum_device(krig_mat, krig_vecs, cond, variance=False))


def calc_field_krige_and_variance(krig_mat, krig_vecs, cond, num_threads=None):
    """Drop-in for the backend ``calc_field_krige_and_variance`` -> (field, error)."""
    _check_numpy(krig_mat=krig_mat, krig_vecs=krig_vecs, cond=cond)
    field, error = krige_sum_device(krig_mat, krig_vecs, cond, variance=True)
    return host_array(field), host_array(error)


def krige_cov_device(model, cond_pos, pos, exact=False, out=None, row0=0):
    """Covariance rows C(|c_i - x_p|) on the GPU for closed-form models -> tensor (K, P), or None.

    ``cond_pos`` (dim, K) and ``pos`` (dim, P) are isometrized positions;
    with ``exact`` the value at distance 0 is the sill (``cov_nugget``).
    Returns None when the model has no device form (the caller evaluates
    it on the host as the reference does).  With ``out`` the rows are written
    to ``out[row0:row0 + K]``.
    """
    kind = _device_cov_kind(model)
    if kind is None:
        return None
    dev = _device()
    c = _as_f64(cond_pos, "cond_pos", dev)
    p = _as_f64(pos, "pos", dev)
    k, n = c.shape[1], p.shape[1]
    if k > 65535:
        return None
    if out is None:
        out = torch.empty((k, n), dtype=torch.float64, device=dev)
        row0 = 0
    dst = out[row0:row0 + k]
    alpha = float(getattr(model, "alpha", 0.0)) if kind == 2 else 0.0
    sill0 = float(model.sill) if exact else -1.0
    _native.check(_native.lib().gsx_krige_cov_f64(kind, float(model.var), float(model.len_rescaled), alpha, sill0,
                                                  c.shape[0], k, n, _ptr(c.contiguous()), _ptr(p.contiguous()),
                                                  _ptr(dst), out.shape[1], _stream()), "krige_cov")
    return dst


def _device_cov_kind(model):
    from . import covmodel as cm

    for cls, kind in ((cm.Gaussian, 0), (cm.Exponential, 1), (cm.Stable, 2), (cm.Spherical, 3)):
        if type(model) is cls:
            return kind
    return None


# ---------------------------------------------------------------- helpers ---
def set_condition(cond_pos, cond_val, dim):
    """Shape checks and removal of non-finite values (``krige/tools.py:16-47``)."""
    cond_val = np.asarray(cond_val, dtype=np.double).reshape(-1)
    cond_pos = np.asarray(cond_pos, dtype=np.double).reshape(dim, -1)
    if len(cond_pos[0]) != len(cond_val):
        raise ValueError("Please check your 'cond_pos' and 'cond_val' parameters. The shapes do not match.")
    mask = np.isfinite(cond_val)
    return cond_pos[:, mask], cond_val[mask]


def _monomial(select):
    def f(*pos):
        res = 1.0
        for i in select:
            res = res * np.asarray(pos[i])
        return res

    return f


def get_drift_functions(dim, drift_type):
    """Polynomial drift monomials up to the given order (``krige/tools.py:50-85``)."""
    if drift_type in ["lin", "linear"]:
        drift_type = 1
    elif drift_type in ["quad", "quadratic"]:
        drift_type = 2
    else:
        drift_type = int(drift_type)
    funcs = []
    for d in range(drift_type):
        for select in combinations_with_replacement(range(dim), d + 1):
            funcs.append(_monomial(select))
    return funcs


def eval_func(func_val, pos, dim, broadcast=False):
    """Mean / trend at positions (d, n) (``tools/misc.py:62-141``): a callable is evaluated,
    ``None`` is 0, any other value is a constant (its first element, as for a scalar field)."""
    pos = np.asarray(pos, dtype=np.double).reshape(dim, -1)
    if callable(func_val):
        return np.asarray(func_val(*pos), dtype=np.double).reshape(-1) * np.ones(pos.shape[1])
    val = 0.0 if func_val is None else float(np.asarray(func_val, dtype=np.double).ravel()[0])
    return np.full(pos.shape[1], val) if broadcast else val


def _positions(dim, pos, mesh_type):
    """(d, n) positions and the field shape of a position tuple."""
    if mesh_type == "structured":
        axes = [np.asarray(a, dtype=np.double).reshape(-1) for a in (pos if dim > 1 else np.atleast_2d(pos))]
        if len(axes) != dim:
            raise ValueError("position tuple does not match the model dimension")
        return generate_grid(axes), tuple(a.size for a in axes)
    if mesh_type != "unstructured":
        raise ValueError(f"unknown mesh_type {mesh_type!r}")
    arr = np.asarray(pos, dtype=np.double)
    if arr.dtype == object:
        raise ValueError("unstructured positions must have equal lengths")
    arr = arr.reshape(dim, -1)
    return arr, (arr.shape[1],)


# ----------------------------------------------------------------- Krige ----
class Krige:
    """Kriging front end (``krige/base.py:64-753``) with the sums on the GPU."""

    def __init__(self, model, cond_pos, cond_val, drift_functions=None, ext_drift=None, mean=None,
                 normalizer=None, trend=None, unbiased=True, exact=False, cond_err="nugget", pseudo_inv=True,
                 pseudo_inv_type="pinv", fit_normalizer=False, fit_variogram=False):
        if normalizer is not None or fit_normalizer:
            raise NotImplementedError("Krige: data normalizers are not part of this build")
        if fit_variogram:
            raise NotImplementedError("Krige: variogram fitting is not part of this build")
        if cond_pos is None or cond_val is None:
            raise ValueError("Krige.set_condition: missing cond_pos/cond_val.")
        self.model = model
        self.mean = mean
        self.trend = trend
        self._unbiased = bool(unbiased)
        self._exact = bool(exact)
        self._pseudo_inv = bool(pseudo_inv)
        self._pseudo_inv_type = None
        self.pseudo_inv_type = pseudo_inv_type
        self._cond_pos = None
        self._cond_val = None
        self._cond_err = None
        self._krige_mat = None
        self._krige_pos = None
        self._cond_ext_drift = np.array([])
        self._drift_functions = []
        self.fields = {}
        self.pos = None
        self.mesh_type = "unstructured"
        self.set_drift_functions(drift_functions)
        self.set_condition(cond_pos, cond_val, ext_drift, cond_err)

    @property
    def dim(self):
        return self.model.dim

    @property
    def name(self):
        return self.__class__.__name__

    # ----------------------------------------------------------- call --
    def __call__(self, pos=None, mesh_type="unstructured", ext_drift=None, chunk_size=None, only_mean=False,
                 return_var=True, post_process=True, store=True):
        """Kriged field (and variance) at ``pos`` (``krige/base.py:209-305``)."""
        return_var &= not only_mean
        if pos is None:
            if self.pos is None:
                raise ValueError("Krige: no positions given")
            pos, mesh_type = self.pos, self.mesh_type
        self.pos, self.mesh_type = pos, mesh_type
        xyz, shape = _positions(self.dim, pos, mesh_type)
        iso_pos = self.model.isometrize(xyz)
        pnt_cnt = iso_pos.shape[1]
        field = np.empty(pnt_cnt, dtype=np.double)
        krige_var = np.empty(pnt_cnt, dtype=np.double) if return_var else None
        if only_mean and self.drift_no == 0:
            field[...] = self.get_mean(post_process=False)
        else:
            chunk_size = pnt_cnt if chunk_size is None else int(chunk_size)
            chunk_no = int(np.ceil(pnt_cnt / chunk_size)) if pnt_cnt else 0
            ext_drift = self._pre_ext_drift(pnt_cnt, ext_drift)
            on_device = not only_mean and _device_cov_kind(self.model) is not None and self.cond_no <= 65535
            for i in range(chunk_no):
                sl = (i * chunk_size, min(pnt_cnt, (i + 1) * chunk_size))
                if on_device:
                    k_vec = self._get_krige_vecs_device(iso_pos, sl, ext_drift)
                else:
                    k_vec = self._get_krige_vecs(iso_pos, sl, ext_drift, only_mean)
                self._summate(field, krige_var, slice(*sl), k_vec, return_var)
        field = np.reshape(field, shape)
        if post_process:
            field = self._post(field, xyz, shape)
        if store:
            self.fields["mean_field" if only_mean else "field"] = field
        if return_var:
            krige_var = np.reshape(np.maximum(self.model.sill - krige_var, 0), shape)
            if store:
                self.fields["krige_var"] = krige_var
            return field, krige_var
        return field

    def _post(self, field, xyz, shape):
        """Add mean and trend like ``apply_mean_norm_trend`` without a normalizer."""
        mean = eval_func(self.mean, xyz, self.dim, broadcast=True).reshape(shape)
        trend = eval_func(self.trend, xyz, self.dim, broadcast=True).reshape(shape)
        return field + mean + trend

    def __getitem__(self, name):
        return self.fields[name]

    @property
    def field_names(self):
        return list(self.fields)

    def _summate(self, field, krige_var, c_slice, k_vec, return_var):
        if return_var:
            f, e = krige_sum_device(self._krige_mat, k_vec, self._krige_cond, variance=True)
            field[c_slice] = host_array(f)
            krige_var[c_slice] = host_array(e)
        else:
            field[c_slice] = host_array(krige_sum_device(self._krige_mat, k_vec, self._krige_cond, variance=False))

    # ------------------------------------------------------- system ----
    def _inv(self, mat):
        if self.pseudo_inv:
            if callable(self.pseudo_inv_type):
                return self.pseudo_inv_type(mat)
            return P_INV[self.pseudo_inv_type](mat)
        return spl.inv(mat)

    def _get_krige_mat(self):
        """Inverse of the kriging matrix (``krige/base.py:330-355``)."""
        res = np.empty((self.krige_size, self.krige_size), dtype=np.double)
        res[: self.cond_no, : self.cond_no] = self.model.covariance(cdist(self._krige_pos.T, self._krige_pos.T))
        res[np.diag_indices(self.cond_no)] += self.cond_err
        if self.unbiased:
            res[self.cond_no, : self.cond_no] = 1
            res[: self.cond_no, self.cond_no] = 1
        for i, f in enumerate(self.drift_functions):
            drift_tmp = f(*self.cond_pos)
            res[-self.drift_no + i, : self.cond_no] = drift_tmp
            res[: self.cond_no, -self.drift_no + i] = drift_tmp
        if self.ext_drift_no > 0:
            ext_size = self.krige_size - self.ext_drift_no
            res[ext_size:, : self.cond_no] = self.cond_ext_drift
            res[: self.cond_no, ext_size:] = self.cond_ext_drift.T
        res[self.cond_no:, self.cond_no:] = 0
        return self._inv(res)

    def _get_krige_vecs(self, pos, chunk_slice=(0, None), ext_drift=None, only_mean=False):
        """Right-hand sides of a chunk of target points (``krige/base.py:357-388``)."""
        chunk_size = (pos.shape[1] if chunk_slice[1] is None else chunk_slice[1]) - chunk_slice[0]
        res = np.empty((self.krige_size, chunk_size), dtype=np.double)
        if only_mean:
            res[: self.cond_no, :] = 0
        else:
            cf = self.model.cov_nugget if self.exact else self.model.covariance
            res[: self.cond_no, :] = cf(cdist(self._krige_pos.T, pos.T[slice(*chunk_slice), ...]))
        if self.unbiased:
            res[self.cond_no, :] = 1
        chunk_pos = None
        if self.int_drift_no > 0:
            chunk_pos = self.model.anisometrize(pos)[:, slice(*chunk_slice)]
        for i, f in enumerate(self.drift_functions):
            res[-self.drift_no + i, :] = f(*chunk_pos)
        if self.ext_drift_no > 0:
            ext_size = self.krige_size - self.ext_drift_no
            res[ext_size:, :] = ext_drift[:, slice(*chunk_slice)]
        return res

    def _get_krige_vecs_device(self, pos, chunk_slice, ext_drift=None):
        """``_get_krige_vecs`` with the covariance rows computed on the GPU (a device tensor)."""
        dev = _device()
        chunk_pos = np.ascontiguousarray(pos[:, slice(*chunk_slice)])
        res = torch.empty((self.krige_size, chunk_pos.shape[1]), dtype=torch.float64, device=dev)
        krige_cov_device(self.model, self._krige_pos, chunk_pos, exact=self.exact, out=res, row0=0)
        extra = self.krige_size - self.cond_no
        if extra:
            rows = np.empty((extra, chunk_pos.shape[1]), dtype=np.double)
            off = 0
            if self.unbiased:
                rows[0] = 1
                off = 1
            if self.int_drift_no > 0:
                drift_pos = self.model.anisometrize(pos)[:, slice(*chunk_slice)]
                for i, f in enumerate(self.drift_functions):
                    rows[off + i] = f(*drift_pos)
            if self.ext_drift_no > 0:
                rows[extra - self.ext_drift_no:] = ext_drift[:, slice(*chunk_slice)]
            res[self.cond_no:] = torch.from_numpy(rows).to(dev)
        return res

    def _pre_ext_drift(self, pnt_cnt, ext_drift=None, set_cond=False):
        """External drift values checked and shaped (``krige/base.py:390-428``)."""
        if ext_drift is not None:
            ext_drift = np.atleast_2d(np.asarray(ext_drift, dtype=np.double))
            if ext_drift.size == 0:
                return np.array([])
            if set_cond:
                if len(ext_drift.shape) > 2 or ext_drift.shape[1] != pnt_cnt:
                    raise ValueError("Krige: wrong number of ext. drifts.")
                return ext_drift
            ext_shape = np.shape(ext_drift)
            shape = (self.ext_drift_no, pnt_cnt)
            if self.drift_no > 1 and ext_shape[0] != self.ext_drift_no:
                raise ValueError("Krige: wrong number of external drifts.")
            if np.prod(ext_shape) != np.prod(shape):
                raise ValueError("Krige: wrong number of ext. drift values.")
            return np.asarray(ext_drift, dtype=np.double).reshape(shape)
        if not set_cond and self._cond_ext_drift.size > 0:
            raise ValueError("Krige: wrong number of ext. drift values.")
        return np.array([])

    def get_mean(self, post_process=True):
        """Estimated mean of the detrended field (``krige/base.py:452-487``)."""
        if not self.has_const_mean and (post_process or self.drift_no > 0):
            return None
        res = 0.0
        mean = 0.0 if self.mean is None else self.mean
        if self.unbiased:
            mean_est = np.concatenate((np.full_like(self.cond_val, 0.0), [1]))
            res = np.einsum("i,ij,j", self._krige_cond, self._krige_mat, mean_est)
        return res + mean if post_process else res

    def set_condition(self, cond_pos=None, cond_val=None, ext_drift=None, cond_err=None):
        """(Re)build the kriging system (``krige/base.py:489-586``, without fitting)."""
        ext_drift = self._cond_ext_drift if (ext_drift is None and cond_pos is None) else ext_drift
        cond_pos = self._cond_pos if cond_pos is None else cond_pos
        cond_val = self._cond_val if cond_val is None else cond_val
        cond_err = self._cond_err if cond_err is None else cond_err
        cond_err = "nugget" if cond_err is None else cond_err
        if cond_pos is None or cond_val is None:
            raise ValueError("Krige.set_condition: missing cond_pos/cond_val.")
        self._cond_pos, self._cond_val = set_condition(cond_pos, cond_val, self.dim)
        self.cond_err = cond_err
        self._cond_ext_drift = self._pre_ext_drift(self.cond_no, ext_drift, set_cond=True)
        self._krige_pos = self.model.isometrize(self.cond_pos)
        self._krige_mat = self._get_krige_mat()
        self.fields = {}

    def set_drift_functions(self, drift_functions=None):
        """Polynomial order, "linear"/"quadratic" or callables (``krige/base.py:588-624``)."""
        if drift_functions is None:
            self._drift_functions = []
        elif isinstance(drift_functions, (str, int)):
            self._drift_functions = get_drift_functions(self.dim, drift_functions)
        else:
            if callable(drift_functions):
                drift_functions = [drift_functions]
            drift_functions = list(drift_functions)
            for f in drift_functions:
                if not callable(f):
                    raise ValueError("Krige: Drift functions not callable")
            self._drift_functions = drift_functions

    # ----------------------------------------------------- properties ----
    @property
    def _krige_cond(self):
        pad_size = self.drift_no + int(self.unbiased)
        val = self.cond_val - self.cond_trend
        val = val - self.cond_mean
        return np.pad(val, (0, pad_size), mode="constant", constant_values=0)

    @property
    def cond_pos(self):
        return self._cond_pos

    @property
    def cond_val(self):
        return self._cond_val

    @property
    def cond_err(self):
        if isinstance(self._cond_err, str) and self._cond_err == "nugget":
            return self.model.nugget
        return self._cond_err

    @cond_err.setter
    def cond_err(self, value):
        if isinstance(value, str) and value == "nugget":
            self._cond_err = value
        else:
            if self.exact:
                raise ValueError("krige.cond_err: measurement errors can't be given, "
                                 "when interpolator should be exact.")
            value = np.asarray(value, dtype=np.double).reshape(-1)
            if value.size == 1:
                self._cond_err = value.item()
            else:
                if value.size != self.cond_no:
                    raise ValueError("krige.cond_err: wrong number of measurement errors.")
                self._cond_err = value

    @property
    def cond_no(self):
        return len(self._cond_val)

    @property
    def cond_ext_drift(self):
        return self._cond_ext_drift

    @property
    def cond_mean(self):
        return eval_func(self.mean, self.cond_pos, self.dim, broadcast=True)

    @property
    def cond_trend(self):
        return eval_func(self.trend, self.cond_pos, self.dim, broadcast=True)

    @property
    def unbiased(self):
        return self._unbiased

    @property
    def exact(self):
        return self._exact

    @property
    def pseudo_inv(self):
        return self._pseudo_inv

    @property
    def pseudo_inv_type(self):
        return self._pseudo_inv_type

    @pseudo_inv_type.setter
    def pseudo_inv_type(self, val):
        if val not in P_INV and not callable(val):
            raise ValueError(f"Krige: pseudo_inv_type not in {sorted(P_INV)}")
        self._pseudo_inv_type = val

    @property
    def drift_functions(self):
        return self._drift_functions

    @property
    def has_const_mean(self):
        return self.drift_no == 0 and not callable(self.mean)

    @property
    def krige_size(self):
        return self.cond_no + self.drift_no + int(self.unbiased)

    @property
    def drift_no(self):
        return self.int_drift_no + self.ext_drift_no

    @property
    def int_drift_no(self):
        return len(self.drift_functions)

    @property
    def ext_drift_no(self):
        return self.cond_ext_drift.shape[0] if self.cond_ext_drift.size else 0

    def structured(self, *args, **kwargs):
        return self(*args, mesh_type="structured", **kwargs)

    def unstructured(self, *args, **kwargs):
        return self(*args, mesh_type="unstructured", **kwargs)

    def __repr__(self):
        return f"{self.name}(model={self.model.name}, cond_no={self.cond_no})"


class Simple(Krige):
    """Simple kriging: known mean, no unbiasedness row (``krige/methods.py:21-119``)."""

    def __init__(self, model, cond_pos, cond_val, mean=0.0, normalizer=None, trend=None, exact=False,
                 cond_err="nugget", pseudo_inv=True, pseudo_inv_type="pinv", fit_normalizer=False,
                 fit_variogram=False):
        super().__init__(model, cond_pos, cond_val, mean=mean, normalizer=normalizer, trend=trend, unbiased=False,
                         exact=exact, cond_err=cond_err, pseudo_inv=pseudo_inv, pseudo_inv_type=pseudo_inv_type,
                         fit_normalizer=fit_normalizer, fit_variogram=fit_variogram)


class Ordinary(Krige):
    """Ordinary kriging: estimated constant mean (``krige/methods.py:122-214``)."""

    def __init__(self, model, cond_pos, cond_val, normalizer=None, trend=None, exact=False, cond_err="nugget",
                 pseudo_inv=True, pseudo_inv_type="pinv", fit_normalizer=False, fit_variogram=False):
        super().__init__(model, cond_pos, cond_val, trend=trend, normalizer=normalizer, exact=exact,
                         cond_err=cond_err, pseudo_inv=pseudo_inv, pseudo_inv_type=pseudo_inv_type,
                         fit_normalizer=fit_normalizer, fit_variogram=fit_variogram)


class Universal(Krige):
    """Universal kriging: drift functions (``krige/methods.py:217-324``)."""

    def __init__(self, model, cond_pos, cond_val, drift_functions, normalizer=None, trend=None, exact=False,
                 cond_err="nugget", pseudo_inv=True, pseudo_inv_type="pinv", fit_normalizer=False,
                 fit_variogram=False):
        super().__init__(model, cond_pos, cond_val, drift_functions=drift_functions, normalizer=normalizer,
                         trend=trend, exact=exact, cond_err=cond_err, pseudo_inv=pseudo_inv,
                         pseudo_inv_type=pseudo_inv_type, fit_normalizer=fit_normalizer,
                         fit_variogram=fit_variogram)


class ExtDrift(Krige):
    """External-drift kriging (``krige/methods.py:327-429``)."""

    def __init__(self, model, cond_pos, cond_val, ext_drift, normalizer=None, trend=None, exact=False,
                 cond_err="nugget", pseudo_inv=True, pseudo_inv_type="pinv", fit_normalizer=False,
                 fit_variogram=False):
        super().__init__(model, cond_pos, cond_val, ext_drift=ext_drift, normalizer=normalizer, trend=trend,
                         exact=exact, cond_err=cond_err, pseudo_inv=pseudo_inv, pseudo_inv_type=pseudo_inv_type,
                         fit_normalizer=fit_normalizer, fit_variogram=fit_variogram)


class Detrended(Krige):
    """Simple kriging of detrended data (``krige/methods.py:432-519``)."""

    def __init__(self, model, cond_pos, cond_val, trend, exact=False, cond_err="nugget", pseudo_inv=True,
                 pseudo_inv_type="pinv", fit_variogram=False):
        super().__init__(model, cond_pos, cond_val, trend=trend, unbiased=False, exact=exact, cond_err=cond_err,
                         pseudo_inv=pseudo_inv, pseudo_inv_type=pseudo_inv_type, fit_variogram=fit_variogram)
