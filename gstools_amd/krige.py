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
    "Krige", "KrigingSystem", "Simple", "Ordinary", "Universal", "ExtDrift", "Detrended",
]


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
def clean_conditions(cond_pos, cond_val, dim):
    """Conditioning data as (dim, K) positions and (K,) values, dropping
    points whose value is not finite (``krige/tools.py:16-47``)."""
    vals = np.asarray(cond_val, dtype=np.double).ravel()
    pts = np.asarray(cond_pos, dtype=np.double).reshape(dim, -1)
    if pts.shape[1] != vals.size:
        raise ValueError(f"Krige: {pts.shape[1]} conditioning positions but {vals.size} values")
    keep = np.isfinite(vals)
    return pts[:, keep], vals[keep]


def polynomial_drifts(dim, order):
    """Monomials x_a x_b ... of total degree 1..order in ``dim`` variables
    (``krige/tools.py:50-85``); ``order`` may be "linear" / "quadratic"."""
    named = {"lin": 1, "linear": 1, "quad": 2, "quadratic": 2}
    order = named[order] if order in named else int(order)
    out = []
    for degree in range(1, order + 1):
        for axes in combinations_with_replacement(range(dim), degree):
            out.append(_Monomial(axes))
    return out


class _Monomial:
    """Product of the coordinates listed in ``axes`` (repeats allowed)."""

    def __init__(self, axes):
        self.axes = tuple(axes)

    def __call__(self, *coords):
        val = 1.0
        for a in self.axes:
            val = val * np.asarray(coords[a], dtype=np.double)
        return val

    def __repr__(self):
        return "*".join(f"x{a}" for a in self.axes)


def eval_func(func_val, pos, dim, broadcast=False):
    """A mean or trend at positions (d, n) (``tools/misc.py:62-141``): callables are
    evaluated, ``None`` is 0, anything else a constant (its first element)."""
    pts = np.asarray(pos, dtype=np.double).reshape(dim, -1)
    n = pts.shape[1]
    if callable(func_val):
        return np.broadcast_to(np.asarray(func_val(*pts), dtype=np.double).ravel(), (n,)).copy()
    const = 0.0 if func_val is None else float(np.ravel(np.asarray(func_val, dtype=np.double))[0])
    return np.full(n, const) if broadcast else const


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


# -------------------------------------------------------- linear system ----
class KrigingSystem:
    """The kriging equations of one conditioning set.

    K conditioning points carry K weights; r = u + m + q constraints (u = 1
    for unbiasedness, m drift functions, q external drifts) carry Lagrange
    multipliers.  With C the covariance between the conditioning points,
    E the diagonal of measurement-error variances and F (K x r) the
    constraint values at the conditioning points,

        A = [[C + E, F],
             [F^T,   0]],       M = A^-1 (or a pseudo-inverse).

    A target x has the right-hand side b(x) = [c(x); f(x)]: its covariances
    to the conditioning points (with the sill at distance 0 for an exact
    interpolator) and its constraint values.  With the residual data
    y = [v - mean - trend; 0 (r entries)],

        estimate(x) = y . M b(x),     error(x) = b(x) . M b(x),

    and the kriging variance is max(sill - error, 0).  This is the system
    ``krige/base.py:330-388`` builds; the per-target sums are
    ``gsx_krige_f64``.
    """

    def __init__(self, model, iso_pos, constraints, errors, inverse):
        self.model = model
        self.iso_pos = iso_pos                  # (d, K) isometrized conditioning points
        self.constraints = constraints          # (K, r)
        k, r = constraints.shape
        a = np.zeros((k + r, k + r), dtype=np.double)
        a[:k, :k] = model.covariance(cdist(iso_pos.T, iso_pos.T))
        a[np.arange(k), np.arange(k)] += errors
        a[:k, k:] = constraints
        a[k:, :k] = constraints.T
        self.matrix = a
        self.inverse = inverse(a)

    @property
    def size(self):
        return self.matrix.shape[0]

    @property
    def cond_no(self):
        return self.constraints.shape[0]

    def rhs_host(self, iso, target_constraints, exact, covariance=True):
        """b(x) for the columns of ``iso`` (d, n) as a host array (K + r, n)."""
        k = self.cond_no
        out = np.empty((self.size, iso.shape[1]), dtype=np.double)
        if covariance:
            cov = self.model.cov_nugget if exact else self.model.covariance
            out[:k] = cov(cdist(self.iso_pos.T, iso.T))
        else:
            out[:k] = 0.0
        out[k:] = target_constraints
        return out

    def rhs_device(self, iso, target_constraints, exact):
        """b(x) on the GPU (covariance rows from ``gsx_krige_cov_f64``), or
        None when the model has no device covariance."""
        if _device_cov_kind(self.model) is None or self.cond_no > 65535:
            return None
        k = self.cond_no
        out = torch.empty((self.size, iso.shape[1]), dtype=torch.float64, device=_device())
        krige_cov_device(self.model, self.iso_pos, np.ascontiguousarray(iso), exact=exact, out=out, row0=0)
        if self.size > k:
            out[k:] = torch.from_numpy(np.ascontiguousarray(target_constraints)).to(out.device)
        return out

    def estimate(self, residual, rhs, variance):
        """``y . M b`` (and ``b . M b``) for every column of ``rhs``, on the GPU."""
        return krige_sum_device(self.inverse, rhs, residual, variance=variance)

    def unit_estimate(self, residual, row):
        """``y . M e_row``: the estimate for a right-hand side that is the unit
        vector of constraint ``row`` (the fitted constant of ordinary kriging)."""
        unit = np.zeros(self.size)
        unit[self.cond_no + row] = 1.0
        return float(residual @ (self.inverse @ unit))


_INVERSES = {"pinv": spl.pinv, "pinvh": spl.pinvh}


# ----------------------------------------------------------------- Krige ----
class Krige:
    """Kriging front end with the per-target sums on the GPU.

    Same constructor, call signature and properties as the reference's
    ``Krige`` (``krige/base.py:64-753``); the equations are documented on
    :class:`KrigingSystem`.  Conditioning points with non-finite values are
    dropped; polynomial / callable drifts are evaluated at the original
    (not isometrized) positions, external drifts are given per point.
    """

    def __init__(self, model, cond_pos, cond_val, drift_functions=None, ext_drift=None, mean=None,
                 normalizer=None, trend=None, unbiased=True, exact=False, cond_err="nugget", pseudo_inv=True,
                 pseudo_inv_type="pinv", fit_normalizer=False, fit_variogram=False):
        if normalizer is not None or fit_normalizer:
            raise NotImplementedError("Krige: data normalizers are not part of this build")
        if fit_variogram:
            raise NotImplementedError("Krige: variogram fitting is not part of this build")
        self.model = model
        self.mean = mean
        self.trend = trend
        self._unbiased = bool(unbiased)
        self._exact = bool(exact)
        self._pseudo_inv = bool(pseudo_inv)
        self.pseudo_inv_type = pseudo_inv_type
        self._drifts = []
        self._cond = None            # (positions, values) after cleaning
        self._err_spec = "nugget"    # "nugget", a scalar or one variance per point
        self._cond_ext = np.empty((0, 0))
        self._system = None
        self.fields = {}
        self.pos = None
        self.mesh_type = "unstructured"
        self.set_drift_functions(drift_functions)
        self.set_condition(cond_pos, cond_val, ext_drift, cond_err)

    # -------------------------------------------------------- setup --
    def set_drift_functions(self, drift_functions=None):
        """Internal drifts: None, a polynomial order ("linear", "quadratic",
        an int) or callables of the coordinates (``krige/base.py:588-624``)."""
        if drift_functions is None:
            drifts = []
        elif isinstance(drift_functions, (str, int, np.integer)):
            drifts = polynomial_drifts(self.dim, drift_functions)
        else:
            drifts = [drift_functions] if callable(drift_functions) else list(drift_functions)
            bad = [f for f in drifts if not callable(f)]
            if bad:
                raise ValueError(f"Krige: drift functions must be callable, got {bad[0]!r}")
        self._drifts = drifts
        if self._cond is not None:
            self._rebuild()

    def set_condition(self, cond_pos=None, cond_val=None, ext_drift=None, cond_err=None):
        """Replace (parts of) the conditioning data and rebuild the system
        (``krige/base.py:489-586``); omitted parts keep their values, and
        the external drift is kept only when the positions are."""
        if cond_pos is None and ext_drift is None:
            ext_drift = self._cond_ext if self._cond_ext.size else None
        pos_in = self._cond[0] if cond_pos is None and self._cond is not None else cond_pos
        val_in = self._cond[1] if cond_val is None and self._cond is not None else cond_val
        if pos_in is None or val_in is None:
            raise ValueError("Krige: conditioning positions and values are required")
        self._cond = clean_conditions(pos_in, val_in, self.dim)
        k = self._cond[1].size
        if ext_drift is None:
            self._cond_ext = np.empty((0, k))
        else:
            ext = np.atleast_2d(np.asarray(ext_drift, dtype=np.double))
            if ext.ndim != 2 or ext.shape[1] != k:
                raise ValueError(f"Krige: external drift at the conditioning points must be (q, {k}), "
                                 f"got shape {ext.shape}")
            self._cond_ext = ext
        self.cond_err = self._err_spec if cond_err is None else cond_err
        self._rebuild()

    def _rebuild(self):
        inv = spl.inv if not self._pseudo_inv else (
            self._pseudo_inv_type if callable(self._pseudo_inv_type) else _INVERSES[self._pseudo_inv_type])
        iso = self.model.isometrize(self.cond_pos)
        errors = np.broadcast_to(np.asarray(self.cond_err, dtype=np.double), (self.cond_no,))
        constraints = self._constraint_rows(self.cond_pos, self._cond_ext).T
        self._system = KrigingSystem(self.model, iso, np.ascontiguousarray(constraints), errors, inv)
        self.fields = {}

    def _constraint_rows(self, orig_pos, ext):
        """(r, n) constraint values: ones, drifts at ``orig_pos``, external drifts."""
        n = orig_pos.shape[1]
        rows = [np.ones(n)] if self._unbiased else []
        rows += [np.broadcast_to(np.asarray(f(*orig_pos), dtype=np.double).ravel(), (n,)) for f in self._drifts]
        rows = np.array(rows, dtype=np.double).reshape(len(rows), n)
        return np.concatenate([rows, ext], axis=0) if ext.size else rows

    def _target_ext_drift(self, n, ext_drift):
        """External drift at n targets as (q, n); required iff the conditioning has one."""
        q = self.ext_drift_no
        if ext_drift is None:
            if q:
                raise ValueError(f"Krige: {q} external drift(s) needed at the targets")
            return np.empty((0, n))
        ext = np.asarray(ext_drift, dtype=np.double)
        if ext.size == 0 and q == 0:
            return np.empty((0, n))
        if ext.size != q * n or (ext.ndim == 2 and q > 1 and ext.shape[0] != q):
            raise ValueError(f"Krige: expected {q} external drift value(s) for each of {n} targets, "
                             f"got shape {ext.shape}")
        return ext.reshape(q, n)

    # ---------------------------------------------------------- call --
    def __call__(self, pos=None, mesh_type="unstructured", ext_drift=None, chunk_size=None, only_mean=False,
                 return_var=True, post_process=True, store=True):
        """Kriged field (and kriging variance) at ``pos`` (``krige/base.py:209-305``)."""
        if pos is None:
            if self.pos is None:
                raise ValueError("Krige: no positions given")
            pos, mesh_type = self.pos, self.mesh_type
        self.pos, self.mesh_type = pos, mesh_type
        want_var = bool(return_var) and not only_mean
        orig, shape = _positions(self.dim, pos, mesh_type)
        iso = self.model.isometrize(orig)
        n = iso.shape[1]
        field = np.empty(n)
        error = np.empty(n) if want_var else None
        if only_mean and self.drift_no == 0:
            field[:] = self.get_mean(post_process=False)
        else:
            ext = self._target_ext_drift(n, ext_drift)
            step = n if chunk_size is None else max(int(chunk_size), 1)
            residual = self._residual()
            for lo in range(0, n, step):
                hi = min(n, lo + step)
                cons = self._constraint_rows(orig[:, lo:hi], ext[:, lo:hi])
                rhs = None if only_mean else self._system.rhs_device(iso[:, lo:hi], cons, self._exact)
                if rhs is None:
                    rhs = self._system.rhs_host(iso[:, lo:hi], cons, self._exact, covariance=not only_mean)
                res = self._system.estimate(residual, rhs, want_var)
                if want_var:
                    field[lo:hi], error[lo:hi] = host_array(res[0]), host_array(res[1])
                else:
                    field[lo:hi] = host_array(res)
        field = field.reshape(shape)
        if post_process:
            field = self._post(field, orig, shape)
        if store:
            self.fields["mean_field" if only_mean else "field"] = field
        if not want_var:
            return field
        variance = np.maximum(self.model.sill - error, 0.0).reshape(shape)
        if store:
            self.fields["krige_var"] = variance
        return field, variance

    def _post(self, field, orig, shape):
        """Add back mean and trend (``apply_mean_norm_trend`` without a normalizer)."""
        mean = eval_func(self.mean, orig, self.dim, broadcast=True).reshape(shape)
        trend = eval_func(self.trend, orig, self.dim, broadcast=True).reshape(shape)
        return field + mean + trend

    def _residual(self):
        """y = [v - trend - mean; 0 per constraint]."""
        y = np.zeros(self._system.size)
        y[: self.cond_no] = (self.cond_val - self.cond_trend) - self.cond_mean
        return y

    def get_mean(self, post_process=True):
        """Estimated constant mean (``krige/base.py:452-487``): the fitted
        constant of the unbiasedness constraint (0 without it), plus the
        given mean when ``post_process``; None when there is no constant
        mean (drifts, or a callable mean asked for with ``post_process``)."""
        if self.drift_no > 0 or (post_process and callable(self.mean)):
            return None
        fitted = self._system.unit_estimate(self._residual(), 0) if self._unbiased else 0.0
        if not post_process:
            return fitted
        return fitted + (0.0 if self.mean is None else self.mean)

    # -------------------------------------------------- properties --
    def __getitem__(self, name):
        return self.fields[name]

    @property
    def field_names(self):
        return list(self.fields)

    @property
    def dim(self):
        return self.model.dim

    @property
    def name(self):
        return type(self).__name__

    @property
    def system(self):
        """The :class:`KrigingSystem` of the current conditioning data."""
        return self._system

    @property
    def cond_pos(self):
        return self._cond[0]

    @property
    def cond_val(self):
        return self._cond[1]

    @property
    def cond_no(self):
        return self._cond[1].size

    @property
    def cond_err(self):
        """Measurement-error variances: the model's nugget unless given."""
        return self.model.nugget if isinstance(self._err_spec, str) else self._err_spec

    @cond_err.setter
    def cond_err(self, value):
        if isinstance(value, str):
            if value != "nugget":
                raise ValueError(f"Krige: cond_err must be 'nugget' or variances, got {value!r}")
            self._err_spec = value
            return
        if self._exact:
            raise ValueError("Krige: an exact interpolator takes no measurement errors (cond_err)")
        arr = np.asarray(value, dtype=np.double).ravel()
        if arr.size == 1:
            self._err_spec = float(arr[0])
        elif arr.size == self.cond_no:
            self._err_spec = arr
        else:
            raise ValueError(f"Krige: cond_err needs 1 or {self.cond_no} values, got {arr.size}")

    @property
    def cond_ext_drift(self):
        return self._cond_ext

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
        if not callable(val) and val not in _INVERSES:
            raise ValueError(f"Krige: pseudo_inv_type must be callable or one of {sorted(_INVERSES)}, got {val!r}")
        self._pseudo_inv_type = val

    @property
    def drift_functions(self):
        return self._drifts

    @property
    def has_const_mean(self):
        return self.drift_no == 0 and not callable(self.mean)

    @property
    def krige_size(self):
        return self.cond_no + self.drift_no + int(self._unbiased)

    @property
    def int_drift_no(self):
        return len(self._drifts)

    @property
    def ext_drift_no(self):
        return self._cond_ext.shape[0]

    @property
    def drift_no(self):
        return self.int_drift_no + self.ext_drift_no

    def structured(self, *args, **kwargs):
        return self(*args, mesh_type="structured", **kwargs)

    def unstructured(self, *args, **kwargs):
        return self(*args, mesh_type="unstructured", **kwargs)

    def __repr__(self):
        return f"{self.name}(model={self.model.name}, cond_no={self.cond_no})"


def _method(name, doc, fixed):
    """A kriging method: ``Krige`` with some constructor arguments fixed."""

    def __init__(self, model, cond_pos, cond_val, *args, **kwargs):
        names = _METHOD_ARGS[name]
        if len(args) > len(names):
            raise TypeError(f"{name}: at most {3 + len(names)} positional arguments")
        kw = dict(fixed)
        kw.update(zip(names, args))
        for key, val in kwargs.items():
            if key in fixed or key not in names:
                raise TypeError(f"{name}: unexpected argument '{key}'")
            if key in kw:
                raise TypeError(f"{name}: '{key}' given twice")
            kw[key] = val
        need = _METHOD_REQUIRED.get(name)
        if need and need not in kw:
            raise TypeError(f"{name}: missing argument '{need}'")
        Krige.__init__(self, model, cond_pos, cond_val, **kw)

    return type(name, (Krige,), {"__init__": __init__, "__doc__": doc})


# positional arguments after (model, cond_pos, cond_val), as in krige/methods.py
_METHOD_ARGS = {
    "Simple": ("mean", "normalizer", "trend", "exact", "cond_err", "pseudo_inv", "pseudo_inv_type",
               "fit_normalizer", "fit_variogram"),
    "Ordinary": ("normalizer", "trend", "exact", "cond_err", "pseudo_inv", "pseudo_inv_type", "fit_normalizer",
                 "fit_variogram"),
    "Universal": ("drift_functions", "normalizer", "trend", "exact", "cond_err", "pseudo_inv", "pseudo_inv_type",
                  "fit_normalizer", "fit_variogram"),
    "ExtDrift": ("ext_drift", "normalizer", "trend", "exact", "cond_err", "pseudo_inv", "pseudo_inv_type",
                 "fit_normalizer", "fit_variogram"),
    "Detrended": ("trend", "exact", "cond_err", "pseudo_inv", "pseudo_inv_type", "fit_variogram"),
}

_METHOD_REQUIRED = {"Universal": "drift_functions", "ExtDrift": "ext_drift", "Detrended": "trend"}

Simple = _method("Simple", "Simple kriging: known mean, no unbiasedness constraint (``krige/methods.py:21-119``).",
                 {"unbiased": False})
Ordinary = _method("Ordinary", "Ordinary kriging: unknown constant mean (``krige/methods.py:122-214``).", {})
Universal = _method("Universal", "Universal kriging: drift functions (``krige/methods.py:217-324``).", {})
ExtDrift = _method("ExtDrift", "Kriging with external drifts (``krige/methods.py:327-429``).", {})
Detrended = _method("Detrended", "Simple kriging of detrended data (``krige/methods.py:432-519``).",
                    {"unbiased": False})
