"""Randomisation-method generators whose summation runs on the GPU.

:class:`RandMeth`, :class:`IncomprRandMeth` and :class:`Fourier` mirror the
reference classes (``src/gstools/field/generator.py:165-446``, ``:449-592``
and ``:595-919``): same
constructor keywords, same ``update`` / ``reset_seed`` / ``seed`` /
``mode_no`` / ``sampling`` / ``model`` semantics, same private mode arrays
``_cov_sample`` / ``_z_1`` / ``_z_2`` drawn from the same ``RNG`` streams,
and the same scaling ``sqrt(var / mode_no) * summed + nugget`` (Fourier: the
spectrum factors, no global scale).  The sum itself is the gfx950 kernel
(``gstools_amd.summate``).

Plugging into the reference: when ``gstools`` is importable the classes
derive from its ``Generator`` ABC, so ``gstools.SRF(model,
generator=RandMeth, ...)`` accepts them unchanged (the seam at
``src/gstools/field/srf.py:169-186``).  Models may be the reference's
``CovModel`` objects or the ones in :mod:`gstools_amd.covmodel`.

Extra keyword (not in the reference): ``precision`` -- ``"fp64"`` (default,
matches the reference CPU sum to 1e-12) or ``"fp32"``.
"""

import warnings
from abc import ABC, abstractmethod
from copy import deepcopy as dcp

import numpy as np
import torch

from .rng import RNG
from .summate import (
    GRID_PRECISIONS,
    INCOMPR_GRID_PRECISIONS,
    PRECISIONS,
    host_array,
    summate_device,
    summate_fourier_tensor_device,
    summate_grid_device,
    summate_incompr_grid_device,
    summate_incompr_device,
)

try:  # register with the reference's plugin seam when it is installed
    from gstools.field.generator import Generator as _RefGenerator  # type: ignore
except Exception:  # pragma: no cover - the reference is not importable here
    _RefGenerator = None

__all__ = ["Generator", "RandMeth", "IncomprRandMeth", "Fourier", "SAMPLING"]

SAMPLING = ["auto", "inversion", "mcmc"]


def _is_covmodel(model):
    if model is None:
        return False
    try:
        from gstools.covmodel.base import CovModel as _RefCov  # type: ignore

        if isinstance(model, _RefCov):
            return True
    except Exception:
        pass
    from .covmodel import CovModel

    return isinstance(model, CovModel)


def _cached_upload(gen, arrays):
    """Device copies of a generator's mode arrays, uploaded again whenever an
    array was replaced or written in place (the reference reads them on
    every call).  The check compares against a host snapshot of what was
    uploaded: O(N) per call, against the O(N P) sum."""
    dev = torch.device("cuda", torch.cuda.current_device())
    cached = gen._dev_modes
    if (cached is None or cached[0] != dev or len(cached[1]) != len(arrays)
            or not all(np.array_equal(a, s, equal_nan=True) for a, s in zip(arrays, cached[1]))):
        snap = tuple(np.array(a, dtype=np.float64, copy=True, order="C") for a in arrays)
        gen._dev_modes = (dev, snap, tuple(torch.from_numpy(a).to(dev) for a in snap))
    return gen._dev_modes[2]


class _GeneratorBase(ABC):
    """Abstract generator (``generator.py:78-162``)."""

    @abstractmethod
    def __init__(self, model, **kwargs):
        pass

    @abstractmethod
    def update(self, model=None, seed=np.nan):
        """Update the model and the seed."""

    @abstractmethod
    def get_nugget(self, shape):
        """Normal nugget noise of the given shape."""

    @abstractmethod
    def __call__(self, pos, add_nugget=True):
        """Generate the field."""

    @property
    @abstractmethod
    def model(self):
        """Covariance model."""

    @property
    def zero_var(self):
        return np.isclose(self.model.var, 0)

    @property
    @abstractmethod
    def value_type(self):
        """Type of the field values (scalar, vector)."""

    @property
    def name(self):
        return self.__class__.__name__


Generator = _RefGenerator if _RefGenerator is not None else _GeneratorBase


class RandMeth(Generator):
    """Randomisation method for isotropic random fields (``generator.py:165-446``)."""

    def __init__(self, model, *, mode_no=1000, seed=None, sampling="auto", precision="fp64", **kwargs):
        if kwargs:
            warnings.warn(f"{type(self).__name__}: unused keyword arguments {sorted(kwargs)}")
        self._mode_no = int(mode_no)
        self._model = None
        self._seed = None
        self._rng = None
        self._z_1 = None
        self._z_2 = None
        self._cov_sample = None
        self._value_type = "scalar"
        self._sampling = None
        self._dev_modes = None
        self.sampling = sampling
        self.precision = precision
        self.update(model, seed)

    # --------------------------------------------------------------- call --
    def _device_modes(self):
        return _cached_upload(self, (self._cov_sample, self._z_1, self._z_2))

    def _folded_modes(self, iso_matrix):
        """Device modes with the isometrize matrix folded in: k' = M^T k, so that
        k'.x = k.(M x) and a rotated or anisotropic grid stays a tensor grid."""
        cov, z1, z2 = self._device_modes()
        mat = np.asarray(iso_matrix, dtype=np.float64)
        if mat.shape != (cov.shape[0],) * 2:
            raise ValueError(f"iso_matrix must be {cov.shape[0]} x {cov.shape[0]}")
        return torch.as_tensor(mat.T.copy(), device=cov.device) @ cov, z1, z2

    def _scale(self):
        return np.sqrt(self.model.var / self._mode_no)

    def __call__(self, pos, add_nugget=True):
        """Field at ``pos`` (d, n): ``sqrt(var / N) * summate(...) + nugget``.

        Mirrors ``generator.py:243-270``.  numpy input gives a float64 numpy
        result; a torch CUDA tensor gives a tensor on the device.
        """
        as_tensor = isinstance(pos, torch.Tensor)
        if not as_tensor:
            pos = np.asarray(pos, dtype=np.double)
        if self.zero_var:
            shp = tuple(pos.shape[1:])
            return self.get_nugget(shp) if add_nugget else np.full(shp, 0.0)
        cov, z1, z2 = self._device_modes()
        summed = summate_device(cov, z1, z2, pos, precision=self.precision)
        field = summed.to(torch.float64) * float(self._scale())
        if not as_tensor:
            field = host_array(field)
        nugget = self.get_nugget(tuple(field.shape)) if add_nugget else 0.0
        if as_tensor and not np.isscalar(nugget):
            nugget = torch.as_tensor(nugget, device=field.device)
        return field + nugget

    def structured(self, axes, precision=None, add_nugget=True, iso_matrix=None):
        """Field on the 'ij' meshgrid of ``axes`` without materialising it.

        ``precision`` may additionally be ``"bf16x3"`` or ``"fp16x3"`` (matrix
        cores, d = 2 or 3) or ``"fp64gemm"`` (FP64 GEMM on the matrix cores; it
        skips the rounding of the phase sum, about |k.x| 2^-52 per term).
        ``iso_matrix`` (d, d): the field at the isometrized positions
        ``iso_matrix @ x`` (rotation / anisotropy, ``covmodel/base.py:572-582``),
        evaluated as k'.x with k' = iso_matrix^T k; the phases then differ from
        the reference's k.(M x) in rounding only (about |k| |x| 2^-52).  Returns
        a torch tensor shaped like the grid.
        """
        precision = precision or self.precision
        if precision not in GRID_PRECISIONS:
            raise ValueError(f"precision must be one of {GRID_PRECISIONS}")
        shape = tuple(int(np.size(a)) for a in axes)
        if self.zero_var:
            out = torch.zeros(shape, dtype=torch.float64, device="cuda")
        else:
            cov, z1, z2 = self._device_modes() if iso_matrix is None else self._folded_modes(iso_matrix)
            if precision in ("bf16x3", "fp16x3"):
                out = summate_grid_device(cov, z1, z2, axes, precision, alpha=float(self._scale()))
            else:
                out = summate_grid_device(cov, z1, z2, axes, precision).to(torch.float64) * float(self._scale())
            out = out.reshape(shape)
        if add_nugget and self.model.nugget > 0:
            out = out + torch.as_tensor(self.get_nugget(shape), device=out.device, dtype=out.dtype)
        return out

    def get_nugget(self, shape):
        """Nugget noise (``generator.py:272-292``)."""
        if self.model.nugget > 0:
            nugget = np.sqrt(self.model.nugget) * self._rng.random.normal(size=shape)
        else:
            nugget = np.full(shape, 0.0) if self.zero_var else 0.0
        return nugget

    # -------------------------------------------------------------- modes --
    def update(self, model=None, seed=np.nan):
        """Update model and seed (``generator.py:294-344``)."""
        if _is_covmodel(model):
            if self.model != model:
                self._model = dcp(model)
                if seed is None or not np.isnan(seed):
                    self.reset_seed(seed)
                else:
                    self.reset_seed(self._seed)
            elif seed is None or not np.isnan(seed):
                self.seed = seed
        elif model is None and (seed is None or not np.isnan(seed)):
            if _is_covmodel(self._model):
                self.seed = seed
            else:
                raise ValueError(f"{self.name}.update: a seed was given but the generator has no model yet")
        elif model is None and np.isnan(seed):
            if not (_is_covmodel(self._model) and self._z_1 is not None
                    and self._z_2 is not None and self._cov_sample is not None):
                raise ValueError(f"{self.name}.update: needs a model or a seed (no modes to keep)")
        else:
            raise ValueError(f"{self.name}.update: model must be a CovModel, got {type(model).__name__}")

    def reset_seed(self, seed=np.nan):
        """Redraw z_1, z_2 and the wavenumbers (``generator.py:346-387``)."""
        if seed is None or not np.isnan(seed):
            self._seed = seed
        self._rng = RNG(self._seed)
        self._z_1 = self._rng.random.normal(size=self._mode_no)
        self._z_2 = self._rng.random.normal(size=self._mode_no)
        if self.zero_var:
            self._cov_sample = np.full((self.model.dim, self._mode_no), 0.0)
        else:
            sph_crd = self._rng.sample_sphere(self.model.dim, self._mode_no)
            if self.sampling == "inversion" or (self.sampling == "auto" and self.model.has_ppf):
                pdf, cdf, ppf = self.model.dist_func
                rad = self._rng.sample_dist(size=self._mode_no, pdf=pdf, cdf=cdf, ppf=ppf, a=0)
            else:
                rad = self._rng.sample_ln_pdf(
                    ln_pdf=self.model.ln_spectral_rad_pdf,
                    size=self._mode_no,
                    sample_around=1.0 / self.model.len_rescaled,
                )
            self._cov_sample = rad * sph_crd
        self._dev_modes = None

    @property
    def sampling(self):
        return self._sampling

    @sampling.setter
    def sampling(self, sampling):
        if sampling not in SAMPLING:
            raise ValueError(f"RandMeth: sampling must be one of {SAMPLING}, got {sampling!r}")
        self._sampling = sampling

    @property
    def precision(self):
        return self._precision

    @precision.setter
    def precision(self, precision):
        if precision not in PRECISIONS:
            raise ValueError(f"RandMeth: precision not in {PRECISIONS}.")
        self._precision = precision

    @property
    def seed(self):
        return self._seed

    @seed.setter
    def seed(self, new_seed):
        if new_seed is not self._seed:
            self.reset_seed(new_seed)

    @property
    def model(self):
        return self._model

    @model.setter
    def model(self, model):
        self.update(model)

    @property
    def mode_no(self):
        return self._mode_no

    @mode_no.setter
    def mode_no(self, mode_no):
        if int(mode_no) != self._mode_no:
            self._mode_no = int(mode_no)
            self.reset_seed(self._seed)

    @property
    def value_type(self):
        return self._value_type

    def __repr__(self):
        return f"{self.name}(model={self.model}, mode_no={self._mode_no}, seed={self.seed})"


class IncomprRandMeth(RandMeth):
    """Incompressible vector fields (``generator.py:449-592``)."""

    def __init__(self, model, *, mean_velocity=1.0, mode_no=1000, seed=None, sampling="auto",
                 precision="fp64", **kwargs):
        if model.dim < 2 or model.dim > 3:
            raise ValueError(f"IncomprRandMeth: incompressible fields need dim 2 or 3, got {model.dim}")
        super().__init__(model=model, mode_no=mode_no, seed=seed, sampling=sampling,
                         precision=precision, **kwargs)
        self.mean_u = mean_velocity
        self._value_type = "vector"

    def __call__(self, pos, add_nugget=True):
        """``mean_u e1 + mean_u sqrt(var / N) summate_incompr + nugget`` (``:529-567``)."""
        as_tensor = isinstance(pos, torch.Tensor)
        if not as_tensor:
            pos = np.asarray(pos, dtype=np.double)
        nugget = self.get_nugget(tuple(pos.shape)) if add_nugget else 0.0
        e1 = self._create_unit_vector(tuple(pos.shape))
        if self.zero_var:
            return self.mean_u * e1 + nugget
        cov, z1, z2 = self._device_modes()
        summed = summate_incompr_device(cov, z1, z2, pos, precision=self.precision).to(torch.float64)
        scale = self.mean_u * np.sqrt(self.model.var / self._mode_no)
        if as_tensor:
            e1t = torch.as_tensor(e1, device=summed.device)
            out = self.mean_u * e1t + scale * summed
            if not np.isscalar(nugget):
                nugget = torch.as_tensor(nugget, device=summed.device)
            return out + nugget
        return self.mean_u * e1 + scale * host_array(summed) + nugget

    def structured(self, axes, precision=None, add_nugget=True):
        """Vector field (d, *shape) on the 'ij' meshgrid of ``axes`` without materialising it.

        ``summate_incompr`` (``generator.py:479-494``) on the grid: the direct
        sums ("fp64" / "fp32") or, on the matrix cores ("bf16x3" / "fp16x3"),
        one GEMM per component with the projector weight
        ``p_c(k_j) = delta_c0 - k_jc k_j0 / |k_j|^2`` per mode; then
        ``mean_u e1 + mean_u sqrt(var / N) * sum`` as in ``__call__``.  A k = 0
        mode makes every value NaN, as the reference's 0 / 0 does.
        """
        precision = precision or self.precision
        if precision not in INCOMPR_GRID_PRECISIONS:
            raise ValueError(f"precision must be one of {INCOMPR_GRID_PRECISIONS}")
        d = self.model.dim
        shape = tuple(int(np.size(a)) for a in axes)
        if len(shape) != d:
            raise ValueError(f"expected {d} axes, got {len(shape)}")
        mfma = precision in ("bf16x3", "fp16x3")
        dtype = torch.float32 if mfma else torch.float64
        if self.zero_var:
            out = torch.zeros((d,) + shape, dtype=dtype, device="cuda")
        else:
            cov, z1, z2 = self._device_modes()
            scale = float(self.mean_u * np.sqrt(self.model.var / self._mode_no))
            if mfma:
                out = summate_incompr_grid_device(cov, z1, z2, axes, precision, alpha=scale).reshape((d,) + shape)
            else:
                out = summate_incompr_grid_device(cov, z1, z2, axes, precision).to(torch.float64)
                out = (out * scale).reshape((d,) + shape)
        out[0] += self.mean_u
        if add_nugget and self.model.nugget > 0:
            out = out + torch.as_tensor(self.get_nugget((d,) + shape), device=out.device, dtype=out.dtype)
        return out

    def _create_unit_vector(self, broadcast_shape, axis=0):
        shape = np.ones(len(broadcast_shape), dtype=int)
        shape[0] = self.model.dim
        e1 = np.zeros(shape)
        e1[axis] = 1.0
        return e1


class Fourier(Generator):
    """Fourier method for periodic fields (``generator.py:595-919``).

    ``u(x) = sum_i sqrt(S(|k_i|) prod(dk)) (z1_i cos(k_i.x) + z2_i sin(k_i.x))``
    over the regular wavenumber grid ``k`` of ``mode_no`` points per axis with
    spacing ``dk = 2 pi / period * [1, anis]``; ``S`` is the model's
    spectrum.  Only even ``mode_no`` are accepted, as in the reference.
    """

    def __init__(self, model, period, mode_no=32, seed=None, precision="fp64", **kwargs):
        if kwargs:
            warnings.warn(f"Fourier: unused keyword arguments {sorted(kwargs)}")
        self._modes = None
        self._period = None
        self._mode_no = None
        self._delta_k = None
        self._model = None
        self._seed = None
        self._rng = None
        self._z_1 = None
        self._z_2 = None
        self._spectrum_factor = None
        self._value_type = "scalar"
        self._dev_modes = None
        self.precision = precision
        self.update(model, seed, period, mode_no)

    # --------------------------------------------------------------- call --
    def _device_modes(self):
        return _cached_upload(self, (self._spectrum_factor, self._modes, self._z_1, self._z_2))

    def __call__(self, pos, add_nugget=True):
        """Field at ``pos`` (d, n): ``summate_fourier(...) + nugget`` (``generator.py:659-691``).

        numpy input gives a float64 numpy result; a torch CUDA tensor gives a
        tensor on the device.
        """
        as_tensor = isinstance(pos, torch.Tensor)
        if not as_tensor:
            pos = np.asarray(pos, dtype=np.double)
        if self.zero_var:
            shp = tuple(pos.shape[1:])
            return self.get_nugget(shp) if add_nugget else np.full(shp, 0.0)
        fac, modes, z1, z2 = self._device_modes()
        field = summate_device(modes, z1, z2, pos, precision=self.precision, spectrum_factor=fac)
        field = field.to(torch.float64)
        if not as_tensor:
            field = host_array(field)
        nugget = self.get_nugget(tuple(field.shape)) if add_nugget else 0.0
        if as_tensor and not np.isscalar(nugget):
            nugget = torch.as_tensor(nugget, device=field.device)
        return field + nugget

    def structured(self, axes, precision=None, add_nugget=True, method="tensor", iso_matrix=None):
        """Field on the 'ij' meshgrid of ``axes`` without materialising it.

        ``method="tensor"`` (default): the modes are a tensor grid, so the sum
        factorises into one FP64 GEMM contraction per axis on the matrix cores
        (``summate_fourier_tensor_device``); FP64 whatever ``precision`` says,
        which only picks the output dtype (float64 for "fp64" / "fp32", float32
        for "bf16x3" / "fp16x3").  ``method="direct"``: the per-point sum as in
        :meth:`RandMeth.structured` (``"bf16x3"`` / ``"fp16x3"`` on the matrix
        cores).  ``iso_matrix`` (d, d): the field at ``iso_matrix @ x``,
        evaluated with the modes k' = iso_matrix^T k; a diagonal matrix
        (anisotropy only) keeps the modes a tensor grid, a rotation turns
        ``method="tensor"`` into the direct sum.  Returns a torch tensor shaped
        like the grid.
        """
        precision = precision or self.precision
        if precision not in GRID_PRECISIONS:
            raise ValueError(f"precision must be one of {GRID_PRECISIONS}")
        if method not in ("tensor", "direct"):
            raise ValueError("method must be 'tensor' or 'direct'")
        shape = tuple(int(np.size(a)) for a in axes)
        mat = None if iso_matrix is None else np.asarray(iso_matrix, dtype=np.float64)
        if mat is not None and mat.shape != (len(shape),) * 2:
            raise ValueError(f"iso_matrix must be {len(shape)} x {len(shape)}")
        diagonal = mat is None or np.array_equal(mat, np.diag(np.diag(mat)))
        if self.zero_var:
            out = torch.zeros(shape, dtype=torch.float64, device="cuda")
        elif method == "tensor" and diagonal:
            fac, _, z1, z2 = self._device_modes()
            kaxes = self._kaxes if mat is None else [k * mat[i, i] for i, k in enumerate(self._kaxes)]
            out = summate_fourier_tensor_device(fac, kaxes, z1, z2, axes)
            if precision in ("bf16x3", "fp16x3"):
                out = out.to(torch.float32)
            out = out.reshape(shape)
        else:
            fac, modes, z1, z2 = self._device_modes()
            if mat is not None:
                modes = torch.as_tensor(mat.T.copy(), device=modes.device) @ modes
            out = summate_grid_device(modes, z1, z2, axes, precision, spectrum_factor=fac)
            if precision in ("fp64", "fp32"):
                out = out.to(torch.float64)
            out = out.reshape(shape)
        if add_nugget and self.model.nugget > 0:
            out = out + torch.as_tensor(self.get_nugget(shape), device=out.device, dtype=out.dtype)
        return out

    def get_nugget(self, shape):
        """Nugget noise (``generator.py:693-713``)."""
        if self.model.nugget > 0:
            nugget = np.sqrt(self.model.nugget) * self._rng.random.normal(size=shape)
        else:
            nugget = np.full(shape, 0.0) if self.zero_var else 0.0
        return nugget

    # -------------------------------------------------------------- modes --
    def update(self, model=None, seed=np.nan, period=None, mode_no=None):
        """Update model, seed, period and mode numbers (``generator.py:715-789``)."""
        tmp_model = model if model is not None else self._model
        dim = tmp_model.dim
        if period is not None:
            self._period = self._fill_to_dim(period, dim)
            anis = np.insert(np.atleast_1d(tmp_model.anis).copy(), 0, 1.0)
            self._delta_k = 2.0 * np.pi / self._period * anis
            if mode_no is None:
                self._set_modes(self._mode_no, dim)
        if mode_no is not None:
            mode_no = self._fill_to_dim(mode_no, dim)
            if (np.asarray([m % 2 for m in mode_no]) != 0).any():
                raise ValueError(f"Fourier: mode_no must be even in every dimension, got {list(mode_no)}")
            self._set_modes(mode_no, dim)

        if _is_covmodel(model):
            if self.model != model:
                self._model = dcp(model)
                if seed is None or not np.isnan(seed):
                    self.reset_seed(seed)
                else:
                    self.reset_seed(self._seed)
            elif seed is None or not np.isnan(seed):
                self.seed = seed
        elif model is None and (seed is None or not np.isnan(seed)):
            if _is_covmodel(self._model):
                self.seed = seed
            else:
                raise ValueError(f"{self.name}.update: a seed was given but the generator has no model yet")
        elif mode_no is not None or period is not None:
            if seed is None or not np.isnan(seed):
                self.reset_seed(seed)
            else:
                self.reset_seed(self._seed)
        elif model is None and np.isnan(seed):
            if (_is_covmodel(self._model) and self._z_1 is not None and self._z_2 is not None
                    and self._spectrum_factor is not None):
                raise ValueError(f"{self.name}.update: needs a model, a seed, a period or mode numbers")
        else:
            raise ValueError(f"{self.name}.update: model must be a CovModel, got {type(model).__name__}")

    def reset_seed(self, seed=np.nan):
        """Redraw z_1, z_2 and recompute the spectrum factors (``generator.py:791-820``)."""
        if seed is None or not np.isnan(seed):
            self._seed = seed
        self._rng = RNG(self._seed)
        n = int(np.prod(self._mode_no))
        self._z_1 = self._rng.random.normal(size=n)
        self._z_2 = self._rng.random.normal(size=n)
        if self.zero_var:
            self._spectrum_factor = np.full(n, 0.0)
        else:
            k_norm = np.linalg.norm(self._modes, axis=0)
            self._spectrum_factor = np.sqrt(self._model.spectrum(k_norm) * np.prod(self._delta_k))
        self._dev_modes = None

    def _fill_to_dim(self, values, dim, dtype=float, default_value=None):
        """Pad ``values`` with its last entry up to ``dim`` entries (``generator.py:822-836``)."""
        r = np.atleast_1d(values)
        if values is None:
            if default_value is None:
                raise ValueError("Fourier: a value or a default is required")
            r = default_value
        r = np.atleast_1d(np.array(r, dtype=dtype))[:dim]
        if len(r) < dim:
            r = np.pad(r, (0, dim - len(r)), "edge")
        return r

    def _set_modes(self, mode_no, dim):
        """Regular wavenumber grid (``generator.py:838-862``): ``arange(-m/2 dk, m/2 dk, dk)`` per axis."""
        axes = [np.arange(-mode_no[d] / 2.0 * self._delta_k[d], mode_no[d] / 2.0 * self._delta_k[d],
                          self._delta_k[d]) for d in range(dim)]
        self._modes = np.asarray(np.meshgrid(*axes, indexing="ij"), dtype=np.double).reshape((dim, -1))
        self._kaxes = [np.asarray(m, dtype=np.double) for m in axes]
        self._mode_no = [len(m) for m in axes]

    @property
    def precision(self):
        return self._precision

    @precision.setter
    def precision(self, precision):
        if precision not in PRECISIONS:
            raise ValueError(f"Fourier: precision not in {PRECISIONS}.")
        self._precision = precision

    @property
    def seed(self):
        return self._seed

    @seed.setter
    def seed(self, new_seed):
        if new_seed is not self._seed:
            self.reset_seed(new_seed)

    @property
    def model(self):
        return self._model

    @model.setter
    def model(self, model):
        self.update(model)

    @property
    def modes(self):
        return self._modes

    @property
    def mode_no(self):
        return self._mode_no

    @mode_no.setter
    def mode_no(self, mode_no):
        self.update(mode_no=mode_no)

    @property
    def period(self):
        return self._period

    @period.setter
    def period(self, period):
        self.update(period=period)

    @property
    def value_type(self):
        return self._value_type

    def __repr__(self):
        return f"{self.name}(model={self.model}, seed={self.seed})"
