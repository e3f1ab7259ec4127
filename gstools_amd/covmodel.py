"""Spectral members of the covariance models the randomisation method samples.

Only what ``RandMeth.reset_seed`` (``src/gstools/field/generator.py:346-387``)
and ``Field.pre_pos`` (``src/gstools/field/base.py:254-297``) read from a
model is restated here:

* parameters ``dim``, ``var``, ``len_scale`` (scalar or per-axis list),
  ``nugget``, ``anis``, ``angles``, ``rescale`` and ``len_rescaled``
  (``covmodel/base.py:934-1000``, ``covmodel/tools.py:186-280``);
* ``spectral_density``, ``spectral_rad_pdf``, ``ln_spectral_rad_pdf``,
  ``dist_func``, ``has_ppf`` (``covmodel/base.py:498-570``,
  ``covmodel/tools.py:113-138``, ``:374-405``);
* ``isometrize`` (``covmodel/base.py:572-582``);
* equality as in ``covmodel/tools.py:570-598``.

Models: Gaussian (``covmodel/models.py:118-189``), Exponential (``:192-273``),
Stable (``:276-348``, spectral density by the Hankel transform) and Matern
(``:351-458``).  The generators in this package also accept the reference's
own ``CovModel`` objects, which expose the same members.
"""

import copy

import numpy as np
import scipy.special as sps

from .geometry import matrix_anisometrize, matrix_isometrize, set_angles, set_anis
from .hankel import SymmetricFourierTransform

__all__ = ["CovModel", "Gaussian", "Exponential", "Stable", "Matern", "set_len_anis", "rad_fac", "HANKEL_DEFAULT"]

# Keyword arguments of the radial Fourier transform behind numerical spectral
# densities: the values GSTools uses (``covmodel/base.py:59``).
HANKEL_DEFAULT = {"a": -1, "b": 1, "N": 200, "h": 0.001, "alt": True}


def rad_fac(dim, r):
    """Volume element of n-dimensional spherical coordinates (``covmodel/tools.py:113-138``)."""
    if dim == 1:
        return 2.0
    if dim == 2:
        return 2 * np.pi * r
    if dim == 3:
        return 4 * np.pi * r**2
    return dim * r ** (dim - 1) * np.sqrt(np.pi) ** dim / sps.gamma(dim / 2 + 1)


def set_len_anis(dim, len_scale, anis):
    """Main length scale and anisotropy ratios (``covmodel/tools.py:186-245``)."""
    ls_tmp = np.atleast_1d(np.array(len_scale, dtype=np.double))[:dim]
    out_len_scale = ls_tmp[0]
    if len(ls_tmp) == 1:
        out_anis = set_anis(dim, anis)
    else:
        if len(ls_tmp) < dim:
            ls_tmp = np.pad(ls_tmp, (0, dim - len(ls_tmp)), "edge")
        out_anis = np.zeros(dim - 1, dtype=np.double)
        for i in range(1, dim):
            out_anis[i - 1] = ls_tmp[i] / ls_tmp[0]
    for ani in out_anis:
        if not ani > 0.0:
            raise ValueError(f"anisotropy-ratios needs to be > 0, got: {out_anis}")
    return out_len_scale, out_anis


class CovModel:
    """Base class: parameters, geometry and the generic spectral pdf."""

    _opt_defaults = {}

    def __init__(self, dim=3, var=1.0, len_scale=1.0, nugget=0.0, anis=1.0, angles=0.0,
                 rescale=None, hankel_kw=None, **opt_arg):
        self._init = False
        self._dim = None
        self._len_scale = None
        self._anis = None
        self._angles = None
        self._hankel_kw = None
        self._sft = None
        self.hankel_kw = hankel_kw
        for name, default in self._opt_defaults.items():
            setattr(self, name, float(opt_arg.pop(name, default)))
        if opt_arg:
            raise ValueError(f"{self.name}: unknown arguments {sorted(opt_arg)}")
        self.dim = dim
        self._len_scale, self._anis = set_len_anis(self.dim, len_scale, anis)
        self._angles = set_angles(self.dim, angles)
        self.var = var
        self.nugget = nugget
        self.rescale = rescale
        self._init = True

    # parameters ---------------------------------------------------------
    @property
    def name(self):
        return self.__class__.__name__

    @property
    def opt_arg(self):
        return list(self._opt_defaults)

    @property
    def dim(self):
        return self._dim

    @dim.setter
    def dim(self, dim):
        if dim < 1:
            raise ValueError("Only dimensions of d >= 1 are supported.")
        self._dim = int(dim)
        if self.needs_fourier_transform:
            self._sft = SymmetricFourierTransform(ndim=self._dim, **self._hankel_kw)
        if self._init:
            self._len_scale, self._anis = set_len_anis(self._dim, self._len_scale, self._anis)
            self._angles = set_angles(self._dim, self._angles)

    @property
    def hankel_kw(self):
        """Keyword arguments of the radial Fourier transform (``covmodel/base.py:1040-1052``).

        ``None`` for models with an analytic spectral density.  Assigning a
        dict updates the current keywords (starting from
        :data:`HANKEL_DEFAULT`), assigning ``None`` restores the defaults;
        either rebuilds the transform.
        """
        return self._hankel_kw

    @hankel_kw.setter
    def hankel_kw(self, hankel_kw):
        if not self.needs_fourier_transform:
            return
        if self._hankel_kw is None or hankel_kw is None:
            self._hankel_kw = dict(HANKEL_DEFAULT)
        if hankel_kw is not None:
            self._hankel_kw.update(hankel_kw)
        if self._dim is not None:
            self._sft = SymmetricFourierTransform(ndim=self._dim, **self._hankel_kw)

    @property
    def var(self):
        return self._var

    @var.setter
    def var(self, var):
        self._var = float(var)

    @property
    def nugget(self):
        return self._nugget

    @nugget.setter
    def nugget(self, nugget):
        self._nugget = float(nugget)

    @property
    def sill(self):
        return self.var + self.nugget

    @property
    def len_scale(self):
        return self._len_scale

    @len_scale.setter
    def len_scale(self, len_scale):
        self._len_scale, self._anis = set_len_anis(self.dim, len_scale, self._anis)

    @property
    def anis(self):
        return self._anis

    @anis.setter
    def anis(self, anis):
        self._len_scale, self._anis = set_len_anis(self.dim, self._len_scale, anis)

    @property
    def angles(self):
        return self._angles

    @angles.setter
    def angles(self, angles):
        self._angles = set_angles(self.dim, angles)

    @property
    def rescale(self):
        return self._rescale

    @rescale.setter
    def rescale(self, rescale):
        rescale = self.default_rescale() if rescale is None else rescale
        self._rescale = abs(float(rescale))

    def default_rescale(self):
        return 1.0

    @property
    def len_rescaled(self):
        return self.len_scale / self.rescale

    # correlation --------------------------------------------------------
    def cor(self, h):
        raise NotImplementedError

    def correlation(self, r):
        return self.cor(np.abs(np.asarray(r, dtype=np.double)) / self.len_rescaled)

    def covariance(self, r):
        """``var * correlation(r)`` (``covmodel/tools.py:65-67``)."""
        return self.var * self.correlation(r)

    def variogram(self, r):
        """``var - covariance(r) + nugget`` (``covmodel/tools.py:61-63``)."""
        return self.var - self.covariance(r) + self.nugget

    def cov_nugget(self, r):
        """Covariance with the nugget at r = 0: the sill there (``covmodel/base.py:313-320``)."""
        r = np.asarray(np.abs(r), dtype=np.double)
        r_gz = np.logical_not(np.isclose(r, 0))
        res = np.empty_like(r, dtype=np.double)
        res[r_gz] = self.covariance(r[r_gz])
        res[np.logical_not(r_gz)] = self.sill
        return res

    # spectrum -----------------------------------------------------------
    @property
    def needs_fourier_transform(self):
        return type(self).spectral_density is CovModel.spectral_density

    def spectrum(self, k):
        """Spectrum ``var * spectral_density(k)`` (``covmodel/base.py:511-532``)."""
        return self.spectral_density(k) * self.var

    def spectral_density(self, k):
        """Spectral density by the radial Fourier transform of the correlation."""
        k = np.asarray(np.abs(k), dtype=np.double)
        return self._sft.transform(self.correlation, k)

    def spectral_rad_pdf(self, r):
        """Radial spectral pdf (``covmodel/tools.py:374-405``)."""
        r = np.asarray(np.abs(r), dtype=np.double)
        if self.dim > 1:
            r_gz = np.logical_not(np.isclose(r, 0))
            res = np.zeros_like(r, dtype=np.double)
            res[r_gz] = rad_fac(self.dim, r[r_gz]) * np.abs(self.spectral_density(r[r_gz]))
        else:
            res = rad_fac(self.dim, r) * np.abs(self.spectral_density(r))
            res = np.asarray(res, dtype=np.double)
        res[np.logical_not(np.isfinite(res))] = 0.0
        return np.maximum(res, 0.0)

    def ln_spectral_rad_pdf(self, r):
        with np.errstate(divide="ignore"):
            return np.log(self.spectral_rad_pdf(r))

    def _has_cdf(self):
        return hasattr(self, "spectral_rad_cdf")

    def _has_ppf(self):
        return hasattr(self, "spectral_rad_ppf")

    @property
    def has_cdf(self):
        return self._has_cdf()

    @property
    def has_ppf(self):
        return self._has_ppf()

    @property
    def dist_func(self):
        pdf = self.spectral_rad_pdf
        cdf = self.spectral_rad_cdf if self.has_cdf else None
        ppf = self.spectral_rad_ppf if self.has_ppf else None
        return pdf, cdf, ppf

    # geometry -----------------------------------------------------------
    @property
    def field_dim(self):
        return self.dim

    def anisometrize(self, pos):
        pos = np.asarray(pos, dtype=np.double).reshape((self.dim, -1))
        return np.dot(matrix_anisometrize(self.dim, self.angles, self.anis), pos)

    def isometrize(self, pos):
        pos = np.asarray(pos, dtype=np.double).reshape((self.field_dim, -1))
        return np.dot(matrix_isometrize(self.dim, self.angles, self.anis), pos)

    # comparison ---------------------------------------------------------
    def __eq__(self, other):
        if not isinstance(other, CovModel):
            return False
        if set(self.opt_arg) != set(other.opt_arg) or self.dim != other.dim:
            return False
        equal = self.name == other.name
        equal &= bool(np.isclose(self.var, other.var))
        equal &= bool(np.isclose(self.nugget, other.nugget))
        equal &= bool(np.isclose(self.len_scale, other.len_scale))
        equal &= bool(np.all(np.isclose(self.anis, other.anis)))
        equal &= bool(np.all(np.isclose(self.angles, other.angles)))
        equal &= bool(np.isclose(self.rescale, other.rescale))
        for opt in self.opt_arg:
            equal &= bool(np.isclose(getattr(self, opt), getattr(other, opt)))
        return equal

    def __ne__(self, other):
        return not self.__eq__(other)

    __hash__ = None

    def copy(self):
        return copy.deepcopy(self)

    def __repr__(self):
        opt = "".join(f", {o}={getattr(self, o):.4g}" for o in self.opt_arg)
        return (f"{self.name}(dim={self.dim}, var={self.var:.4g}, len_scale={self.len_scale:.4g}, "
                f"nugget={self.nugget:.4g}, anis={list(self.anis)}, angles={list(self.angles)}{opt})")


class Gaussian(CovModel):
    """Gaussian model, rescale sqrt(pi)/2 (``covmodel/models.py:118-189``)."""

    def cor(self, h):
        return np.exp(-(h**2))

    def default_rescale(self):
        return np.sqrt(np.pi) / 2.0

    def spectral_density(self, k):
        k = np.asarray(k, dtype=np.double)
        return (self.len_rescaled / 2.0 / np.sqrt(np.pi)) ** self.dim * np.exp(
            -((k * self.len_rescaled / 2.0) ** 2))

    def spectral_rad_cdf(self, r):
        r = np.asarray(r, dtype=np.double)
        if self.dim == 1:
            return sps.erf(r * self.len_rescaled / 2.0)
        if self.dim == 2:
            return 1.0 - np.exp(-((r * self.len_rescaled / 2.0) ** 2))
        if self.dim == 3:
            return sps.erf(r * self.len_rescaled / 2.0) - r * self.len_rescaled / np.sqrt(np.pi) * np.exp(
                -((r * self.len_rescaled / 2.0) ** 2))
        return None

    def spectral_rad_ppf(self, u):
        u = np.asarray(u, dtype=np.double)
        if self.dim == 1:
            return 2.0 / self.len_rescaled * sps.erfinv(u)
        if self.dim == 2:
            return 2.0 / self.len_rescaled * np.sqrt(-np.log(1.0 - u))
        return None

    def _has_cdf(self):
        return self.dim in [1, 2, 3]

    def _has_ppf(self):
        return self.dim in [1, 2]


class Spherical(CovModel):
    """Spherical model (``covmodel/models.py:745-781``); spectral density by the Hankel transform."""

    def cor(self, h):
        h = np.minimum(np.abs(h, dtype=np.double), 1.0)
        return 1.0 - 1.5 * h + 0.5 * h**3


class Exponential(CovModel):
    """Exponential model (``covmodel/models.py:192-273``)."""

    def cor(self, h):
        return np.exp(-h)

    def spectral_density(self, k):
        k = np.asarray(k, dtype=np.double)
        return (self.len_rescaled**self.dim * sps.gamma((self.dim + 1) / 2.0)
                / (np.pi * (1.0 + (k * self.len_rescaled) ** 2)) ** ((self.dim + 1) / 2.0))

    def spectral_rad_cdf(self, r):
        r = np.asarray(r, dtype=np.double)
        if self.dim == 1:
            return np.arctan(r * self.len_rescaled) * 2.0 / np.pi
        if self.dim == 2:
            return 1.0 - 1.0 / np.sqrt(1.0 + (r * self.len_rescaled) ** 2)
        if self.dim == 3:
            return ((np.arctan(r * self.len_rescaled)
                     - r * self.len_rescaled / (1.0 + (r * self.len_rescaled) ** 2)) * 2.0 / np.pi)
        return None

    def spectral_rad_ppf(self, u):
        u = np.asarray(u, dtype=np.double)
        if self.dim == 1:
            return np.tan(np.pi / 2 * u) / self.len_rescaled
        if self.dim == 2:
            u_power = np.divide(1, u**2, out=np.full_like(u, np.inf),
                                where=np.logical_not(np.isclose(u, 0)))
            return np.sqrt(u_power - 1.0) / self.len_rescaled
        return None

    def _has_cdf(self):
        return self.dim in [1, 2, 3]

    def _has_ppf(self):
        return self.dim in [1, 2]


def stable_density_tail(y, d, alpha, nmax=300):
    """Large-|y| series of the d-dimensional isotropic alpha-stable density.

    ``exp(-|t|^alpha)`` is the characteristic function of an isotropic
    alpha-stable law, whose density p_d(y) = (2 pi)^-d int exp(-|t|^alpha)
    e^{-i t.y} d^d t has the series (convergent for alpha < 1, asymptotic for
    1 <= alpha < 2)

        p_d(y) = pi^-(d/2+1) sum_{n>=1} (-1)^(n+1) / n! 2^(n alpha)
                 Gamma(n alpha / 2 + 1) Gamma((n alpha + d) / 2)
                 sin(n pi alpha / 2) |y|^-(n alpha + d).

    The sum stops before the first growing term.  Returns (value, estimated
    relative error = smallest term used / |value|).
    """
    y = np.atleast_1d(np.asarray(y, dtype=np.double))
    out = np.zeros_like(y)
    last = np.full_like(y, np.inf)
    stop = np.zeros(y.shape, dtype=bool)
    logy = np.log(y)
    for n in range(1, nmax + 1):
        s = np.sin(n * np.pi * alpha / 2.0)
        if abs(s) < 1e-12:
            continue
        logc = (n * alpha * np.log(2.0) + sps.gammaln(n * alpha / 2.0 + 1.0)
                + sps.gammaln((n * alpha + d) / 2.0) - sps.gammaln(n + 1.0))
        mag = np.exp(np.minimum(logc - (n * alpha + d) * logy, 700.0)) * abs(s)
        stop |= mag > last
        out = np.where(stop, out, out + (1.0 if n % 2 else -1.0) * np.sign(s) * mag)
        last = np.where(stop, last, mag)
    val = out / np.pi ** (d / 2.0 + 1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        err = np.where(val > 0, last / np.pi ** (d / 2.0 + 1.0) / np.abs(val), np.inf)
    return val, err


class Stable(CovModel):
    """Stable model exp(-h^alpha) (``covmodel/models.py:276-348``).

    The reference has no analytic spectral density for this model and takes
    the radial Fourier transform of the correlation by Ogata's quadrature
    with ``HANKEL_DEFAULT`` (``covmodel/base.py:59``, ``:535-551``).  That is
    the default here too (``density="hankel"``), so the drop-in draws the
    reference's modes.  Its nodes reach r = 197 / k only, so beyond
    k L ~ 20 the quadrature truncates the integral and returns an O(k^-2)
    artefact (the transform of the correlation's value near r = 0); through
    the |.| in the radial pdf (``covmodel/tools.py:395``) the pdf then
    decays like 1/k, which is not normalisable, and MCMC radii drift without
    bound (2D, alpha 1.5, len 15, 2048 modes: median |k| ~ 21 where ~0.1 is
    expected).

    ``density="series"`` (opt-in) keeps the quadrature where it converges
    and switches to the stable-density series (:func:`stable_density_tail`)
    where that series is accurate to 1e-8; alpha = 2 is the Gaussian closed
    form.  Both branches are checked against mpmath in the tests.
    """

    _opt_defaults = {"alpha": 1.5}
    DENSITIES = ("hankel", "series")

    def __init__(self, *args, density="hankel", **kwargs):
        if density not in self.DENSITIES:
            raise ValueError(f"Stable: density must be one of {self.DENSITIES}, got {density!r}")
        self._density = density
        super().__init__(*args, **kwargs)

    @property
    def density(self):
        """``"hankel"`` (the reference's quadrature) or ``"series"`` (corrected tail)."""
        return self._density

    @property
    def needs_fourier_transform(self):
        return True  # both densities use the quadrature at small k L

    def cor(self, h):
        return np.exp(-np.power(h, self.alpha))

    def spectral_density(self, k):
        if self._density == "hankel":
            return CovModel.spectral_density(self, k)
        k = np.asarray(np.abs(k), dtype=np.double)
        shape = k.shape
        k = np.atleast_1d(k).ravel()
        ell = self.len_rescaled
        d = self.dim
        if self.alpha >= 2.0:
            res = (ell / 2.0 / np.sqrt(np.pi)) ** d * np.exp(-((k * ell / 2.0) ** 2))
            return res.reshape(shape)
        res = np.empty_like(k)
        y = k * ell
        big = y > 1.0
        use_tail = np.zeros_like(big)
        if np.any(big):
            val, err = stable_density_tail(y[big], d, self.alpha)
            ok = err < 1e-8
            idx = np.flatnonzero(big)
            res[idx[ok]] = ell**d * val[ok]
            use_tail[idx[ok]] = True
        rest = ~use_tail
        if np.any(rest):
            res[rest] = self._sft.transform(self.correlation, k[rest])
        return res.reshape(shape)

    def __eq__(self, other):
        # a model without the switch (the reference's) uses the quadrature
        return super().__eq__(other) and getattr(other, "density", "hankel") == self._density

    __hash__ = None

    def __repr__(self):
        base = super().__repr__()
        return base if self._density == "hankel" else base[:-1] + f", density={self._density!r})"


class Matern(CovModel):
    """Matern model (``covmodel/models.py:351-458``)."""

    _opt_defaults = {"nu": 1.0}

    def cor(self, h):
        h = np.asarray(np.abs(h), dtype=np.double)
        if self.nu > 20.0:
            return np.exp(-((h / 2.0) ** 2))
        h_gz = h[h > 0.0]
        res = np.ones_like(h)
        res[h > 0.0] = np.exp((1.0 - self.nu) * np.log(2) - sps.loggamma(self.nu)
                              + self.nu * np.log(np.sqrt(self.nu) * h_gz)) * sps.kv(self.nu, np.sqrt(self.nu) * h_gz)
        res[np.logical_not(np.isfinite(res))] = 0.0
        return np.maximum(res, 0.0)

    def spectral_density(self, k):
        k = np.asarray(k, dtype=np.double)
        x = (k * self.len_rescaled) ** 2
        if self.nu > 20.0:
            return ((self.len_rescaled / np.sqrt(np.pi)) ** self.dim * np.exp(-x)
                    * (1 + 0.5 * x**2 / self.nu) * np.sqrt(1 + x / self.nu) ** (-self.dim))
        return (self.len_rescaled / np.sqrt(np.pi)) ** self.dim * np.exp(
            -(self.nu + self.dim / 2.0) * np.log(1.0 + x / self.nu)
            + sps.loggamma(self.nu + self.dim / 2.0) - sps.loggamma(self.nu)
            - self.dim * np.log(np.sqrt(self.nu)))
