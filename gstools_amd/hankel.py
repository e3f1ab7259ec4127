"""Symmetric n-dimensional Fourier transform of radial functions (Ogata 2005).

The reference computes spectral densities of models without an analytic one
(the Stable model) with ``hankel.SymmetricFourierTransform(ndim, **hankel_kw)``
where ``hankel_kw`` defaults to ``HANKEL_DEFAULT = {"a": -1, "b": 1, "N": 200,
"h": 0.001, "alt": True}`` (``src/gstools/covmodel/base.py:59``, ``:535-551``,
``:1040-1052``); the ``hankel`` package (``>=1.0.0``, ``pyproject.toml:47``)
is not vendored.  This module restates its published method.  The transform
with Fourier convention (a, b) is

    F(k) = (|b| / (2 pi)^(1-a))^(n/2) int f(|r|) e^(i b k.r) d^n r
         = (|b| / (2 pi)^(1-a))^(n/2) (2 pi)^(n/2) (|b| k)^(1-n/2)
           int_0^inf r^(n/2) f(r) J_nu(|b| k r) dr,     nu = n/2 - 1,

so (a, b) = (-1, 1) gives the spectral density (2 pi)^-n int C(r) e^(ik.r).
The radial integral is H. Ogata's quadrature ("A numerical integration
formula based on the Bessel functions", Publ. RIMS 41 (2005) 949-970) with
the double-exponential change of variables psi(t) = t tanh(pi/2 sinh t):

    int_0^inf g(x) J_nu(x) dx ~ pi sum_m w_m g(pi psi(h xi_m)/h) J_nu(.) psi'(h xi_m),
    xi_m = j_{nu,m} / pi,  w_m = Y_nu(pi xi_m) / J_{nu+1}(pi xi_m).

The nodes reach x = pi psi(h xi_N)/h only (about 197 for N = 200,
h = 0.001), so for k beyond ~197 / (the correlation's range) the integral is
truncated: the reference's Stable density carries that truncation.

The quadrature's values are not pinned by any reference test; the tests check
it against the closed-form Gaussian, Exponential and Matern densities, and
the (a, b) convention against the direct d-dimensional integral.
"""

import numpy as np
import scipy.special as sps

__all__ = ["SymmetricFourierTransform"]


def _bessel_zeros_over_pi(nu, n):
    if nu == 0.5:
        return np.arange(1, n + 1, dtype=np.double)
    if nu == -0.5:
        return np.arange(1, n + 1, dtype=np.double) - 0.5
    if float(nu).is_integer():
        return sps.jn_zeros(int(nu), n) / np.pi
    # McMahon's asymptotic start, refined by Newton steps on J_nu.
    m = np.arange(1, n + 1, dtype=np.double)
    beta = (m + nu / 2 - 0.25) * np.pi
    x = beta - (4 * nu**2 - 1) / (8 * beta)
    for _ in range(20):
        x = x - sps.jv(nu, x) / sps.jvp(nu, x)
    return x / np.pi


class SymmetricFourierTransform:
    """Radial Fourier transform in ``ndim`` dimensions, convention (a, b).

    Defaults follow the ``hankel`` package's signature (a = 1, b = 1,
    N = 200, h = 0.05); the covariance models pass ``HANKEL_DEFAULT``.
    ``alt`` selects the package's arrangement of the Bessel factor; GSTools
    always passes ``alt=True``, the only arrangement restated here.
    """

    def __init__(self, ndim=2, a=1, b=1, N=200, h=0.05, alt=True):
        if not alt:
            raise NotImplementedError("SymmetricFourierTransform: only alt=True (GSTools' setting) is restated")
        if b == 0:
            raise ValueError("SymmetricFourierTransform: b must not be 0")
        self.ndim = int(ndim)
        self.a = float(a)
        self.b = float(b)
        self.alt = True
        self.nu = self.ndim / 2.0 - 1.0
        self.N = int(N)
        self.h = float(h)
        xi = _bessel_zeros_over_pi(self.nu, self.N)
        t = self.h * xi
        sinh_t = np.sinh(t)
        psi = t * np.tanh(0.5 * np.pi * sinh_t)
        dpsi = (np.pi * t * np.cosh(t) + np.sinh(np.pi * sinh_t)) / (1.0 + np.cosh(np.pi * sinh_t))
        dpsi = np.where(np.isfinite(dpsi), dpsi, 1.0)
        self._x = np.pi * psi / self.h
        w = sps.yv(self.nu, np.pi * xi) / sps.jv(self.nu + 1.0, np.pi * xi)
        self._series = np.pi * w * sps.jv(self.nu, self._x) * dpsi
        # (|b| / (2 pi)^(1-a))^(n/2) (2 pi)^(n/2)
        self._norm = (abs(self.b) / (2 * np.pi) ** (1.0 - self.a)) ** (self.ndim / 2.0) * (2 * np.pi) ** (
            self.ndim / 2.0)

    @property
    def r_max_times_k(self):
        """Largest node x = k r: beyond r = x / k the integral is cut off."""
        return float(self._x[-1])

    def transform(self, f, k):
        """F(k) for scalar or array ``k`` (taken by absolute value)."""
        k = np.asarray(np.abs(k), dtype=np.double)
        shape = k.shape
        k = np.atleast_1d(k).ravel() * abs(self.b)
        # the limit k -> 0 is approximated by the smallest positive k the
        # quadrature handles
        kk = np.where(k == 0.0, 1e-12, k)
        n = self.ndim
        r = self._x[None, :] / kk[:, None]
        g = r ** (n / 2.0) * f(r)
        integral = np.sum(self._series * g, axis=-1) / kk
        out = self._norm * kk ** (1.0 - n / 2.0) * integral
        return out.reshape(shape)
