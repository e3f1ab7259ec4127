"""Symmetric n-dimensional Fourier transform of radial functions (Ogata 2005).

The reference computes spectral densities of models without an analytic one
(the Stable model) with ``hankel.SymmetricFourierTransform(ndim, a=-1, b=1,
N=200, h=0.001)`` (``src/gstools/covmodel/base.py:59``, ``:535-551``,
``:1044-1052``); the ``hankel`` package is not vendored.  This module restates
the published method:

    S(k) = (2 pi)^-n  (2 pi)^(n/2) k^(1-n/2) int_0^inf r^(n/2) f(r) J_nu(k r) dr,
    nu = n/2 - 1,

with the integral evaluated by H. Ogata, "A numerical integration formula
based on the Bessel functions", Publ. RIMS 41 (2005) 949-970, using the
double-exponential change of variables psi(t) = t tanh(pi/2 sinh t):

    int_0^inf g(x) J_nu(x) dx ~ pi sum_m w_m g(pi psi(h xi_m)/h) J_nu(.) psi'(h xi_m),
    xi_m = j_{nu,m} / pi,  w_m = Y_nu(pi xi_m) / J_{nu+1}(pi xi_m).

The quadrature's values are not pinned by any reference test; the tests check
it against the closed-form Gaussian, Exponential and Matern densities.
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
    """Radial Fourier transform in ``ndim`` dimensions, normalised by (2 pi)^-n."""

    def __init__(self, ndim=2, N=200, h=0.001):
        self.ndim = int(ndim)
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

    def transform(self, f, k):
        """S(k) for scalar or array ``k`` (taken by absolute value)."""
        k = np.asarray(np.abs(k), dtype=np.double)
        shape = k.shape
        k = np.atleast_1d(k).ravel()
        out = np.empty_like(k)
        n = self.ndim
        norm = (2 * np.pi) ** (-n) * (2 * np.pi) ** (n / 2.0)
        for idx, kk in enumerate(k):
            if kk == 0.0:
                # limit k -> 0: (2pi)^-n * surface(n) * int r^(n-1) f(r) dr
                # is approximated by the smallest positive k the quadrature handles
                kk = 1e-12
            r = self._x / kk
            g = r ** (n / 2.0) * f(r)
            integral = np.sum(self._series * g) / kk
            out[idx] = norm * kk ** (1.0 - n / 2.0) * integral
        return out.reshape(shape)
