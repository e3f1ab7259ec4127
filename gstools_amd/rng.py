"""Seed streams and mode samplers of the randomisation method.

Mirrors ``gstools.random`` for the draws that ``RandMeth.reset_seed``
(``src/gstools/field/generator.py:346-387``) makes:

* :class:`MasterRNG` -- ``random/tools.py:19-37``: a legacy numpy
  ``RandomState(seed)`` whose ``randint(1, 2**16)`` seeds a fresh
  ``RandomState`` on every access of :attr:`RNG.random`
  (``random/rng.py:195-203``).
* :meth:`RNG.sample_sphere` -- ``random/rng.py:142-193``.
* :meth:`RNG.sample_dist` -- ``random/rng.py:106-140`` -> scipy
  ``rv_continuous.rvs``, whose default ``_rvs`` is ``ppf(U)`` with
  ``U = random_state.uniform(size)``; ``loc = 0`` and ``scale = 1`` leave it
  unchanged.  Only the inversion route (a ppf is given) is needed here.
* :meth:`RNG.sample_ln_pdf` -- ``random/rng.py:38-104``: an ensemble MCMC
  run with the affine-invariant stretch move (see :mod:`.mcmc`).

All randomness comes from numpy's legacy ``RandomState`` (MT19937), whose
streams numpy keeps frozen, so identical seeds give identical modes.
"""

import numpy as np
import numpy.random as rand

from .mcmc import StretchEnsemble

__all__ = ["MasterRNG", "RNG"]


class MasterRNG:
    """Seed generator: ``RandomState(seed).randint(1, 2**16)`` per call.

    Mirrors ``random/tools.py:19-37``.
    """

    def __init__(self, seed):
        self._seed = seed
        self._state = rand.RandomState(seed)

    def __call__(self):
        return self._state.randint(1, 2**16)

    @property
    def seed(self):
        return self._seed

    def __repr__(self):
        return f"MasterRNG(seed={self.seed})"


class RNG:
    """Random number generator with independent sub-streams.

    Mirrors ``random/rng.py:22-220``.
    """

    def __init__(self, seed=None):
        self._master_rng = None
        self.seed = seed

    @property
    def random(self):
        """A fresh ``RandomState`` seeded from the master stream."""
        return rand.RandomState(self._master_rng())

    @property
    def seed(self):
        return self._master_rng.seed

    @seed.setter
    def seed(self, new_seed=None):
        self._master_rng = MasterRNG(new_seed)

    def sample_sphere(self, dim, size=None):
        """Uniform directions on the unit sphere, shape (dim, size).

        Mirrors ``random/rng.py:142-193`` for the dimensions the covariance
        models here support (1D: random sign; 2D: one angle; 3D: azimuth plus
        uniform cosine).
        """
        if size is None:
            coord = np.empty((dim, 1), dtype=np.double)
        else:
            coord = np.empty((dim,) + tuple(np.atleast_1d(size)), dtype=np.double)
        if dim == 1:
            coord[0] = self.random.choice([-1, 1], size=size)
        elif dim == 2:
            ang1 = self.random.uniform(0.0, 2 * np.pi, size)
            coord[0] = np.cos(ang1)
            coord[1] = np.sin(ang1)
        elif dim == 3:
            ang1 = self.random.uniform(0.0, 2 * np.pi, size)
            ang2 = self.random.uniform(-1.0, 1.0, size)
            coord[0] = np.sqrt(1.0 - ang2**2) * np.cos(ang1)
            coord[1] = np.sqrt(1.0 - ang2**2) * np.sin(ang1)
            coord[2] = ang2
        else:
            raise ValueError(f"sample_sphere: dimensions 1 to 3 are supported, got {dim}")
        return np.reshape(coord, dim) if size is None else coord

    def sample_dist(self, pdf=None, cdf=None, ppf=None, size=None, **kwargs):
        """Inverse-transform sampling: ``ppf(RandomState.uniform(size))``.

        Mirrors ``random/rng.py:106-140`` for the case where a ppf is given
        (the only case ``RandMeth`` uses, ``generator.py:373-379``).  The
        ``a`` (support) keyword is accepted and has no effect on the draw.
        """
        if ppf is None:
            raise ValueError("sample_dist: only inversion sampling (ppf given) is supported")
        state = self.random
        u = state.uniform(size=size)
        return np.asarray(ppf(u), dtype=np.double) * 1.0 + 0.0

    def sample_ln_pdf(self, ln_pdf, size=None, sample_around=1.0, nwalkers=50,
                      burn_in=20, oversampling_factor=10):
        """Sample radii from ``exp(ln_pdf)`` by ensemble MCMC.

        Mirrors ``random/rng.py:38-104``: initial walkers
        ``rand(nwalkers) * sample_around`` from one sub-stream, a burn-in run
        seeded from the next, a production run of
        ``int(max(burn_in, size / nwalkers * oversampling_factor))`` steps
        seeded from the next, and ``choice(flat_chain, size)`` from the next.
        """
        if size is None:
            sample_size = burn_in
        else:
            sample_size = max(burn_in, (size / nwalkers) * oversampling_factor)
        sample_size = int(sample_size)
        init_guess = self.random.rand(nwalkers).reshape((nwalkers, 1)) * sample_around
        sampler = StretchEnsemble(nwalkers, ln_pdf)
        coords, log_prob = sampler.run(init_guess, None, burn_in, self.random.get_state())
        sampler.reset()
        sampler.run(coords, log_prob, sample_size, self.random.get_state())
        samples = sampler.flat_chain()[:, 0]
        return self.random.choice(samples, size)

    def __repr__(self):
        return f"RNG(seed={self.seed})"
