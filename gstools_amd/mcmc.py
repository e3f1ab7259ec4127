"""Affine-invariant ensemble MCMC (stretch move) for radial mode sampling.

``RNG.sample_ln_pdf`` of the reference (``src/gstools/random/rng.py:38-104``)
drives ``emcee.EnsembleSampler(nwalkers, 1, ln_pdf, vectorize=True)``
(emcee >= 3.0.0, ``pyproject.toml:46``), which is not vendored.  This module
restates the published algorithm -- Goodman & Weare (2010), "Ensemble
samplers with affine invariance", with the red/blue parallel split of
Foreman-Mackey et al. (2013) -- and emcee 3's default move parameters
(stretch scale a = 2, two randomly shuffled halves, one move drawn per step
from a one-element move list).  Every random number is drawn from a legacy
numpy ``RandomState`` in this order per step:

1. one ``random_sample()`` choosing the move (the move list has one entry);
2. ``shuffle`` of the half labels ``arange(nwalkers) % 2``;
3. for each half in turn: ``rand(Ns)`` for the stretch factors,
   ``randint(Nc, size=Ns)`` for the partner walkers, then one uniform per
   proposed walker, in walker order, for the Metropolis test; the half is
   updated before the other half proposes.

Only this order is pinned, by the 3D known values of the reference tests
(``tests/test_randmeth.py:43-46``, ``tests/test_incomprrandmeth.py:40-44``).
"""

import numpy as np

__all__ = ["StretchEnsemble"]


class StretchEnsemble:
    """Ensemble of ``nwalkers`` one-dimensional walkers on ``exp(ln_pdf)``."""

    def __init__(self, nwalkers, ln_pdf, a=2.0):
        self.nwalkers = int(nwalkers)
        self.ln_pdf = ln_pdf
        self.a = float(a)
        self._random = np.random.RandomState()
        self._chain = []

    def reset(self):
        self._chain = []

    def _log_prob(self, coords):
        if np.any(np.isinf(coords)):
            raise ValueError("At least one parameter value was infinite")
        if np.any(np.isnan(coords)):
            raise ValueError("At least one parameter value was NaN")
        res = np.asarray(self.ln_pdf(coords), dtype=np.double)
        log_prob = res.reshape(len(coords), -1)[:, 0].copy()
        if np.any(np.isnan(log_prob)):
            raise ValueError("Probability function returned NaN")
        return log_prob

    def _propose(self, coords, log_prob):
        rs = self._random
        nw = self.nwalkers
        accepted = np.zeros(nw, dtype=bool)
        all_inds = np.arange(nw)
        inds = all_inds % 2
        rs.shuffle(inds)
        for split in range(2):
            sel = inds == split
            sets = [coords[inds == j] for j in range(2)]
            s = sets[split]
            c = np.concatenate(sets[:split] + sets[split + 1:], axis=0)
            ns, nc = len(s), len(c)
            ndim = s.shape[1]
            zz = ((self.a - 1.0) * rs.rand(ns) + 1) ** 2.0 / self.a
            factors = (ndim - 1.0) * np.log(zz)
            rint = rs.randint(nc, size=(ns,))
            q = c[rint] - (c[rint] - s) * zz[:, None]
            new_lp = self._log_prob(q)
            # one uniform per proposed walker, in walker order (rand(ns) draws
            # the same sequence as ns scalar rand() calls)
            lnpdiff = factors + new_lp - log_prob[sel]
            with np.errstate(invalid="ignore"):
                accepted[sel] = lnpdiff > np.log(rs.rand(ns))
            m1 = sel & accepted
            m2 = accepted[sel]
            coords[m1] = q[m2]
            log_prob[m1] = new_lp[m2]
        return coords, log_prob

    def run(self, coords, log_prob, nsteps, random_state):
        """Advance ``nsteps`` steps from ``coords``; returns the final state."""
        self._random.set_state(random_state)
        coords = np.array(coords, dtype=np.double, copy=True).reshape(self.nwalkers, -1)
        if np.all(coords == coords[0]):
            raise ValueError("Initial state has walkers that are not linearly independent")
        if log_prob is None:
            log_prob = self._log_prob(coords)
        else:
            log_prob = np.array(log_prob, dtype=np.double, copy=True)
        if np.any(np.isnan(log_prob)):
            raise ValueError("The initial log_prob was NaN")
        for _ in range(int(nsteps)):
            self._random.random_sample()  # choice of the (single) move
            coords, log_prob = self._propose(coords, log_prob)
            self._chain.append(coords.copy())
        return coords, log_prob

    def flat_chain(self):
        """Chain flattened step-major, walker-minor: (steps * nwalkers, 1)."""
        if not self._chain:
            return np.empty((0, 1))
        return np.concatenate(self._chain, axis=0)
