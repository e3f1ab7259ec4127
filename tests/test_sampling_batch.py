"""Batch mode sampling (CPU): native stretch-move draws and the lockstep sampler.

* ``gsx_stretch_draws`` against numpy's own legacy ``RandomState`` calls in
  the order the stretch move makes them (``random_sample``, ``shuffle``,
  ``rand``, ``randint``, ``rand`` per half), draw for draw, including the
  generator state left behind;
* ``sample_modes_batch`` against one ``RandMeth`` per realization, bit for
  bit, for every model family and for the inversion and zero-variance paths;
* the reference's 3D MCMC known values (``tests/test_randmeth.py:43-46``,
  ``tests/test_incomprrandmeth.py:40-44``) from modes of the batch sampler.
"""

import ctypes

import numpy as np
import pytest

import oracle
from gstools_amd import Exponential, Gaussian, IncomprRandMeth, Matern, RandMeth, Stable, _native
from gstools_amd.sampling import _mt_state, sample_ln_pdf_batch, sample_modes_batch
from gstools_amd.rng import RNG

SEED = 19031977
X_T = np.linspace(0.0, 10.0, 10)
Y_T = np.linspace(-5.0, 5.0, 10)
Z_T = np.linspace(-6.0, 8.0, 10)


def _vp(a):
    return ctypes.c_void_p(a.ctypes.data)


def _numpy_draws(seed, steps, nw):
    rs = np.random.RandomState(seed)
    lab, uz, par, ua = [], [], [], []
    for _ in range(steps):
        rs.random_sample()
        inds = np.arange(nw) % 2
        rs.shuffle(inds)
        lab.append(inds.copy())
        z, p, u = [], [], []
        for split in range(2):
            ns = int(np.sum(inds == split))
            z.append(rs.rand(ns))
            p.append(rs.randint(nw - ns, size=(ns,)))
            u.append(rs.rand(ns))
        uz.append(np.concatenate(z))
        par.append(np.concatenate(p))
        ua.append(np.concatenate(u))
    return np.array(lab), np.array(uz), np.array(par), np.array(ua), rs.get_state()


@pytest.mark.parametrize("nw", [50, 7, 2])
def test_native_draws_equal_numpy(nw):
    seeds = [1, 65535, 12345, 4242]
    steps = np.array([40, 3, 17, 0], dtype=np.int64)
    ms = int(steps.max())
    states = np.stack([_mt_state(np.random.RandomState(s)) for s in seeds])
    shape = (len(seeds), ms, nw)
    lab, uz, par, ua = np.zeros(shape, np.int8), np.zeros(shape), np.zeros(shape, np.int64), np.zeros(shape)
    st = _native.lib().gsx_stretch_draws(len(seeds), _vp(states), _vp(steps), ms, nw, _vp(lab), _vp(uz), _vp(par),
                                         _vp(ua), 3)
    assert st == 0
    for r, seed in enumerate(seeds):
        n = int(steps[r])
        l_np, z_np, p_np, u_np, state = _numpy_draws(seed, n, nw)
        if n:
            assert np.array_equal(l_np, lab[r, :n])
            assert np.array_equal(z_np, uz[r, :n])
            assert np.array_equal(p_np, par[r, :n])
            assert np.array_equal(u_np, ua[r, :n])
        assert np.array_equal(state[1], states[r, :624]) and state[2] == states[r, 624]


def test_native_draws_reject_bad_input():
    lib = _native.lib()
    states = np.stack([_mt_state(np.random.RandomState(3))])
    steps = np.array([5], dtype=np.int64)
    buf = [np.zeros((1, 4, 10), t) for t in (np.int8, np.double, np.int64, np.double)]
    assert lib.gsx_stretch_draws(1, _vp(states), _vp(steps), 4, 10, *map(_vp, buf), 1) == -1  # steps > max
    states[0, 624] = 700
    steps[0] = 2
    assert lib.gsx_stretch_draws(1, _vp(states), _vp(steps), 4, 10, *map(_vp, buf), 1) == -1  # bad position
    assert lib.gsx_stretch_draws(1, _vp(states), _vp(steps), 4, 1, *map(_vp, buf), 1) == -1  # one walker


CASES = [
    (lambda: Exponential(dim=3, var=2.0, len_scale=[50, 20, 5], angles=[0.5, 0.2, 0.1]), [1000, 300, 64], [11, 22, 44]),
    (lambda: Gaussian(dim=3, var=1.0, len_scale=10.0), [500, 100, 7], [5, 6, 7]),
    (lambda: Matern(dim=3, var=1.0, len_scale=8.0, nu=1.5), [256, 257], [7, 8]),
    (lambda: Stable(dim=2, var=1.0, len_scale=15.0, alpha=1.5), [120, 64], [9, 10]),
    (lambda: Gaussian(dim=2, var=1.0, len_scale=10.0), [100, 50], [1, 2]),     # inversion
    (lambda: Gaussian(dim=3, var=0.0, len_scale=10.0), [40], [3]),             # zero variance
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_batch_equals_one_at_a_time(case):
    make, mode_nos, seeds = CASES[case]
    model = make()
    got = sample_modes_batch(model, mode_nos, seeds)
    assert len(got) == len(mode_nos)
    for (cov, z1, z2), n, s in zip(got, mode_nos, seeds):
        rm = RandMeth(model, mode_no=n, seed=s)
        assert np.array_equal(cov, rm._cov_sample)
        assert np.array_equal(z1, rm._z_1) and np.array_equal(z2, rm._z_2)


def test_forced_mcmc_and_odd_walkers():
    """``sampling="mcmc"`` on a model with a ppf, and an odd walker count."""
    model = Gaussian(dim=2, var=1.0, len_scale=4.0)
    got = sample_modes_batch(model, [90, 33], [13, 14], sampling="mcmc")
    for (cov, _, _), n, s in zip(got, [90, 33], [13, 14]):
        assert np.array_equal(cov, RandMeth(model, mode_no=n, seed=s, sampling="mcmc")._cov_sample)
    ln_pdf = model.ln_spectral_rad_pdf
    rngs = [RNG(21), RNG(22)]
    batch = sample_ln_pdf_batch(rngs, ln_pdf, [40, 25], sample_around=0.25, nwalkers=7, burn_in=5)
    for seed, n, b in zip((21, 22), (40, 25), batch):
        one = RNG(seed).sample_ln_pdf(ln_pdf, size=n, sample_around=0.25, nwalkers=7, burn_in=5)
        assert np.array_equal(one, b)


def test_reference_known_values_through_batch():
    """3D MCMC values of ``tests/test_randmeth.py:43-46`` and ``tests/test_incomprrandmeth.py:40-44``."""
    model = Gaussian(dim=3, var=1.5, len_scale=3.5)
    (cov, z1, z2), _ = sample_modes_batch(model, [100, 200], [SEED, SEED + 1])
    pos = np.array([X_T, Y_T, Z_T])
    f = np.sqrt(1.5 / 100) * oracle.summate_c(cov, z1, z2, pos)
    assert f[0] == pytest.approx(1.3240234883187239, abs=1e-12)
    assert f[1] == pytest.approx(1.6367244277732766, abs=1e-12)
    m3 = Gaussian(dim=3, var=1.5, len_scale=2.5)
    (cov, z1, z2), = sample_modes_batch(m3, [100], [SEED])
    assert np.array_equal(cov, IncomprRandMeth(m3, mode_no=100, seed=SEED)._cov_sample)
    v = oracle.summate_incompr_c(cov, z1, z2, pos) * np.sqrt(1.5 / 100)
    v[0] += 1.0  # mean velocity 1 along e1
    assert [v[0, 0], v[0, 1], v[1, 0]] == pytest.approx([0.7924546333550331, 1.660747056686244,
                                                         -0.28049855754819514], abs=1e-12)


def test_batch_argument_errors():
    model = Gaussian(dim=3)
    with pytest.raises(ValueError):
        sample_modes_batch(model, [10, 20], [1])
    with pytest.raises(ValueError):
        sample_modes_batch(model, [10], [1], sampling="gibbs")
    assert sample_modes_batch(model, [], []) == []


@pytest.mark.parametrize("dim,kind", [(1, 1), (2, 1), (3, 2), (2, 2), (1, 2), (3, 0)])
def test_native_mode_draws_equal_numpy(dim, kind):
    """``gsx_mode_draws`` against numpy's own legacy calls, realization by
    realization: master ``RandomState(seed).randint(1, 2**16)`` stream seeds,
    ``normal``, the sphere's ``choice`` / ``uniform`` draws, ``uniform`` for the
    inverse cdf, initial walkers, fresh stream states and the final ``choice``."""
    seeds = np.array([0, 1, 65535, 2**32 - 1, 123456789], dtype=np.int64)
    n = np.array([5, 1, 0, 333, 64], dtype=np.int64)
    nw, around = 7, 0.37
    n_flat = np.array([70, 7, 21, 1400, 140], dtype=np.int64)
    tot = int(n.sum())
    z1, z2, a1, a2, ui = (np.full(tot, np.nan) for _ in range(5))
    init = np.zeros((len(n), nw))
    states = np.zeros((len(n), 2, 625), dtype=np.uint32)
    picks = np.full(tot, -1, dtype=np.int64)
    st = _native.lib().gsx_mode_draws(len(n), _vp(seeds), _vp(n), dim, kind, nw, around, _vp(n_flat), _vp(z1),
                                      _vp(z2), _vp(a1), _vp(a2), _vp(ui), _vp(init), _vp(states), _vp(picks), 3)
    assert st == 0
    off = np.concatenate([[0], np.cumsum(n)])
    for r, seed in enumerate(seeds):
        master = np.random.RandomState(int(seed))
        stream = lambda: np.random.RandomState(master.randint(1, 2**16))
        sl = slice(off[r], off[r + 1])
        assert np.array_equal(z1[sl], stream().normal(size=n[r]))
        assert np.array_equal(z2[sl], stream().normal(size=n[r]))
        if kind == 0:
            continue
        if dim == 1:
            assert np.array_equal(a1[sl], stream().choice([-1, 1], size=n[r]).astype(float))
        else:
            assert np.array_equal(a1[sl], stream().uniform(0.0, 2 * np.pi, n[r]))
            if dim == 3:
                assert np.array_equal(a2[sl], stream().uniform(-1.0, 1.0, n[r]))
        if kind == 1:
            assert np.array_equal(ui[sl], stream().uniform(size=n[r]))
            continue
        assert np.array_equal(init[r], stream().rand(nw) * around)
        for k in range(2):
            assert np.array_equal(states[r, k], _mt_state(stream()))
        assert np.array_equal(picks[sl], stream().choice(int(n_flat[r]), n[r]))


def test_many_groups_on_ready_host():
    """``sample_modes_many(..., on_ready)`` on the host sampler: the callback sees
    every group once (empty groups included), with the modes the call returns."""
    from gstools_amd.sampling import sample_modes_many

    groups = [(Gaussian(dim=2, var=1.0, len_scale=10.0), [300, 50], [1, 2]),  # inversion
              (Exponential(dim=3, var=2.0, len_scale=5.0), [], []),           # empty
              (Matern(dim=3, var=1.0, len_scale=8.0, nu=1.5), [64, 80], [3, 4])]  # MCMC
    seen = {}

    def ready(i, modes):
        assert i not in seen
        seen[i] = [tuple(np.copy(a) for a in m) for m in modes]

    out = sample_modes_many(groups, device=False, on_ready=ready)
    assert sorted(seen) == [0, 1, 2]
    for i, modes in enumerate(out):
        assert len(seen[i]) == len(modes)
        for a, b in zip(seen[i], modes):
            for u, v in zip(a, b):
                assert np.array_equal(u, v)
    # the same modes as one group at a time
    for (model, ns, ss), modes in zip(groups, out):
        ref = sample_modes_batch(model, ns, ss, device=False)
        for a, b in zip(ref, modes):
            for u, v in zip(a, b):
                assert np.array_equal(u, v)
