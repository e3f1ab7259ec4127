"""Kriging scenarios of the reference tests, shared by the CPU and GPU suites.

Mirrors ``tests/test_krige.py:14-288`` and ``tests/test_condition.py:9-110``
of the reference: interpolation at the conditioning points (2 decimals),
structured = unstructured, pseudo-inverse averaging of redundant points,
kriging variance = nugget where the measurement error is 0, the kriged mean,
and conditioned random fields reproducing the conditioning values.  The
normalizer part of ``test_krige_mean`` is not restated (normalizers are not
part of this build).  Each function takes the kriging module to exercise.
"""

import numpy as np
import pytest

DATA = np.array([
    [0.3, 1.2, 0.5, 0.47],
    [1.9, 0.6, 1.0, 0.56],
    [1.1, 3.2, 1.5, 0.74],
    [3.3, 4.4, 2.0, 1.47],
    [4.7, 3.8, 2.5, 1.74],
])
COND_POS = (DATA[:, 0], DATA[:, 1], DATA[:, 2])
COND_VAL = DATA[:, 3]
COND_ERR = np.array([0.01, 0.0, 0.1, 0.05, 0])
MEAN = np.mean(COND_VAL)
DATA_IDX = tuple(np.array(DATA[:, :3] * 10, dtype=int).T)
X, Y, Z = np.linspace(0, 5, 51), np.linspace(0, 6, 61), np.linspace(0, 7, 71)
POS = (X, Y, Z)
GRIDS = [X, np.meshgrid(X, Y, indexing="ij"), np.meshgrid(X, Y, Z, indexing="ij")]
GRID_SHAPE = [51, 61, 71]


def trend(*xyz):
    return xyz[0]


def mean_func(*xyz):
    return 2 * xyz[0]


def _models(gs):
    return [gs.Gaussian, gs.Exponential, gs.Spherical]


def _check(krig, dim, ext=None):
    kw = {} if ext is None else {"ext_drift": ext}
    field_1, _ = krig.unstructured(GRIDS[dim - 1], **kw)
    field_1 = field_1.reshape(GRID_SHAPE[:dim])
    field_2, _ = krig.structured(POS[:dim], **kw)
    assert np.max(np.abs(field_1 - field_2)) == pytest.approx(0.0, abs=5e-3)
    for i, val in enumerate(COND_VAL):
        assert field_2[DATA_IDX[:dim]][i] == pytest.approx(val, abs=5e-3)


def simple(gs, dims=(1, 2, 3)):
    for Model in _models(gs):
        for dim in dims:
            model = Model(dim=dim, var=2, len_scale=2, anis=[0.9, 0.8], angles=[2, 1, 0.5])
            _check(gs.krige.Simple(model, COND_POS[:dim], COND_VAL, MEAN), dim)


def ordinary(gs, dims=(1, 2, 3)):
    for trend_func in [None, trend]:
        for Model in _models(gs):
            for dim in dims:
                model = Model(dim=dim, var=5, len_scale=10, anis=[0.9, 0.8], angles=[2, 1, 0.5])
                _check(gs.krige.Ordinary(model, COND_POS[:dim], COND_VAL, trend=trend_func), dim)


def universal(gs, dims=(1, 2, 3)):
    for drift in ["linear", 0, 1, trend]:
        for Model in _models(gs):
            for dim in dims:
                model = Model(dim=dim, var=2, len_scale=10, anis=[0.9, 0.8], angles=[2, 1, 0.5])
                _check(gs.krige.Universal(model, COND_POS[:dim], COND_VAL, drift), dim)


def detrended(gs, dims=(1, 2, 3)):
    for Model in _models(gs):
        for dim in dims:
            model = Model(dim=dim, var=2, len_scale=10, anis=[0.5, 0.2], angles=[0.4, 0.2, 0.1])
            _check(gs.krige.Detrended(model, COND_POS[:dim], COND_VAL, trend), dim)


def extdrift(gs, dims=(1, 2, 3)):
    """External drift: a smooth field sampled on the grid (the reference draws an SRF)."""
    for dim in dims:
        grid = np.asarray(GRIDS[dim - 1], dtype=np.double).reshape(dim, -1)
        drift_field = np.sin(0.7 * grid[0]) + 0.3 * np.sum(grid, axis=0)
        cond_drift = drift_field.reshape(GRID_SHAPE[:dim])[DATA_IDX[:dim]]
        for Model in _models(gs):
            model = Model(dim=dim, var=2, len_scale=10, anis=[0.5, 0.2], angles=[0.4, 0.2, 0.1])
            krig = gs.krige.ExtDrift(model, COND_POS[:dim], COND_VAL, cond_drift)
            _check(krig, dim, ext=drift_field)


def pseudo(gs, dims=(1, 2, 3)):
    p_data = np.zeros((3, 3))
    p_vals = np.array([1.0, 2.0, 6.0])
    for Model in _models(gs):
        for dim in dims:
            model = Model(dim=dim, var=2, len_scale=10, anis=[0.5, 0.2], angles=[0.4, 0.2, 0.1])
            krig = gs.krige.Krige(model, p_data[:dim], p_vals, unbiased=False)
            field, _ = krig([0, 0, 0][:dim])
            assert field[0] == pytest.approx(np.mean(p_vals), abs=5e-3)


def error(gs, dims=(1, 2, 3)):
    for Model in _models(gs):
        for dim in dims:
            model = Model(dim=dim, var=5, len_scale=10, nugget=0.1, anis=[0.9, 0.8], angles=[2, 1, 0.5])
            ordinary = gs.krige.Ordinary(model, COND_POS[:dim], COND_VAL, exact=False, cond_err=COND_ERR)
            _, err = ordinary(COND_POS[:dim])
            assert err[1] == pytest.approx(model.nugget, abs=5e-3)
            assert err[4] == pytest.approx(model.nugget, abs=5e-3)


def krige_mean(gs):
    krige = gs.krige.Simple(gs.Gaussian(dim=3), COND_POS, COND_VAL)
    assert np.all(np.isclose(krige.structured(POS, only_mean=True), 0))
    krige = gs.krige.Simple(gs.Gaussian(dim=3), COND_POS, COND_VAL, mean=mean_func, trend=trend)
    mean_f = krige.structured(POS, only_mean=True)
    grid = np.meshgrid(*POS, indexing="ij")
    assert np.allclose(mean_f, mean_func(*grid) + trend(*grid))
    krige = gs.krige.Ordinary(gs.Gaussian(dim=3), COND_POS, COND_VAL)
    assert np.all(np.isclose(krige.structured(POS, only_mean=True), krige.get_mean()))


def condition(gs, kinds=("Simple", "Ordinary"), dims=(1, 2, 3)):
    """``tests/test_condition.py:31-95``."""
    grid = np.linspace(5, 20, 10)
    pos = tuple(np.concatenate((c, grid)) for c in COND_POS)
    for kind in kinds:
        for Model in (gs.Gaussian, gs.Exponential):
            for dim in dims:
                model = Model(dim=dim, var=0.5, len_scale=2, anis=[0.1, 1], angles=[0.5, 0, 0])
                if kind == "Simple":
                    krige = gs.krige.Simple(model, COND_POS[:dim], COND_VAL, MEAN)
                else:
                    krige = gs.krige.Ordinary(model, COND_POS[:dim], COND_VAL)
                crf = gs.CondSRF(krige, seed=19970221)
                field_1 = crf.unstructured(pos[:dim])
                field_2 = crf.structured(pos[:dim])
                if dim > 1:
                    raw = np.copy(crf["raw_krige"])
                    crf(seed=19970222)
                    assert np.allclose(raw, crf["raw_krige"])
                for i, val in enumerate(COND_VAL):
                    assert field_1[i] == pytest.approx(val, abs=5e-3)
                    assert field_2[dim * (i,)] == pytest.approx(val, abs=5e-3)


def condition_nugget(gs):
    grid = np.linspace(5, 20, 10)
    pos = tuple(np.concatenate((c, grid)) for c in COND_POS)
    model = gs.Gaussian(dim=3, nugget=0.01, var=0.5, len_scale=2, anis=[0.1, 1], angles=[0.5, 0, 0])
    krige = gs.krige.Ordinary(model, COND_POS, COND_VAL, exact=True)
    crf = gs.CondSRF(krige, seed=19970221)
    field_1 = crf.unstructured(pos)
    field_2 = crf.structured(pos)
    for i, val in enumerate(COND_VAL):
        assert field_1[i] == pytest.approx(val, abs=5e-3)
        assert field_2[3 * (i,)] == pytest.approx(val, abs=5e-3)
