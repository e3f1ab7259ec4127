"""Kriging on the CPU: the oracle's kriging sums and the host kriging front end.

``oracle.krige_c`` (the C restatement of ``calc_field_krige(_and_variance)``,
``src/gstools/krige/base.py:42-61``) is pinned by matrix identities computed
independently by numpy, by simple kriging's interpolation property, and by a
planted slip.  The host front end (``gstools_amd.krige``) runs the
reference's kriging scenarios (``tests/krige_cases.py``) with the oracle in
place of the device sums -- test infrastructure only; the product calls
``gsx_krige_f64``, compared with this oracle in ``test_gpu_krige.py``.
"""

import numpy as np
import pytest
import torch

import gstools_amd as gs
import oracle
from gstools_amd import krige as krige_mod

import krige_cases as cases


def _identity_ok(f, e, m, v, c):
    w = m @ v
    scale = np.abs(m) @ np.abs(v)
    ok_f = np.all(np.abs(f - c @ w) <= 1e-13 * (np.abs(c) @ scale))
    ok_e = np.all(np.abs(e - np.sum(v * w, axis=0)) <= 1e-13 * np.sum(np.abs(v) * scale, axis=0))
    return bool(ok_f and ok_e)


@pytest.mark.parametrize("k,p", [(1, 5), (7, 13), (64, 33), (130, 300)])
def test_oracle_matrix_identities(k, p):
    rs = np.random.RandomState(k + p)
    m, v, c = rs.normal(size=(k, k)), rs.normal(size=(k, p)), rs.normal(size=k)
    f, e = oracle.krige_c(m, v, c)
    assert _identity_ok(f, e, m, v, c)
    assert np.array_equal(oracle.krige_c(m, v, c, variance=False), f)


def test_oracle_simple_kriging_interpolates():
    """M = C^-1 and v = columns of C: field = cond, error = diag(C)."""
    rs = np.random.RandomState(3)
    pts = rs.uniform(0, 10, size=(2, 12))
    d = np.sqrt(((pts[:, :, None] - pts[:, None, :]) ** 2).sum(0))
    cov = 1.5 * np.exp(-d / 3.0)
    vals = rs.normal(size=12)
    f, e = oracle.krige_c(np.linalg.inv(cov), cov, vals)
    assert np.allclose(f, vals, atol=1e-10)
    assert np.allclose(e, np.diag(cov), atol=1e-10)


def test_planted_slips_are_caught():
    """A transposed kriging matrix, or the variance term weighted by cond, fails the identities."""
    rs = np.random.RandomState(9)
    m, v, c = rs.normal(size=(9, 9)), rs.normal(size=(9, 20)), rs.normal(size=9)

    def loop(mat, weight_e_by_cond=False):
        f, e = np.zeros(20), np.zeros(20)
        for k in range(20):
            for i in range(9):
                w = sum(mat[i, j] * v[j, k] for j in range(9))
                f[k] += c[i] * w
                e[k] += (c[i] if weight_e_by_cond else v[i, k]) * w
        return f, e

    assert _identity_ok(*loop(m), m, v, c)
    assert not _identity_ok(*loop(m.T), m, v, c)
    assert not _identity_ok(*loop(m, True), m, v, c)


@pytest.fixture
def oracle_sums(monkeypatch):
    """Replace the device kriging sums by the oracle (host-logic tests only)."""

    def fake(krig_mat, krig_vecs, cond, variance=True):
        res = oracle.krige_c(np.asarray(krig_mat), np.asarray(krig_vecs), np.asarray(cond), variance)
        if variance:
            return torch.from_numpy(res[0]), torch.from_numpy(res[1])
        return torch.from_numpy(res)

    monkeypatch.setattr(krige_mod, "krige_sum_device", fake)
    monkeypatch.setattr(krige_mod, "_device_cov_kind", lambda model: None)  # host right-hand sides


@pytest.mark.parametrize("scenario", ["simple", "ordinary", "universal", "detrended", "extdrift", "pseudo",
                                      "error"])
def test_reference_scenarios(oracle_sums, scenario):
    getattr(cases, scenario)(gs, dims=(1, 2))


def test_reference_scenarios_3d(oracle_sums):
    cases.simple(gs, dims=(3,))
    cases.pseudo(gs, dims=(3,))


def test_krige_mean(oracle_sums):
    cases.krige_mean(gs)


def test_chunking_and_errors(oracle_sums):
    model = gs.Exponential(dim=2, var=2, len_scale=3)
    k = gs.krige.Ordinary(model, cases.COND_POS[:2], cases.COND_VAL)
    full, var_full = k.structured(cases.POS[:2])
    part, var_part = k.structured(cases.POS[:2], chunk_size=97)
    assert np.array_equal(full, part) and np.array_equal(var_full, var_part)
    assert k["field"] is part and k.field_names == ["field", "krige_var"]
    with pytest.raises(ValueError):
        gs.krige.Krige(gs.Stable(dim=3), None, None)
    with pytest.raises(NotImplementedError):
        gs.krige.Ordinary(model, cases.COND_POS[:2], cases.COND_VAL, normalizer=object())
    with pytest.raises(ValueError):
        gs.krige.Ordinary(model, cases.COND_POS[:2], cases.COND_VAL, exact=True, cond_err=0.1)
    with pytest.raises(ValueError):
        gs.krige.Krige(model, cases.COND_POS[:2], cases.COND_VAL, pseudo_inv_type="lu")
    with pytest.raises(ValueError):
        gs.CondSRF(model)
    with pytest.raises(ValueError):
        gs.krige.calc_field_krige(np.eye(3, dtype=np.float32), np.ones((3, 4)), np.ones(3))
