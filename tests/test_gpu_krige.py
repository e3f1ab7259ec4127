"""GPU kriging: ``gsx_krige_f64`` against the oracle, and the reference's
kriging and conditioning scenarios with the real kernel.

Bar: every field and variance term within 1e-13 of the oracle relative to
the absolute sum of its terms (``|c| |M| |v|`` and ``|v| |M| |v|``): the GPU
sums in a different order (matrix-core tiles), so differences are rounding
of the same terms.  Reference scenarios: ``tests/krige_cases.py``.
"""

import ctypes
import sys

import numpy as np
import pytest
import torch

import gstools_amd as gs
import oracle
from gstools_amd import _native
from gstools_amd.krige import calc_field_krige, calc_field_krige_and_variance, krige_sum_device

import krige_cases as cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bounds(m, v, c):
    scale = np.abs(m) @ np.abs(v)
    return np.abs(c) @ scale, np.sum(np.abs(v) * scale, axis=0)


@pytest.mark.parametrize("k,p", [(1, 1), (5, 7), (64, 128), (65, 129), (33, 130), (130, 1000), (256, 258), (600, 3001),
                                 (17, 0)])
def test_kernel_matches_oracle(k, p):
    _say(f"krige K={k} P={p}")
    rs = np.random.RandomState(k * 7 + p)
    m, v, c = rs.normal(size=(k, k)), rs.normal(size=(k, p)), rs.normal(size=k)
    f_ref, e_ref = oracle.krige_c(m, v, c)
    f, e = calc_field_krige_and_variance(m, v, c)
    bf, be = _bounds(m, v, c)
    assert f.shape == (p,) and e.shape == (p,)
    assert np.all(np.abs(f - f_ref) <= 1e-13 * bf + 1e-300)
    assert np.all(np.abs(e - e_ref) <= 1e-13 * be + 1e-300)
    f2 = calc_field_krige(m, v, c)
    assert np.all(np.abs(f2 - f_ref) <= 1e-13 * bf + 1e-300)
    f3, e3 = calc_field_krige_and_variance(m, v, c)
    assert np.array_equal(f, f3) and np.array_equal(e, e3)  # fixed reduction order


def test_empty_conditions_and_errors():
    _say("krige empty / errors")
    f, e = calc_field_krige_and_variance(np.zeros((0, 0)), np.zeros((0, 6)), np.zeros(0))
    assert np.array_equal(f, np.zeros(6)) and np.array_equal(e, np.zeros(6))
    with pytest.raises(ValueError):
        calc_field_krige(np.ones((3, 3)), np.ones((4, 5)), np.ones(3))
    lib = _native.lib()
    dev = torch.device("cuda")
    m = torch.eye(8, dtype=torch.float64, device=dev)
    v = torch.ones((8, 300), dtype=torch.float64, device=dev)
    c = torch.ones(8, dtype=torch.float64, device=dev)
    out = torch.empty(300, dtype=torch.float64, device=dev)
    ws_need = lib.gsx_krige_workspace(8, 300, 1)
    ws = torch.empty(ws_need, dtype=torch.uint8, device=dev)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    assert lib.gsx_krige_f64(8, 300, p(m), p(v), p(c), p(out), p(out), p(ws), ws_need - 8, None) == -4
    assert lib.gsx_krige_f64(8, 300, p(m), p(v), p(c), None, None, p(ws), ws_need, None) == -1
    assert lib.gsx_krige_f64(-1, 300, p(m), p(v), p(c), p(out), None, p(ws), ws_need, None) == -1
    torch.cuda.synchronize()


def test_kriging_system_sizes():
    """A covariance-built system with 900 conditioning points onto 20000 targets."""
    _say("krige 900 x 20000")
    model = gs.Exponential(dim=2, var=1.3, len_scale=7.0)
    rs = np.random.RandomState(2)
    cond_pos = rs.uniform(0, 100, size=(2, 900))
    cond_val = rs.normal(size=900)
    krig = gs.krige.Ordinary(model, cond_pos, cond_val)
    pos = rs.uniform(0, 100, size=(2, 20000))
    field, var = krig(pos)
    system = krig.system
    vecs = system.rhs_host(model.isometrize(pos), krig._constraint_rows(pos, np.empty((0, pos.shape[1]))),
                           krig.exact)
    f_ref, e_ref = oracle.krige_c(system.inverse, vecs, krig._residual())
    bf, be = _bounds(system.inverse, vecs, krig._residual())
    assert np.all(np.abs(field - f_ref) <= 1e-13 * bf)
    assert np.all(np.abs((model.sill - var) - e_ref) <= 1e-13 * be + 1e-12 * model.sill)


@pytest.mark.parametrize("exact", [False, True])
def test_device_covariance_rows(exact):
    """``gsx_krige_cov_f64`` against the host ``covariance`` / ``cov_nugget`` on scipy's cdist."""
    from scipy.spatial.distance import cdist

    from gstools_amd.krige import krige_cov_device

    _say(f"krige covariance rows exact={exact}")
    rs = np.random.RandomState(11)
    for model in (gs.Gaussian(dim=3, var=1.7, len_scale=4.0, nugget=0.2), gs.Exponential(dim=2, var=2.0, len_scale=3.0),
                  gs.Stable(dim=3, var=0.8, len_scale=5.0, alpha=1.3), gs.Spherical(dim=1, var=1.1, len_scale=6.0)):
        d = model.dim
        cpos = rs.uniform(0, 10, size=(d, 37))
        pos = np.concatenate([rs.uniform(0, 10, size=(d, 500)), cpos[:, :5]], axis=1)  # includes r = 0
        got = krige_cov_device(model, cpos, pos, exact=exact).cpu().numpy()
        fn = model.cov_nugget if exact else model.covariance
        ref = fn(cdist(cpos.T, pos.T))
        # absolute bar: the Spherical polynomial cancels to 0 at the range
        assert np.max(np.abs(got - ref)) <= 8 * 2.3e-16 * model.sill, type(model).__name__
    assert krige_cov_device(gs.Matern(dim=2), cpos[:2], pos[:2]) is None  # host path for other models


@pytest.mark.parametrize("scenario", ["simple", "ordinary", "universal", "detrended", "extdrift", "pseudo",
                                      "error"])
def test_reference_krige_scenarios(scenario):
    _say(f"krige scenario {scenario}")
    getattr(cases, scenario)(gs)


def test_reference_krige_mean():
    _say("krige mean")
    cases.krige_mean(gs)


def test_reference_condition():
    _say("condsrf simple/ordinary")
    cases.condition(gs)


def test_reference_condition_nugget():
    _say("condsrf nugget")
    cases.condition_nugget(gs)


def test_condsrf_composition():
    """field = krige + sqrt(krige_var / var) * unconditioned field (``cond_srf.py:136-153``)."""
    _say("condsrf composition")
    model = gs.Gaussian(dim=2, var=0.5, len_scale=2)
    krige = gs.krige.Ordinary(model, cases.COND_POS[:2], cases.COND_VAL)
    crf = gs.CondSRF(krige, seed=7)
    pos = (np.linspace(0, 6, 40), np.linspace(0, 5, 30))
    field = crf.structured(pos)
    kr, kv = krige.structured(pos, post_process=False)
    rm = gs.RandMeth(model, seed=7)
    raw = rm(model.isometrize(gs.geometry.generate_grid(pos)), add_nugget=False).reshape(40, 30)
    assert np.allclose(crf["raw_krige"], kr, atol=1e-12)
    assert np.allclose(crf["raw_field"], raw, atol=1e-12)
    assert np.allclose(field, kr + np.sqrt(kv / model.var) * raw, atol=1e-12)
