"""Variogram scenarios of the reference tests, shared by the CPU and GPU suites.

Mirrors ``tests/test_variogram_unstructured.py:20-295`` and
``tests/test_variogram_structured.py:16-278`` of the reference: the
literature known values (Matheron and Cressie), uncorrelated fields, point
sampling, stacked fields, no-data values, masks, and directional variograms
(along axes with a narrow band, and from angles vs vectors).  Where the
reference draws its field from an SRF, ``field3d`` is passed in (an SRF
field on the GPU, a seeded smooth field on the CPU): those checks compare
two estimators of the same field, whatever it is.  Each function takes the
module under test (``gstools_amd``).
"""

import numpy as np
import pytest

Z_LIT = np.array((41.2, 40.2, 39.7, 39.2, 40.1, 38.3, 39.1, 40.0, 41.1, 40.3), dtype=np.double)


def known_values(gs):
    x = np.arange(1, 11, 1, dtype=np.double)
    bins = np.arange(1, 11, 1, dtype=np.double)
    _, gamma = gs.vario_estimate([x], Z_LIT, bins)
    assert gamma[0] == pytest.approx(0.4917, abs=5e-5)
    assert gamma[1] == pytest.approx(0.7625, abs=5e-5)
    _, gamma = gs.vario_estimate([np.arange(1, 5)], np.array((10, 20, 30, 40)), np.arange(1, 11))
    assert gamma[0] == pytest.approx(50.0, abs=5e-5)
    _, gamma = gs.vario_estimate([x], list(Z_LIT), bins)
    assert gamma[1] == pytest.approx(0.7625, abs=5e-5)
    gamma = gs.vario_estimate_axis(Z_LIT)
    assert gamma[0] == pytest.approx(0.0, abs=5e-5)
    assert gamma[1] == pytest.approx(0.4917, abs=5e-5)
    assert gamma[2] == pytest.approx(0.7625, abs=5e-5)
    assert gs.vario_estimate_axis(np.array((10, 20, 30, 40)))[1] == pytest.approx(50.0, abs=5e-5)
    assert gs.vario_estimate_axis(list(Z_LIT), estimator="cressie")[1] == pytest.approx(1.546 / 2.0, abs=5e-4)
    gamma = gs.vario_estimate_axis(np.ma.masked_array(Z_LIT, mask=[0] * 10))
    assert gamma[1] == pytest.approx(0.4917, abs=5e-5)
    gamma = gs.vario_estimate_axis(np.ma.masked_array(Z_LIT, mask=[1] + [0] * 9))
    assert gamma[1] == pytest.approx(0.4906, abs=5e-5)
    assert gamma[2] == pytest.approx(0.7107, abs=5e-5)


def uncorrelated(gs):
    var = 1.0 / 12.0
    x_c = np.linspace(0.0, 100.0, 60)
    x, y = np.meshgrid(x_c, x_c)
    field = np.random.RandomState(1479373475).rand(x.size)
    _, gamma = gs.vario_estimate((x.ravel(), y.ravel()), field, np.arange(0, 100, 10))
    for g in (gamma[0], gamma[len(gamma) // 2], gamma[-1]):
        assert g == pytest.approx(var, abs=5e-3)
    c3 = np.linspace(0.0, 100.0, 15)
    x, y, z = np.meshgrid(c3, c3, c3)
    field = np.random.RandomState(1479373475).rand(x.size)
    _, gamma = gs.vario_estimate((x.ravel(), y.ravel(), z.ravel()), field, np.arange(0, 100, 10))
    for g in (gamma[0], gamma[len(gamma) // 2], gamma[-1]):
        assert g == pytest.approx(var, abs=5e-3)
    # structured, along axes
    f2 = np.random.RandomState(1479373475).rand(100, 100)
    for axis in ("x", "y"):
        g = gs.vario_estimate_axis(f2, axis)
        assert g[0] == pytest.approx(0.0, abs=5e-3)
        assert g[len(g) // 2] == pytest.approx(var, abs=5e-3)


def sampling(gs):
    var = 1.0 / 12.0
    x = np.linspace(0.0, 100.0, 21000)
    field = np.random.RandomState(1479373475).rand(len(x))
    _, gamma = gs.vario_estimate([x], field, np.arange(0, 100, 10), sampling_size=5000, sampling_seed=1479373475)
    for g in (gamma[0], gamma[len(gamma) // 2], gamma[-1]):
        assert g == pytest.approx(var, abs=5e-3)
    c = np.linspace(0.0, 100.0, 600)
    x, y = np.meshgrid(c, c)
    field = np.random.RandomState(1479373475).rand(x.size)
    _, gamma = gs.vario_estimate((x.ravel(), y.ravel()), field, np.arange(0, 100, 10), sampling_size=2000,
                                 sampling_seed=1479373475)
    for g in (gamma[0], gamma[len(gamma) // 2], gamma[-1]):
        assert g == pytest.approx(var, abs=5e-3)


def assertions(gs):
    x, x_e = np.arange(0, 10), np.arange(0, 11)
    y, y_e = np.arange(0, 11), np.arange(0, 12)
    z, z_e = np.arange(0, 12), np.arange(0, 15)
    bins = np.arange(0, 3)
    field, field_e = np.arange(0, 10), np.arange(0, 9)
    for args in ([[x_e], field], [(x, y_e), field], [(x, y_e, z), field], [(x, y, z_e), field],
                 [(x_e, y, z), field], [(x, y, z), field_e], [[x], field_e]):
        with pytest.raises(ValueError):
            gs.vario_estimate(*args, bins)
    with pytest.raises(ValueError):
        gs.vario_estimate([x], field, bins, estimator="bla")


def stacked_and_no_data(gs, smooth1d):
    x = np.random.RandomState(19970221).rand(100) * 100.0
    f1, f2 = smooth1d(x, 19970221), smooth1d(x, 20011012)
    bins = np.arange(20) * 2
    _, g1 = gs.vario_estimate(x, f1, bins)
    _, g2 = gs.vario_estimate(x, f2, bins)
    _, g = gs.vario_estimate(x, [f1, f2], bins)
    assert np.allclose(g, 0.5 * (g1 + g2), atol=5e-3)
    x1 = np.random.RandomState(19970221).rand(100) * 100.0
    fn = np.random.RandomState(20011012).rand(100) * 100.0
    fn[:10] = np.nan
    _, ga = gs.vario_estimate(x1, fn, bins)
    _, gb = gs.vario_estimate(x1[10:], fn[10:], bins)
    assert np.allclose(ga, gb, atol=5e-3)


def direction_axis(gs, field3d):
    pos = (range(10), range(10), range(10))
    field = np.ma.array(field3d)
    field.mask = np.abs(field) < 0.1
    bins = range(10)
    _, vario_u = gs.vario_estimate(*(pos, field, bins), direction=((1, 0, 0), (0, 1, 0), (0, 0, 1)),
                                   bandwidth=0.25, mesh_type="structured")
    for k, axis in enumerate("xyz"):
        vario_s = gs.vario_estimate_axis(field, axis)
        assert np.allclose(vario_u[k][: len(bins) - 1], vario_s[: len(bins) - 1], atol=5e-8)


def direction_angle(gs, field3d):
    pos = (range(10), range(10))
    bins = range(0, 10, 2)
    _, v2, c2 = gs.vario_estimate(*(pos, field3d[0], bins), angles=np.pi / 4, mesh_type="structured",
                                  return_counts=True)
    _, v1, c1 = gs.vario_estimate(*(pos, field3d[0], bins), direction=(1, 1), mesh_type="structured",
                                  return_counts=True)
    assert np.allclose(v1, v2, atol=5e-8)
    assert np.array_equal(c1, c2)


def direction_assertions(gs):
    """``tests/test_variogram_unstructured.py:275-300``."""
    pos = [[1, 2, 3], [1, 2, 3]]
    bins = [1, 2]
    fld = np.ma.array([1, 2, 3])
    for kw in ({"direction": [0, 0]}, {"direction": [[[3, 1]]]}, {"direction": [[3, 1, 2]]},
               {"angles": [[[1]]]}, {"angles": [[1, 1]]}, {"direction": [1, 0], "latlon": True}):
        with pytest.raises(ValueError):
            gs.vario_estimate(pos, fld, bins, **kw)


def latlon_auto_binning(gs, field3d):
    """Lat-lon part of ``tests/test_variogram_unstructured.py:362-398`` and the 1D check of
    ``:400-405``: Sturges bins (15 for 100 points) up to a third of the great-circle
    diameter, increasing and positive."""
    pos = (range(10), range(10), range(10))
    bin_center, emp_vario = gs.vario_estimate(pos[:2], field3d[..., 0], mesh_type="structured", latlon=True)
    assert len(bin_center) == 15
    assert np.all(bin_center[1:] > bin_center[:-1]) and np.all(bin_center > 0)
    assert np.all(np.isfinite(emp_vario))
    with pytest.raises(ValueError):
        gs.vario_estimate([[1, 2]], [1, 2], latlon=True)


def latlon_scales(gs):
    """``geo_scale``: bins given in km (or degrees) select the same pairs as the same bins
    in radians; the returned centers stay in the caller's unit (``variogram.py:421-431``)."""
    rs = np.random.RandomState(77)
    lat, lon = rs.uniform(-60, 60, 300), rs.uniform(-170, 170, 300)
    fld = rs.normal(size=300)
    rad = np.linspace(0, 1.5, 12)
    bc_r, g_r, c_r = gs.vario_estimate((lat, lon), fld, rad, latlon=True, return_counts=True)
    for scale in (gs.KM_SCALE, gs.DEGREE_SCALE):
        edges = rad * scale
        kept = edges.copy()
        bc, g, c = gs.vario_estimate((lat, lon), fld, edges, latlon=True, geo_scale=scale, return_counts=True)
        assert np.array_equal(edges, kept)  # the caller's edges are not divided in place
        assert np.allclose(bc, bc_r * scale)
        assert np.array_equal(c, c_r) and np.allclose(g, g_r, rtol=1e-12)


def mean_and_trend(gs):
    """``mean`` / ``trend`` removal (``variogram.py:433-440``): a constant mean leaves the
    variogram unchanged; a callable trend equals removing it from the field first."""
    rs = np.random.RandomState(5)
    x = rs.uniform(0, 50, size=(2, 400))
    fld = rs.normal(size=400) + 0.3 * x[0] - 0.1 * x[1]
    bins = np.linspace(0, 20, 9)
    _, g0, c0 = gs.vario_estimate(x, fld, bins, return_counts=True)
    _, g1, c1 = gs.vario_estimate(x, fld, bins, mean=3.5, return_counts=True)
    assert np.array_equal(c0, c1) and np.allclose(g0, g1, rtol=1e-12)

    def trend(a, b):
        return 0.3 * a - 0.1 * b

    _, gt = gs.vario_estimate(x, fld, bins, trend=trend, mean=[1.0, 2.0])
    _, gm = gs.vario_estimate(x, fld - trend(*x), bins)
    assert np.allclose(gt, gm, rtol=1e-12)
    assert np.all(gt < g0)  # the linear trend inflates the raw variogram
