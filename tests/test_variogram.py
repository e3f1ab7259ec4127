"""Variograms on the CPU: the oracle's variogram kernels and the host front end.

``oracle.vario_*_c`` (C restatement of the external ``unstructured`` /
``directional`` / ``structured`` / ``ma_structured`` kernels,
``src/gstools/variogram/variogram.py:44-113``) is pinned by the reference's
literature known values, by brute force over explicit pair lists (numpy
distances and ``searchsorted`` bins), by the structured kernel agreeing with
the unstructured one on integer grids, by direction tests against explicit
angles, and by planted slips.  The host front end runs the reference's
variogram scenarios (``tests/vario_cases.py``) with the oracle in place of
the device kernels -- test infrastructure only; the product calls
``gsx_vario_*``, compared with this oracle in ``test_gpu_variogram.py``.
"""

import numpy as np
import pytest
import torch

import gstools_amd as gs
import oracle
from gstools_amd import variogram as vmod

import vario_cases as cases

Z = cases.Z_LIT


def test_oracle_known_values():
    x = np.arange(1, 11, dtype=np.double)
    g, c = oracle.vario_unstructured_c(Z, np.arange(1, 11, dtype=np.double), x)
    assert g[0] == pytest.approx(0.4917, abs=5e-5) and g[1] == pytest.approx(0.7625, abs=5e-5)
    assert list(c[:3]) == [9, 8, 7]
    g, _ = oracle.vario_structured_c(Z, estimator_type="c")
    assert g[1] == pytest.approx(1.546 / 2.0, abs=5e-4)
    m = np.zeros(10, bool)
    m[0] = True
    g, c = oracle.vario_structured_c(Z, mask=m)
    assert g[1] == pytest.approx(0.4906, abs=5e-5) and g[2] == pytest.approx(0.7107, abs=5e-5)
    assert list(c[:3]) == [0, 8, 7]


def _brute(f, edges, pos, cressie=False):
    p = pos.shape[1]
    j, k = np.triu_indices(p, 1)
    d = np.sqrt(((pos[:, j] - pos[:, k]) ** 2).sum(0))
    b = np.searchsorted(edges, d, side="right") - 1
    ok = (b >= 0) & (b < len(edges) - 1)
    sums, counts = np.zeros(len(edges) - 1), np.zeros(len(edges) - 1, dtype=np.int64)
    for fm in np.atleast_2d(f):
        diff = fm[k] - fm[j]
        good = ok & ~np.isnan(diff)
        val = np.sqrt(np.abs(diff)) if cressie else diff ** 2
        np.add.at(sums, b[good], val[good])
        np.add.at(counts, b[good], 1)
    n = np.maximum(counts, 1)
    if cressie:
        return 0.5 * (sums / n) ** 4 / (0.457 + 0.494 / n + 0.045 / n ** 2), counts
    return sums / (2 * n), counts


@pytest.mark.parametrize("dim,cressie", [(1, False), (2, True), (3, False)])
def test_oracle_brute_force(dim, cressie):
    rs = np.random.RandomState(dim)
    pos = rs.uniform(0, 10, size=(dim, 300))
    f = rs.normal(size=(2, 300))
    f[1, :7] = np.nan
    edges = np.linspace(0, 6, 9)
    g, c = oracle.vario_unstructured_c(f, edges, pos, "c" if cressie else "m")
    gb, cb = _brute(f, edges, pos, cressie)
    assert np.array_equal(c, cb)
    assert np.allclose(g, gb, rtol=1e-12, atol=0)


def test_oracle_structured_equals_unstructured_on_grid():
    rs = np.random.RandomState(4)
    f = rs.normal(size=(20, 6))
    g_s, c_s = oracle.vario_structured_c(f)
    # along axis 0 of a 20 x 6 grid with unit spacing: one column at a time
    edges = np.arange(20) + 0.5  # bin b holds lag b + 1
    sums = np.zeros(19)
    counts = np.zeros(19, dtype=np.int64)
    for col in range(6):
        g, c = oracle.vario_unstructured_c(f[:, col], edges, np.arange(20.0))
        sums += g * 2 * np.maximum(c, 1)
        counts += c
    assert np.array_equal(counts, c_s[1:])
    assert np.allclose(sums / (2 * np.maximum(counts, 1)), g_s[1:], rtol=1e-13)


def test_oracle_directional_against_explicit_angles():
    rs = np.random.RandomState(8)
    pos = rs.uniform(0, 10, size=(2, 120))
    f = rs.normal(size=120)
    edges = np.linspace(0, 8, 5)
    dirs = np.array([[1.0, 0.0], [np.sqrt(0.5), np.sqrt(0.5)]])
    tol = np.pi / 8
    g, c = oracle.vario_directional_c(f, edges, pos, dirs, tol, -1.0, False)
    j, k = np.triu_indices(120, 1)
    dv = pos[:, k] - pos[:, j]
    d = np.hypot(*dv)
    b = np.searchsorted(edges, d, side="right") - 1
    for di, u in enumerate(dirs):
        ang = np.arccos(np.minimum(np.abs(u @ dv) / d, 1.0))
        sel = (ang < tol) & (b >= 0) & (b < 4)
        assert np.array_equal(c[di], np.bincount(b[sel], minlength=4))
    # separate directions: each pair at most once
    gs_, cs = oracle.vario_directional_c(f, edges, pos, dirs, np.pi / 4, -1.0, True)
    g_all, _ = oracle.vario_unstructured_c(f, edges, pos)
    assert np.all(cs.sum(0) <= oracle.vario_unstructured_c(f, edges, pos)[1])
    del gs_, g_all


def test_planted_slips_are_caught():
    """Inclusive upper edges, or pairs counted in both orders, miss the known counts/values."""
    x = np.arange(1, 11, dtype=np.double)
    edges = np.arange(1, 11, dtype=np.double)
    j, k = np.triu_indices(10, 1)
    d = np.abs(x[k] - x[j])
    incl = np.searchsorted(edges, d, side="left") - 1  # edges[b] < d <= edges[b+1]
    counts_slip = np.bincount(incl[(incl >= 0) & (incl < 9)], minlength=9)
    assert list(counts_slip[:3]) != [9, 8, 7]
    g, c = oracle.vario_unstructured_c(Z, edges, x)
    both = np.bincount(np.concatenate([np.searchsorted(edges, d, side="right") - 1] * 2), minlength=10)[:9]
    assert list(both[:3]) != list(c[:3])


def _sq_threshold(e):
    """Smallest double x with sqrt(x) >= e (the squared-distance thresholds of the pair kernel)."""
    if e <= 0:
        return -1.0
    x = e * e
    while np.sqrt(x) < e:
        x = np.nextafter(x, np.inf)
    while True:
        y = np.nextafter(x, 0.0)
        if y >= 0 and np.sqrt(y) >= e and y < x:
            x = y
        else:
            return x


def test_squared_distance_thresholds_decide_like_sqrt():
    """sqrt(d2) >= e exactly when d2 >= T(e): bins found on squared distances are the reference's."""
    rs = np.random.RandomState(2)
    edges = np.concatenate([rs.uniform(0, 50, 300), np.sqrt(rs.uniform(0, 2500, 300)), [1.0, 2.0, 0.5, 1e-300]])
    for e in edges:
        t = _sq_threshold(e)
        assert np.sqrt(t) >= e and not np.sqrt(np.nextafter(t, 0.0)) >= e
        for x in (t, np.nextafter(t, 0.0), np.nextafter(t, np.inf), e * e):
            assert (np.sqrt(x) >= e) == (x >= t)


@pytest.fixture
def oracle_kernels(monkeypatch):
    """Replace the device kernels by the oracle (host-logic tests only)."""

    def pairs(f, bin_edges, pos, estimator_type, direction=None, angles_tol=np.pi / 8, bandwidth=-1.0,
              separate_dirs=False, haversine=False):
        f = np.atleast_2d(np.asarray(f, dtype=np.double))
        pos = np.atleast_2d(np.asarray(pos, dtype=np.double))
        if direction is None:
            g, c = oracle.vario_unstructured_c(f, bin_edges, pos, estimator_type, "h" if haversine else "e")
        else:
            g, c = oracle.vario_directional_c(f, bin_edges, pos, direction, angles_tol, bandwidth, separate_dirs,
                                              estimator_type)
        return torch.from_numpy(g), torch.from_numpy(c)

    def lags(f, mask, estimator_type):
        g, c = oracle.vario_structured_c(f, mask, estimator_type)
        return torch.from_numpy(g), torch.from_numpy(c)

    monkeypatch.setattr(vmod, "_pairs", pairs)
    monkeypatch.setattr(vmod, "_lags", lags)


def _smooth1d(x, seed):
    rs = np.random.RandomState(seed)
    k, ph = rs.normal(size=30) * 0.2, rs.uniform(0, 2 * np.pi, 30)
    return np.sqrt(2.0 / 30) * np.cos(np.outer(x, k) + ph).sum(1)


def _field3d():
    rs = np.random.RandomState(123456)
    g = np.meshgrid(*[np.arange(10.0)] * 3, indexing="ij")
    k, ph = rs.normal(size=(3, 40)) * 0.3, rs.uniform(0, 2 * np.pi, 40)
    return np.sqrt(2.0 / 40) * np.cos(np.tensordot(np.stack(g, -1), k, 1) + ph).sum(-1)


@pytest.mark.parametrize("scenario", ["known_values", "uncorrelated", "sampling", "assertions",
                                      "direction_assertions", "latlon_scales", "mean_and_trend"])
def test_reference_scenarios(oracle_kernels, scenario):
    getattr(cases, scenario)(gs)


def test_reference_fields(oracle_kernels):
    cases.stacked_and_no_data(gs, _smooth1d)
    cases.direction_axis(gs, _field3d())
    cases.direction_angle(gs, _field3d())
    cases.latlon_auto_binning(gs, _field3d())


def test_standard_bins_and_unsupported(oracle_kernels):
    pos = np.random.RandomState(1).uniform(0, 30, size=(2, 100))
    edges = gs.standard_bins(pos, dim=2)
    diam = np.linalg.norm(pos.max(1) - pos.min(1))
    assert len(edges) == int(np.ceil(2 * np.log2(100) + 1)) + 1 and edges[-1] == pytest.approx(diam / 3)
    bc, g = gs.vario_estimate(pos, np.ones(100))
    assert np.allclose(bc, (edges[:-1] + edges[1:]) / 2) and np.all(g == 0)
    with pytest.raises(NotImplementedError):
        gs.vario_estimate(pos, np.ones(100), normalizer=object())
    with pytest.raises(ValueError):
        gs.variogram.unstructured(np.ones(100), edges, pos, distance_type="x")


def test_standard_bins_latlon():
    """Lat-lon bins (``binning.py:87-104``): points on a sphere of radius ``geo_scale``, the
    box diagonal turned into a great-circle distance, a third of it as the last edge."""
    lat, lon = np.array([10.0, 20.0, 50.0]), np.array([0.0, 30.0, 100.0])
    xyz = np.stack([np.cos(np.radians(lat)) * np.cos(np.radians(lon)),
                    np.cos(np.radians(lat)) * np.sin(np.radians(lon)), np.sin(np.radians(lat))])
    chord = np.linalg.norm(xyz.max(1) - xyz.min(1))
    for scale in (gs.RADIAN_SCALE, gs.KM_SCALE):
        edges = gs.standard_bins((lat, lon), latlon=True, geo_scale=scale, bin_no=4)
        assert len(edges) == 5 and edges[-1] == pytest.approx(2 * scale * np.arcsin(chord / 2) / 3, rel=1e-12)
    # chordal distances beyond the diameter are clipped (geometric.py:728-733)
    assert gs.geometry.chordal_to_great_circle(3.0, 1.0) == pytest.approx(np.pi)


def _great_circle(lat, lon):
    """Pairwise great-circle angles (j < k) by the chordal formula 2 asin(|u_j - u_k| / 2),
    independent of the haversine form."""
    la, lo = np.radians(lat), np.radians(lon)
    u = np.stack([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)])
    j, k = np.triu_indices(len(lat), 1)
    return 2 * np.arcsin(np.minimum(np.linalg.norm(u[:, j] - u[:, k], axis=0) / 2, 1.0)), j, k


def test_haversine_oracle_known_angles():
    """Great-circle angles the geometry fixes: quarter circles, the pole-to-pole half
    circle, 45 and 135 degrees (degrees in, radians out)."""
    pos = np.array([[0.0, 0.0, 90.0, -90.0, 45.0], [0.0, 90.0, 0.0, 10.0, 180.0]])
    edges = np.array([0.0, 0.5, 1.0, 1.6, 2.0, 3.2])
    g, c = oracle.vario_unstructured_c(np.arange(5.0), edges, pos, "m", "h")
    # 45 deg: (2,4); 90 deg: (0,1) (0,2) (0,3) (1,2) (1,3) (1,4); 135 deg: (0,4) (3,4); 180 deg: (2,3)
    assert list(c) == [0, 1, 6, 0, 3]
    assert g[1] == pytest.approx(0.5 * (4 - 2) ** 2 / 1)
    assert g[4] == pytest.approx(0.5 * ((4 - 0) ** 2 + (4 - 3) ** 2 + (3 - 2) ** 2) / 3)


def test_haversine_oracle_matches_chordal_form():
    """The oracle's haversine distances bin every pair like the chordal formula (random
    points over the whole sphere, the dateline and the poles included)."""
    rs = np.random.RandomState(9)
    lat = np.concatenate([rs.uniform(-90, 90, 250), [90.0, -90.0, 0.0, 0.0]])
    lon = np.concatenate([rs.uniform(-180, 180, 250), [0.0, 33.0, 179.9, -179.9]])
    d, j, k = _great_circle(lat, lon)
    edges = np.linspace(0, np.pi, 41)
    c_ref = np.histogram(d, edges)[0]
    g, c = oracle.vario_unstructured_c(np.zeros(len(lat)), edges, np.stack([lat, lon]), "m", "h")
    assert np.array_equal(c[:-1], c_ref[:-1]) and abs(int(c[-1]) - int(c_ref[-1])) <= 1  # pi: last edge
    # a planted slip (lat and lon swapped) does not pass
    _, c_swap = oracle.vario_unstructured_c(np.zeros(len(lat)), edges, np.stack([lon, lat]), "m", "h")
    assert not np.array_equal(c_swap, c)


def test_eval_func_values():
    """``eval_func`` (``tools/misc.py:62-141``): callables evaluated, None is 0, other
    values are constants taken from their first element."""
    from gstools_amd.krige import eval_func

    pos = np.arange(6.0).reshape(2, 3)
    assert np.array_equal(eval_func(lambda a, b: a + b, pos, 2), [3.0, 5.0, 7.0])
    assert eval_func(None, pos, 2) == 0.0
    assert eval_func([2.5, 9.0], pos, 2) == 2.5
    assert np.array_equal(eval_func(1.5, pos, 2, broadcast=True), [1.5] * 3)
