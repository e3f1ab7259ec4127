"""GPU variograms: ``gsx_vario_*`` against the oracle, and the reference's
variogram scenarios with the real kernels.

Bars: counts identical to the oracle's (integers; the distances carry the
reference's rounding, so every pair lands in the same bin); estimates within
1e-12 relative (the sums are the same terms added in another order).
Covered: 1-3 dimensions, stacked fields beyond one launch's field chunk,
NaN values and positions, non-monotone edges, more (direction, bin) slots
than one launch holds, directional tolerance / bandwidth / separated
directions, Cressie, structured and masked lags, great-circle distances
(lat-lon), mean / trend removal.
"""

import sys

import numpy as np
import pytest

import gstools_amd as gs
import oracle
from gstools_amd import variogram as v

import vario_cases as cases

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _close(g, g_ref):
    assert np.allclose(g, g_ref, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("dim,p,nf,est", [(1, 1000, 1, "m"), (2, 2500, 3, "c"), (3, 3001, 11, "m"), (2, 2, 1, "m"),
                                          (3, 1, 1, "m")])
def test_unstructured_matches_oracle(dim, p, nf, est):
    _say(f"vario unstructured dim={dim} P={p} F={nf} est={est}")
    rs = np.random.RandomState(p + dim)
    pos = rs.uniform(0, 20, size=(dim, p))
    f = rs.normal(size=(nf, p))
    f[0, :5] = np.nan
    edges = np.linspace(0, 12, 15)
    g, c = v.unstructured(f, edges, pos, est)
    g_ref, c_ref = oracle.vario_unstructured_c(f, edges, pos, est)
    assert np.array_equal(c, c_ref)
    _close(g, g_ref)
    g2, c2 = v.unstructured(f, edges, pos, est)
    assert np.array_equal(g, g2) and np.array_equal(c, c2)  # fixed reduction order


def test_edges_nan_positions_and_many_bins():
    _say("vario odd edges / nan positions / many bins")
    rs = np.random.RandomState(3)
    pos = rs.uniform(0, 10, size=(2, 700))
    pos[0, 3] = np.nan  # its pairs have a NaN distance: counted in every bin
    f = rs.normal(size=700)
    for edges in (np.array([0.0, 2.0, 1.0, 5.0, 4.0, 9.0]), np.linspace(0, 14, 400), np.array([1.0, 1.0, 3.0])):
        g, c = v.unstructured(f, edges, pos)
        g_ref, c_ref = oracle.vario_unstructured_c(f, edges, pos)
        assert np.array_equal(c, c_ref)
        _close(g, g_ref)


def test_infinite_positions_and_values():
    """Infinite coordinates and field values: inf - inf is a NaN distance (every bin), one infinite
    coordinate an infinite distance (no bin); an infinite field value makes its bins' sums inf or NaN."""
    _say("vario infinite positions / values")
    rs = np.random.RandomState(21)
    pos = rs.uniform(0, 10, size=(3, 900))
    pos[1, 7] = np.inf
    pos[1, 300] = np.inf  # with point 7: a NaN distance
    pos[2, 650] = -np.inf
    f = rs.normal(size=900)
    for fv in (f, np.where(np.arange(900) == 500, np.inf, f)):
        for edges in (np.linspace(0, 12, 15), np.array([0.0, 1.0, 4.0, 4.5, 11.0])):
            g, c = v.unstructured(fv, edges, pos)
            g_ref, c_ref = oracle.vario_unstructured_c(fv, edges, pos)
            assert np.array_equal(c, c_ref)
            np.testing.assert_allclose(g, g_ref, rtol=1e-12, atol=1e-300, equal_nan=True)


def test_edges_on_pair_distances():
    """Edges placed exactly on pair distances (the reference's rounding): those pairs go to the upper bin."""
    _say("vario edges on pair distances")
    rs = np.random.RandomState(12)
    pos = np.round(rs.uniform(0, 10, size=(3, 400)), 3)
    f = rs.normal(size=400)
    j, k = np.triu_indices(400, 1)
    d = np.sqrt(((pos[:, j] - pos[:, k]) ** 2).sum(0))
    edges = np.unique(np.concatenate([[0.0], rs.choice(d, 30, replace=False), [20.0]]))
    for e in (edges, np.linspace(0.0, d[17], 25), np.concatenate([[d[5]], d[5] + np.arange(1, 12) * 0.3])):
        g, c = v.unstructured(f, e, pos)
        g_ref, c_ref = oracle.vario_unstructured_c(f, e, pos)
        assert np.array_equal(c, c_ref)
        _close(g, g_ref)


@pytest.mark.parametrize("separate,bandwidth", [(False, -1.0), (True, -1.0), (False, 1.5), (True, 0.8)])
def test_directional_matches_oracle(separate, bandwidth):
    _say(f"vario directional separate={separate} bandwidth={bandwidth}")
    rs = np.random.RandomState(5)
    pos = rs.uniform(0, 10, size=(3, 1500))
    f = rs.normal(size=(2, 1500))
    dirs = rs.normal(size=(4, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    edges = np.linspace(0, 9, 10)
    g, c = v.directional(f, edges, pos, dirs, np.pi / 6, bandwidth, separate, "m")
    g_ref, c_ref = oracle.vario_directional_c(f, edges, pos, dirs, np.pi / 6, bandwidth, separate, "m")
    assert np.array_equal(c, c_ref)
    _close(g, g_ref)


@pytest.mark.parametrize("separate,bandwidth", [(False, -1.0), (True, 1.2), (False, 0.7)])
def test_directional_one_field_matches_oracle(separate, bandwidth):
    """One field (the straight-line directional loop): uniform and non-uniform edges,
    with NaN values in some waves (those take the general loop), duplicate points."""
    _say(f"vario directional one field separate={separate} bandwidth={bandwidth}")
    rs = np.random.RandomState(17)
    pos = rs.uniform(0, 10, size=(3, 2000))
    pos[:, 5] = pos[:, 6]  # a distance-0 pair
    f = rs.normal(size=2000)
    f[1500:1503] = np.nan
    dirs = rs.normal(size=(3, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    for edges in (np.linspace(0, 9, 13), np.array([0.0, 0.5, 0.6, 2.0, 4.5, 8.0])):
        for est in ("m", "c"):
            g, c = v.directional(f, edges, pos, dirs, np.pi / 7, bandwidth, separate, est)
            g_ref, c_ref = oracle.vario_directional_c(f, edges, pos, dirs, np.pi / 7, bandwidth, separate, est)
            assert np.array_equal(c, c_ref)
            _close(g, g_ref)


def test_directional_slot_chunks():
    """(direction, bin) slots beyond one launch's LDS: launches over slot ranges."""
    _say("vario directional many slots")
    rs = np.random.RandomState(6)
    pos = rs.uniform(0, 10, size=(2, 900))
    f = rs.normal(size=900)
    dirs = np.stack([np.cos(np.linspace(0, np.pi, 9)), np.sin(np.linspace(0, np.pi, 9))], 1)
    edges = np.linspace(0, 12, 61)
    for separate in (False, True):
        g, c = v.directional(f, edges, pos, dirs, np.pi / 16, -1.0, separate, "c")
        g_ref, c_ref = oracle.vario_directional_c(f, edges, pos, dirs, np.pi / 16, -1.0, separate, "c")
        assert np.array_equal(c, c_ref)
        _close(g, g_ref)


@pytest.mark.parametrize("n0,n1,masked,est", [(10, 1, False, "m"), (300, 5000, False, "c"), (257, 1031, True, "m"),
                                             (1000, 3, True, "c"), (777, 1, True, "m"), (64, 200, False, "m"),
                                             (2000, 64, False, "m"), (1100, 130, True, "c"), (33, 65, True, "m")])
def test_structured_matches_oracle(n0, n1, masked, est):
    _say(f"vario structured {n0}x{n1} masked={masked}")
    rs = np.random.RandomState(n0)
    f = rs.normal(size=(n0, n1))
    mask = rs.uniform(size=(n0, n1)) < 0.1 if masked else None
    g = v.ma_structured(f, mask, est) if masked else v.structured(f, est)
    g_ref, c_ref = oracle.vario_structured_c(f, mask, est)
    _close(g, g_ref)
    g2, c2 = v._lags(f, mask, est)
    assert np.array_equal(c2.cpu().numpy(), c_ref)


@pytest.mark.parametrize("n0,n1,masked", [(100_000, 1, False), (70_001, 2, True)])
def test_structured_long_axis(n0, n1, masked):
    """More lags than a grid dimension's 65535 blocks, on a narrow array:
    counts by closed form (n0 - k unmasked pairs per column, or the unmasked
    pairs of the mask), sums against numpy on a spread of lags."""
    _say(f"vario structured long axis {n0}x{n1} masked={masked}")
    rs = np.random.RandomState(n0)
    f = rs.normal(size=(n0, n1))
    mask = rs.uniform(size=(n0, n1)) < 0.05 if masked else None
    gamma, counts = v._lags(f, mask, "m")
    gamma, counts = gamma.cpu().numpy(), counts.cpu().numpy()
    lags = np.unique(np.concatenate([np.arange(0, 40), rs.randint(1, n0, 60), [65534, 65535, 65536, 65537],
                                     [n0 - 2, n0 - 1]]))
    lags = lags[lags < n0]
    ok = np.ones((n0, n1), dtype=bool) if mask is None else ~mask
    for k in lags:
        if k == 0:
            assert counts[0] == 0
            continue
        both = ok[k:] & ok[:-k]
        diff = (f[k:] - f[:-k])[both]
        assert counts[k] == both.sum(), k
        ref = np.sum(diff**2) / (2.0 * max(both.sum(), 1))
        assert np.isclose(gamma[k], ref, rtol=1e-11, atol=1e-300), (k, gamma[k], ref)
    if mask is None:
        assert np.array_equal(counts[1:], (n0 - np.arange(1, n0)) * n1)


@pytest.mark.parametrize("nf,est", [(1, "m"), (3, "c"), (9, "m")])
def test_haversine_matches_oracle(nf, est):
    """Great-circle pair binning (``distance_type`` "h") against the oracle: points over the
    whole sphere with the poles, the dateline, duplicates and a NaN latitude (NaN distances
    count in every bin); more fields than one launch stages."""
    _say(f"vario haversine F={nf} est={est}")
    rs = np.random.RandomState(31 + nf)
    p = 1500
    lat, lon = rs.uniform(-90, 90, p), rs.uniform(-180, 180, p)
    lat[:4], lon[:4] = [90.0, -90.0, 12.5, 12.5], [0.0, 45.0, 179.99, -179.99]
    lat[4], lon[4] = lat[5], lon[5]  # a distance-0 pair
    lat[7] = np.nan
    pos = np.stack([lat, lon])
    f = rs.normal(size=(nf, p))
    f[0, 10:13] = np.nan
    for edges in (np.linspace(0, np.pi, 31), np.array([0.0, 0.1, 0.15, 1.0, 2.5, 3.2]), np.array([1.0, 0.5, 2.0])):
        g, c = v.unstructured(f, edges, pos, est, distance_type="h")
        g_ref, c_ref = oracle.vario_unstructured_c(f, edges, pos, est, "h")
        assert np.array_equal(c, c_ref)
        _close(g, g_ref)


@pytest.mark.parametrize("scenario", ["known_values", "uncorrelated", "sampling", "assertions", "direction_assertions",
                                      "latlon_scales", "mean_and_trend"])
def test_reference_scenarios(scenario):
    _say(f"vario scenario {scenario}")
    getattr(cases, scenario)(gs)


def test_reference_fields_from_srf():
    """``tests/test_variogram_unstructured.py:13-17``: the 10^3 Exponential SRF field, seed 123456."""
    _say("vario srf fields")
    model = gs.Exponential(dim=3, len_scale=[12, 6, 3])
    srf = gs.SRF(model, seed=123456)
    field = srf((range(10), range(10), range(10)), mesh_type="structured")

    def smooth1d(x, seed):
        return gs.SRF(gs.Exponential(dim=1, var=2, len_scale=10))(x, seed=seed)

    cases.stacked_and_no_data(gs, smooth1d)
    cases.direction_axis(gs, field)
    cases.direction_angle(gs, field)
    cases.latlon_auto_binning(gs, field)
