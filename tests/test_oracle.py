"""Pin the oracle (CPU restatements of the reference summation).

Each check is chosen so that a plausible slip -- a dropped sin term, a sign
error, transposed (d, N) / (d, P) indices, a wrong projector -- fails it:
mpmath brute force, symmetries of the formula in
``src/gstools/field/generator.py:193-199`` / ``:479-494``, the tensor-grid
GEMM identity, divergence-freeness, and finite differences for the adjoint.
The last test plants such slips in a copy of the oracle and shows the
checks catch them.
"""

import numpy as np
import pytest

import oracle


def _case(d, n, p, seed, kscale=1.0, xscale=10.0):
    rs = np.random.RandomState(seed)
    cov = rs.normal(size=(d, n)) * kscale
    z1 = rs.normal(size=n)
    z2 = rs.normal(size=n)
    pos = rs.uniform(-xscale, xscale, size=(d, p))
    return cov, z1, z2, pos


@pytest.mark.parametrize("d", [1, 2, 3])
def test_c_matches_mpmath(d):
    cov, z1, z2, pos = _case(d, 17, 23, 100 + d)
    ref = oracle.summate_mp(cov, z1, z2, pos)
    got = oracle.summate_c(cov, z1, z2, pos)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.sum(np.hypot(z1, z2))


@pytest.mark.parametrize("d", [2, 3])
def test_incompr_c_matches_mpmath(d):
    cov, z1, z2, pos = _case(d, 11, 13, 200 + d)
    ref = oracle.summate_incompr_mp(cov, z1, z2, pos)
    got = oracle.summate_incompr_c(cov, z1, z2, pos)
    assert got.shape == (d, 13)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.sum(np.hypot(z1, z2))


@pytest.mark.parametrize("d", [1, 2, 3])
def test_c_and_numpy_agree(d):
    cov, z1, z2, pos = _case(d, 300, 500, 300 + d, kscale=3.0, xscale=1e3)
    a = oracle.summate_c(cov, z1, z2, pos)
    b = oracle.summate_np(cov, z1, z2, pos)
    rel_max, _ = oracle.rel_errors(a, b)
    assert rel_max < 1e-13
    ai = oracle.summate_incompr_c(cov, z1, z2, pos) if d > 1 else None
    if ai is not None:
        bi = oracle.summate_incompr_np(cov, z1, z2, pos)
        assert oracle.rel_errors(ai, bi)[0] < 1e-13


def test_empty_and_zero_modes():
    cov, z1, z2, pos = _case(2, 0, 5, 1)
    assert np.array_equal(oracle.summate_c(cov, z1, z2, pos), np.zeros(5))
    cov, z1, z2, pos = _case(2, 4, 0, 1)
    assert oracle.summate_c(cov, z1, z2, pos).shape == (0,)
    cov, z1, z2, pos = _case(3, 9, 7, 2)
    cov[:] = 0.0  # k = 0: every term is z1
    assert np.allclose(oracle.summate_c(cov, z1, z2, pos), np.sum(z1), rtol=1e-14, atol=0)


def test_linearity_and_symmetry():
    cov, z1, z2, pos = _case(3, 40, 60, 3)
    f = lambda a, b: oracle.summate_c(cov, a, b, pos)
    u = f(z1, z2)
    assert np.allclose(f(2.5 * z1, 2.5 * z2), 2.5 * u, rtol=1e-13, atol=1e-13)
    assert np.allclose(f(z1, 0 * z2) + f(0 * z1, z2), u, rtol=1e-13, atol=1e-12)
    # (k, z1, z2) -> (-k, z1, -z2) leaves every term unchanged
    assert np.allclose(oracle.summate_c(-cov, z1, -z2, pos), u, rtol=1e-13, atol=1e-12)
    # sin term has the sign of k: flipping k alone changes u unless z2 = 0
    assert not np.allclose(oracle.summate_c(-cov, z1, z2, pos), u)


def test_translation_rotates_amplitudes():
    cov, z1, z2, pos = _case(2, 30, 40, 4)
    t = np.array([[1.7], [-0.6]])
    kt = (cov * t).sum(axis=0)
    # z1 cos(k.(x+t)) + z2 sin(k.(x+t)) = z1' cos(k.x) + z2' sin(k.x)
    z1r = z1 * np.cos(kt) + z2 * np.sin(kt)
    z2r = z2 * np.cos(kt) - z1 * np.sin(kt)
    a = oracle.summate_c(cov, z1, z2, pos + t)
    b = oracle.summate_c(cov, z1r, z2r, pos)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12)


def test_single_mode_period():
    rs = np.random.RandomState(5)
    k = rs.normal(size=(3, 1))
    z1, z2 = np.array([0.7]), np.array([-1.3])
    pos = rs.uniform(-3, 3, size=(3, 20))
    shift = (2 * np.pi / np.sum(k**2)) * k  # one period along k
    assert np.allclose(oracle.summate_c(k, z1, z2, pos + shift), oracle.summate_c(k, z1, z2, pos),
                       rtol=1e-12, atol=1e-12)


def test_tensor_grid_gemm_identity():
    """On a tensor grid the sum is Re(A . diag(c) . B^T): an independent route."""
    rs = np.random.RandomState(6)
    n = 64
    cov = rs.normal(size=(2, n))
    z1, z2 = rs.normal(size=n), rs.normal(size=n)
    x = np.linspace(-4, 5, 13)
    y = np.linspace(2, 9, 11)
    pos = np.asarray(np.meshgrid(x, y, indexing="ij")).reshape(2, -1)
    a = np.exp(1j * np.outer(x, cov[0]))  # (nx, N)
    b = np.exp(1j * np.outer(y, cov[1]))  # (ny, N)
    c = z1 - 1j * z2
    gemm = np.real(np.einsum("xj,j,yj->xy", a, c, b)).reshape(-1)
    assert np.allclose(oracle.summate_c(cov, z1, z2, pos), gemm, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("d", [2, 3])
def test_incompr_is_divergence_free(d):
    cov, z1, z2, pos = _case(d, 25, 8, 7, kscale=0.7, xscale=3.0)
    h = 1e-5
    div = np.zeros(pos.shape[1])
    for c in range(d):
        e = np.zeros((d, 1))
        e[c] = h
        up = oracle.summate_incompr_c(cov, z1, z2, pos + e)[c]
        dn = oracle.summate_incompr_c(cov, z1, z2, pos - e)[c]
        div += (up - dn) / (2 * h)
    scale = np.sum(np.hypot(z1, z2) * np.linalg.norm(cov, axis=0))
    assert np.max(np.abs(div)) < 1e-8 * scale
    # the e1 component carries the projector 1 - k0^2/|k|^2, the others -k_c k0/|k|^2
    u = oracle.summate_incompr_c(cov, z1, z2, pos)
    w = oracle.summate_c(cov, z1, z2, pos)
    k2 = np.sum(cov**2, axis=0)
    w0 = oracle.summate_c(cov, z1 * (1 - cov[0] ** 2 / k2), z2 * (1 - cov[0] ** 2 / k2), pos)
    assert np.allclose(u[0], w0, rtol=1e-12, atol=1e-12)
    assert not np.allclose(u[0], w)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_adjoint_matches_finite_differences(d):
    cov, z1, z2, pos = _case(d, 12, 50, 8 + d, kscale=0.8, xscale=4.0)
    g = np.random.RandomState(9).normal(size=pos.shape[1])
    gz1, gz2, gk = oracle.summate_adjoint_c(cov, z1, z2, pos, g)
    n1, n2, nk = oracle.summate_adjoint_np(cov, z1, z2, pos, g)
    assert np.allclose(gz1, n1, rtol=1e-12, atol=1e-12)
    assert np.allclose(gz2, n2, rtol=1e-12, atol=1e-12)
    assert np.allclose(gk, nk, rtol=1e-12, atol=1e-11)
    loss = lambda c, a, b: float(g @ oracle.summate_c(c, a, b, pos))
    h = 1e-6
    for j in (0, 5, 11):
        e = np.zeros(12)
        e[j] = h
        assert abs((loss(cov, z1 + e, z2) - loss(cov, z1 - e, z2)) / (2 * h) - gz1[j]) < 1e-6
        assert abs((loss(cov, z1, z2 + e) - loss(cov, z1, z2 - e)) / (2 * h) - gz2[j]) < 1e-6
        for k in range(d):
            ck = cov.copy()
            ck[k, j] += h
            cm = cov.copy()
            cm[k, j] -= h
            fd = (loss(ck, z1, z2) - loss(cm, z1, z2)) / (2 * h)
            assert abs(fd - gk[k, j]) < 1e-5 * max(1.0, abs(fd))


def test_planted_mistakes_are_caught():
    """Slips planted in a copy of the oracle fail the brute-force check."""
    cov, z1, z2, pos = _case(3, 9, 11, 10)
    ref = oracle.summate_mp(cov, z1, z2, pos)

    def close(u):
        return np.max(np.abs(u - ref)) <= 1e-12 * np.sum(np.hypot(z1, z2))

    phase = cov.T @ pos
    assert close((z1[:, None] * np.cos(phase) + z2[:, None] * np.sin(phase)).sum(0))
    assert not close((z1[:, None] * np.cos(phase)).sum(0))                          # dropped sin term
    assert not close((z1[:, None] * np.cos(phase) - z2[:, None] * np.sin(phase)).sum(0))  # sign
    swapped = cov[::-1].T @ pos                                                      # axis mix-up
    assert not close((z1[:, None] * np.cos(swapped) + z2[:, None] * np.sin(swapped)).sum(0))
    assert not close((z2[:, None] * np.cos(phase) + z1[:, None] * np.sin(phase)).sum(0))  # z1/z2 swap
