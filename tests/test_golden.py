"""Golden fixtures (``tests/golden/make_golden.py``) against the oracle and,
on a GPU, against the kernels."""

import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


@pytest.mark.parametrize("d", [1, 2, 3])
def test_oracle_tiny_mpmath(d):
    g = load("tiny_mpmath.npz")
    cov, z1, z2, pos = g[f"d{d}_cov"], g[f"d{d}_z1"], g[f"d{d}_z2"], g[f"d{d}_pos"]
    scale = np.sum(np.hypot(z1, z2))
    assert np.max(np.abs(oracle.summate_c(cov, z1, z2, pos) - g[f"d{d}_sum"])) <= 1e-14 * scale
    if d > 1:
        assert np.max(np.abs(oracle.summate_incompr_c(cov, z1, z2, pos) - g[f"d{d}_incompr"])) <= 1e-14 * scale


def test_oracle_adjoint_mpmath():
    g = load("adjoint_case.npz")
    gz1, gz2, gk = oracle.summate_adjoint_c(g["cov"], g["z1"], g["z2"], g["pos"], g["g"])
    assert np.allclose(gz1, g["gz1"], rtol=0, atol=1e-13)
    assert np.allclose(gz2, g["gz2"], rtol=0, atol=1e-13)
    assert np.allclose(gk, g["gk"], rtol=0, atol=1e-11)


def test_config1_modes_are_reproduced():
    """The host sampler redraws the config-1 modes bit for bit."""
    from gstools_amd import Gaussian, RandMeth

    g = load("config1.npz")
    rm = RandMeth(Gaussian(dim=2, var=1.0, len_scale=10), mode_no=1000, seed=20170519)
    assert np.array_equal(rm._cov_sample, g["cov"])
    assert np.array_equal(rm._z_1, g["z1"]) and np.array_equal(rm._z_2, g["z2"])


@pytest.mark.gpu
@pytest.mark.parametrize("d", [1, 2, 3])
def test_gpu_tiny_mpmath(d):
    from gstools_amd import summate, summate_device, summate_incompr

    g = load("tiny_mpmath.npz")
    cov, z1, z2, pos = g[f"d{d}_cov"], g[f"d{d}_z1"], g[f"d{d}_z2"], g[f"d{d}_pos"]
    scale = np.sum(np.hypot(z1, z2))
    assert np.max(np.abs(summate(cov, z1, z2, pos) - g[f"d{d}_sum"])) <= 1e-14 * scale
    f32 = summate_device(cov, z1, z2, pos, "fp32").cpu().numpy()
    assert np.max(np.abs(f32 - g[f"d{d}_sum"])) <= 1e-5 * scale
    if d > 1:
        assert np.max(np.abs(summate_incompr(cov, z1, z2, pos) - g[f"d{d}_incompr"])) <= 1e-14 * scale


@pytest.mark.gpu
def test_gpu_config1_golden():
    from gstools_amd import summate, summate_grid_device

    g = load("config1.npz")
    got = summate(g["cov"], g["z1"], g["z2"], np.asarray(np.meshgrid(np.arange(100.0), np.arange(100.0),
                                                                   indexing="ij")).reshape(2, -1))
    assert oracle.rel_errors(got, g["summed"])[0] <= 1e-12
    mf = summate_grid_device(g["cov"], g["z1"], g["z2"], [np.arange(100.0)] * 2, "fp16x3").cpu().numpy()
    assert oracle.rel_errors(mf, g["summed"])[0] <= 1e-5


@pytest.mark.gpu
def test_gpu_adjoint_golden():
    from gstools_amd import summate_adjoint_device

    g = load("adjoint_case.npz")
    gz1, gz2, gk = (t.cpu().numpy() for t in summate_adjoint_device(g["cov"], g["z1"], g["z2"], g["pos"], g["g"]))
    assert np.allclose(gz1, g["gz1"], rtol=0, atol=1e-13)
    assert np.allclose(gz2, g["gz2"], rtol=0, atol=1e-13)
    assert np.allclose(gk, g["gk"], rtol=0, atol=1e-11)
