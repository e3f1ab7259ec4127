"""The reference's SRF anisotropy / rotation tests through the GPU SRF.

Mirrors ``tests/test_srf.py:85-225`` (Gaussian, var 1.5, len 4, mean 0.3,
mode_no 100, seed 825718662; the same grids and index maps) for every
precision.  "fp64" takes the reference's route on rotated or anisotropic
grids (meshgrid, isometrize, per-point sum) and is held to the reference's
7 decimals.  "fp32", "bf16x3" and "fp16x3" never build the meshgrid: the
isometrize matrix is folded into the wavenumbers (k' = M^T k) and the grid
kernels (matrix cores for the split precisions) run on the original axes;
their maps hold to the precision's own error.  Also: the folded grid path
against the FP64 oracle on the isometrized meshgrid (3D Exponential with the
rotation of the unstructured benchmark workload), and device modes that
follow in-place edits of ``_z_1`` (the reference reads the arrays on every
call, ``generator.py:266-268``).
"""

import sys

import numpy as np
import pytest
import torch

import oracle
from gstools_amd import SRF, Exponential, Fourier, Gaussian, RandMeth
from gstools_amd import srf as srf_module
from gstools_amd.geometry import generate_grid

pytestmark = pytest.mark.gpu

MEAN, MODE_NO, SEED = 0.3, 100, 825718662
X_GRID = np.linspace(0.0, 12.0, 48)
Y_GRID = np.linspace(0.0, 10.0, 46)
Z_GRID = np.linspace(0.0, 10.0, 40)
C_GRID = np.linspace(-6.0, 6.0, 8)
PRECISIONS = ["fp64", "fp32", "bf16x3", "fp16x3"]
# |difference| allowed between two fields at mapped indices: the reference's
# assertAlmostEqual (7 places) for FP64; the precision's error on fields of
# magnitude ~3 otherwise
ATOL = {"fp64": 5e-8, "fp32": 2e-6, "bf16x3": 1e-4, "fp16x3": 5e-5}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _forbid_meshgrid(pos):
    raise AssertionError("structured SRF built a meshgrid")


@pytest.fixture
def no_meshgrid(monkeypatch):
    """Fail if the SRF builds a meshgrid (the folded path must not)."""
    monkeypatch.setattr(srf_module, "generate_grid", _forbid_meshgrid)


def _srf_call(model, pos, mesh_type, precision, monkeypatch):
    """One SRF call; structured grids outside "fp64" must not build a meshgrid."""
    with monkeypatch.context() as mp:
        if precision != "fp64" and mesh_type == "structured":
            mp.setattr(srf_module, "generate_grid", _forbid_meshgrid)
        srf = SRF(model, mean=MEAN, mode_no=MODE_NO)
        return srf(pos, seed=SEED, mesh_type=mesh_type, precision=precision)


def _close(a, b, precision):
    assert abs(a - b) <= ATOL[precision], (a, b, abs(a - b))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_anisotropy_2d(precision, monkeypatch):
    _say(f"srf geometry: anisotropy 2d {precision}")
    m = Gaussian(dim=2, var=1.5, len_scale=4.0)
    iso = _srf_call(m, (X_GRID, Y_GRID), "structured", precision, monkeypatch)
    m.anis = 0.5
    ani = _srf_call(m, (X_GRID, Y_GRID), "structured", precision, monkeypatch)
    assert iso.shape == ani.shape == (48, 46)
    for a, b in (((0, 0), (0, 0)), ((0, 4), (0, 2)), ((0, 10), (0, 5))):
        _close(iso[a], ani[b], precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_anisotropy_3d(precision, monkeypatch):
    _say(f"srf geometry: anisotropy 3d {precision}")
    m = Gaussian(dim=2, var=1.5, len_scale=4.0)
    m.dim = 3
    grid = (X_GRID, Y_GRID, Z_GRID)
    iso = _srf_call(m, grid, "structured", precision, monkeypatch)
    m.anis = (0.5, 4.0)
    ani = _srf_call(m, grid, "structured", precision, monkeypatch)
    for a, b in (((0, 0, 0), (0, 0, 0)), ((0, 4, 0), (0, 2, 0)), ((0, 10, 0), (0, 5, 0)),
                 ((0, 0, 1), (0, 0, 4)), ((0, 0, 3), (0, 0, 12))):
        _close(iso[a], ani[b], precision)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_rotation_unstruct_2d(precision, monkeypatch):
    _say(f"srf geometry: rotation unstructured 2d {precision}")
    m = Gaussian(dim=2, var=1.5, len_scale=4.0)
    x_u, y_u = np.meshgrid(C_GRID, C_GRID)
    pos = (x_u.ravel(), y_u.ravel())
    m.anis = 0.25
    f = _srf_call(m, pos, "unstructured", precision, monkeypatch).reshape(8, 8)
    m.angles = -np.pi / 2.0
    r = _srf_call(m, pos, "unstructured", precision, monkeypatch).reshape(8, 8)
    _close(f[0, 0], r[-1, 0], precision)
    _close(f[1, 2], r[-3, 1], precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_rotation_struct_2d(precision, monkeypatch):
    _say(f"srf geometry: rotation structured 2d {precision}")
    m = Gaussian(dim=2, var=1.5, len_scale=4.0)
    m.anis = 0.25
    f = _srf_call(m, (C_GRID, C_GRID), "structured", precision, monkeypatch)
    m.angles = -np.pi / 2.0
    r = _srf_call(m, (C_GRID, C_GRID), "structured", precision, monkeypatch)
    _close(f[0, 0], r[0, -1], precision)
    _close(f[1, 2], r[2, 6], precision)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_rotation_unstruct_3d(precision, monkeypatch):
    _say(f"srf geometry: rotation unstructured 3d {precision}")
    m = Gaussian(dim=3, var=1.5, len_scale=4.0, anis=(0.25, 0.5))
    x_u, y_u, z_u = np.meshgrid(C_GRID, C_GRID, C_GRID)
    pos = (x_u.ravel(), y_u.ravel(), z_u.ravel())
    f = _srf_call(m, pos, "unstructured", precision, monkeypatch).reshape(8, 8, 8)
    m.angles = (-np.pi / 2.0, -np.pi / 2.0)
    r = _srf_call(m, pos, "unstructured", precision, monkeypatch).reshape(8, 8, 8)
    _close(f[0, 0, 0], r[-1, -1, 0], precision)
    _close(f[1, 2, 0], r[-3, -1, 1], precision)
    _close(f[0, 0, 1], r[-1, -2, 0], precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_rotation_struct_3d(precision, monkeypatch):
    _say(f"srf geometry: rotation structured 3d {precision}")
    m = Gaussian(dim=2, var=1.5, len_scale=4.0)
    m.dim = 3
    m.anis = 0.25
    grid = (C_GRID, C_GRID, C_GRID)
    f = _srf_call(m, grid, "structured", precision, monkeypatch)
    m.angles = -np.pi / 2.0
    r = _srf_call(m, grid, "structured", precision, monkeypatch)
    _close(f[0, 0, 0], r[0, 7, 0], precision)
    _close(f[0, 0, 1], r[0, 7, 1], precision)
    m.angles = (0, -np.pi / 2.0)
    r = _srf_call(m, grid, "structured", precision, monkeypatch)
    _close(f[0, 0, 0], r[7, 0, 0], precision)
    _close(f[0, 1, 0], r[7, 1, 0], precision)
    _close(f[1, 1, 0], r[7, 1, 1], precision)


@pytest.mark.parametrize("precision,bar", [("fp32", 1e-5), ("bf16x3", 5e-5), ("fp16x3", 2e-5),
                                           ("fp64gemm", 1e-10)])
def test_folded_grid_against_oracle(precision, bar, no_meshgrid):
    """Rotated anisotropic 3D grid on the folded path vs the FP64 oracle on
    the isometrized meshgrid (rel_max; rel_rms within 1e-5 for every split)."""
    _say(f"srf geometry: folded 3D grid vs oracle {precision}")
    model = Exponential(dim=3, var=2.0, len_scale=[50, 20, 5], angles=[0.5, 0.2, 0.1])
    axes = (np.arange(0.0, 400.0, 10.0), np.arange(0.0, 300.0, 10.0), np.arange(0.0, 200.0, 10.0))
    srf = SRF(model, mean=0.0, mode_no=1024, seed=20170519)
    got = srf(axes, mesh_type="structured", precision=precision)
    assert got.shape == (40, 30, 20)
    rm = RandMeth(model, mode_no=1024, seed=20170519)
    iso = model.isometrize(np.asarray(np.meshgrid(*axes, indexing="ij")).reshape(3, -1))
    ref = np.sqrt(2.0 / 1024) * oracle.summate_c(rm._cov_sample, rm._z_1, rm._z_2, iso)
    rmax, rrms = oracle.rel_errors(got.ravel(), ref)
    _say(f"  {precision}: rel_max {rmax:.2e} rel_rms {rrms:.2e}")
    assert rmax <= bar and rrms <= 1e-5


@pytest.mark.parametrize("angles", [0.0, 0.3])
def test_fourier_folded_grid(angles, no_meshgrid):
    """Fourier generator on an anisotropic (and rotated) grid: anisotropy keeps
    the modes a tensor grid, a rotation takes the direct sum; both against
    the oracle on the isometrized meshgrid."""
    _say(f"srf geometry: fourier folded angles={angles}")
    model = Gaussian(dim=2, var=1.0, len_scale=[20.0, 10.0], angles=angles)
    axes = (np.arange(0.0, 128.0), np.arange(0.0, 96.0))
    srf = SRF(model, generator="Fourier", period=[128.0, 96.0], mode_no=[32, 32], seed=7)
    got = srf(axes, mesh_type="structured", precision="fp32")
    gen = srf.generator
    iso = model.isometrize(np.asarray(np.meshgrid(*axes, indexing="ij")).reshape(2, -1))
    ref = oracle.summate_fourier_c(gen._spectrum_factor, gen._modes, gen._z_1, gen._z_2, iso)
    rmax, rrms = oracle.rel_errors(got.ravel(), ref)
    _say(f"  fourier angles={angles}: rel_max {rmax:.2e} rel_rms {rrms:.2e}")
    assert rmax <= 1e-5 and rrms <= 1e-5


def test_in_place_mode_edits_reach_the_device():
    """Writing ``_z_1`` / ``_cov_sample`` in place changes the next field."""
    _say("generator: in-place mode edits")
    model = Gaussian(dim=2, var=1.0, len_scale=10.0)
    rm = RandMeth(model, mode_no=256, seed=3)
    pos = np.random.RandomState(0).uniform(0.0, 50.0, size=(2, 500))
    first = rm(pos)
    rm._z_1[:] *= -1.0
    second = rm(pos)
    ref = np.sqrt(1.0 / 256) * oracle.summate_c(rm._cov_sample, rm._z_1, rm._z_2, pos)
    assert not np.allclose(first, second)
    assert oracle.rel_errors(second, ref)[0] <= 1e-12
    rm._cov_sample[0] *= 2.0
    third = rm(pos)
    ref = np.sqrt(1.0 / 256) * oracle.summate_c(rm._cov_sample, rm._z_1, rm._z_2, pos)
    assert oracle.rel_errors(third, ref)[0] <= 1e-12
    fields = rm.structured((np.arange(0.0, 20.0), np.arange(0.0, 10.0)), precision="fp64").cpu().numpy()
    grid = generate_grid((np.arange(0.0, 20.0), np.arange(0.0, 10.0)))
    ref = np.sqrt(1.0 / 256) * oracle.summate_c(rm._cov_sample, rm._z_1, rm._z_2, grid)
    assert oracle.rel_errors(fields.ravel(), ref)[0] <= 1e-12
