"""GPU tests of the SRF caller and of the C ABI's error statuses.

SRF: structured grids through the grid kernels agree with the meshgrid route
(bit for bit in FP64), the mean is applied like ``apply_mean_norm_trend``
(``normalizer/tools.py:35-110``), vector fields carry ``mean_u e1``, and the
reference's incompressible SRF known values (``tests/test_srf.py:259-275``)
come out of the GPU path.  Errors: wrong workspaces, shapes and precisions
map to ``ValueError`` / ``GsxError`` before or instead of a launch.
"""

import ctypes
import sys

import numpy as np
import pytest
import torch

import oracle
from gstools_amd import SRF, Exponential, Gaussian, RandMeth, _native, summate_device, summate_grid_device
from gstools_amd.geometry import generate_grid

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_srf_structured_equals_meshgrid_route():
    _say("srf structured")
    model = Gaussian(dim=2, var=2.0, len_scale=7.0)
    x, y = np.arange(0.0, 40.0), np.linspace(-5.0, 30.0, 33)
    srf = SRF(model, mean=1.5, mode_no=512, seed=5)
    grid = srf((x, y), mesh_type="structured")
    assert grid.shape == (40, 33)
    rm = RandMeth(model, mode_no=512, seed=5)
    ref = np.sqrt(2.0 / 512) * oracle.summate_c(rm._cov_sample, rm._z_1, rm._z_2, generate_grid((x, y))) + 1.5
    assert oracle.rel_errors(grid.ravel(), ref)[0] <= 1e-12
    pts = srf(generate_grid((x, y)))
    assert np.array_equal(pts, grid.ravel())  # same phases, same accumulation order
    bf = srf((x, y), mesh_type="structured", precision="bf16x3")
    assert oracle.rel_errors(bf.ravel() - 1.5, ref - 1.5)[1] <= 1e-5


def test_srf_anisotropic_structured_uses_isometrized_grid():
    _say("srf anisotropic")
    model = Exponential(dim=3, var=2.0, len_scale=[50, 20, 5], angles=[0.5, 0.2, 0.1])
    axes = (np.linspace(0, 100, 9), np.linspace(0, 50, 7), np.linspace(0, 20, 5))
    srf = SRF(model, mean=0.0, mode_no=256, seed=7)
    field = srf(axes, mesh_type="structured")
    rm = RandMeth(model, mode_no=256, seed=7)
    ref = np.sqrt(2.0 / 256) * oracle.summate_c(rm._cov_sample, rm._z_1, rm._z_2,
                                               model.isometrize(generate_grid(axes)))
    assert field.shape == (9, 7, 5)
    assert oracle.rel_errors(field.ravel(), ref)[0] <= 1e-12


def test_srf_incompressible_known_values():
    """``tests/test_srf.py:259-275`` through SRF on the GPU."""
    _say("srf incompr known values")
    rs = np.random.RandomState(123018)
    xt, yt = rs.uniform(0.0, 10, 100), rs.uniform(0.0, 10, 100)
    model = Gaussian(dim=2, var=0.5, len_scale=1.0)
    srf = SRF(model, mean=0.3, mode_no=100, generator="IncomprRandMeth", mean_velocity=0.5)
    field = srf((xt, yt), seed=476356)
    assert field[0, 0] == pytest.approx(1.23693272, abs=5e-8)
    assert field[0, 1] == pytest.approx(0.89242284, abs=5e-8)
    xg, yg = np.linspace(0.0, 12.0, 48), np.linspace(0.0, 10.0, 46)
    field = srf((xg, yg), seed=4734654, mesh_type="structured")
    assert field.shape == (2, 48, 46)
    assert field[0, 0, 0] == pytest.approx(1.07812013, abs=5e-8)
    assert field[0, 1, 0] == pytest.approx(1.06180674, abs=5e-8)


@pytest.mark.parametrize("dim,shape", [(2, (37, 29)), (3, (24, 20, 18)), (3, (30, 9, 5)), (3, (6, 140, 20))])
def test_incompressible_structured_all_precisions(dim, shape):
    """``IncomprRandMeth.structured``: each component is a grid sum with the per-mode
    projector weight; FP64 within 1e-12 of the incompressible sum on the materialised
    grid (the oracle), FP32 within 1e-6, the matrix-core splits within 1e-5 (rms)."""
    _say(f"incompr structured dim={dim}")
    from gstools_amd import IncomprRandMeth

    model = Gaussian(dim=dim, var=1.5, len_scale=4.0)
    axes = [np.linspace(0.0, 30.0, n) for n in shape]
    gen = IncomprRandMeth(model, mode_no=384, seed=11, mean_velocity=0.7)
    scale = 0.7 * np.sqrt(1.5 / 384)
    ref = scale * oracle.summate_incompr_c(gen._cov_sample, gen._z_1, gen._z_2, generate_grid(axes))
    ref[0] += 0.7
    for precision, bar in (("fp64", 1e-12), ("fp32", 1e-6), ("bf16x3", 1e-5), ("fp16x3", 1e-5)):
        out = gen.structured(axes, precision=precision).cpu().numpy().astype(np.float64)
        assert out.shape == (dim,) + shape
        for c in range(dim):
            err = oracle.rel_errors(out[c].ravel() - (0.7 if c == 0 else 0.0),
                                    ref[c] - (0.7 if c == 0 else 0.0))
            assert (err[0] if precision in ("fp64", "fp32") else err[1]) <= bar, (precision, c, err)
    pts = gen(generate_grid(axes))
    # the direct FP64 grid form has the phases of the materialised grid: same bits
    assert np.array_equal(gen.structured(axes, precision="fp64").cpu().numpy().reshape(dim, -1), pts)
    # the C ABI's incompressible grid entries reject 1D grids
    lib = _native.lib()
    shp = (ctypes.c_int64 * 3)(5, 1, 1)
    assert lib.gsx_summate_incompr_grid_mfma_workspace(1, 4, shp) < 0


def test_zero_variance_and_nugget():
    model = Gaussian(dim=2, var=0.0, len_scale=3.0)
    rm = RandMeth(model, mode_no=10, seed=1)
    assert np.array_equal(rm(np.zeros((2, 7)), add_nugget=False), np.zeros(7))
    model = Gaussian(dim=2, var=1.0, len_scale=3.0, nugget=0.25)
    rm = RandMeth(model, mode_no=50, seed=1)
    a = rm(np.zeros((2, 5)), add_nugget=False)
    b = rm(np.zeros((2, 5)))
    assert not np.allclose(a, b)  # nugget noise added


def test_error_statuses():
    _say("error statuses")
    lib = _native.lib()
    dev = torch.device("cuda")
    cov = torch.zeros((3, 32), dtype=torch.float64, device=dev)
    z = torch.zeros(32, dtype=torch.float64, device=dev)
    pos = torch.zeros((3, 100), dtype=torch.float64, device=dev)
    out = torch.empty(32, dtype=torch.float64, device=dev)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    ws_need = lib.gsx_summate_adjoint_workspace(3, 32, 100)
    ws = torch.empty(max(ws_need, 1), dtype=torch.uint8, device=dev)
    g = torch.zeros(100, dtype=torch.float64, device=dev)
    gk = torch.empty((3, 32), dtype=torch.float64, device=dev)
    st = lib.gsx_summate_adjoint_f64(3, 32, 100, p(cov), p(z), p(z), p(pos), p(g), p(out), p(out), p(gk), p(ws),
                                     ws_need - 8, None)
    assert st == -4  # workspace too small
    shape = (ctypes.c_int64 * 3)(4, 4, 4)
    st = lib.gsx_summate_grid_mfma(1, 32, shape, p(z), p(cov), p(z), p(z), 0, 1.0, p(out), p(ws), 0, None)
    assert st == -5  # 1D grids are not run on the matrix cores
    st = lib.gsx_summate_grid_mfma(3, 32, shape, p(z), p(cov), p(z), p(z), 7, 1.0, p(out), p(ws), 0, None)
    assert st == -1  # unknown variant
    with pytest.raises(ValueError):
        summate_device(cov.cpu().numpy(), z.cpu().numpy(), z.cpu().numpy(), np.zeros((2, 5)))
    with pytest.raises(ValueError):
        summate_device(cov, z, z, pos, precision="fp16")
    with pytest.raises(ValueError):
        summate_grid_device(cov, z, z, [np.arange(3.0)] * 2)
    torch.cuda.synchronize()
