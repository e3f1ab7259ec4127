"""GPU parity of the Fourier method against the oracle.

Bars as for the randomisation method: FP64 <= 1e-12 relative (rel_max)
against ``oracle.summate_fourier_c``; FP32 <= 1e-5; split-precision matrix
cores <= 1e-5 rms (bf16x3 <= 5e-5 max).  The reference's known values and
periodicity checks (``tests/test_fouriergen.py:47-81``) run through SRF on the
GPU, on the meshgrid route and on the grid kernels.
"""

import ctypes
import sys

import numpy as np
import pytest
import torch

import oracle
from gstools_amd import (SRF, Fourier, Gaussian, _native, summate_fourier, summate_fourier_device,
                         summate_fourier_grid_device)
from gstools_amd.geometry import generate_grid

pytestmark = pytest.mark.gpu

SEED = 19900408
L = [80, 30, 91]
X = np.linspace(0, L[0], 11)
Y = np.linspace(0, L[1], 31)
Z = np.linspace(0, L[2], 13)
KNOWN = {
    1: (lambda: Gaussian(dim=1, var=0.5, len_scale=10.0), [12], [80], (X,), 0.6236929351309081),
    2: (lambda: Gaussian(dim=2, var=2.0, len_scale=30.0), [12, 6], L[:2], (X, Y), -0.1431996611581266),
    3: (lambda: Gaussian(dim=3, var=2.1, len_scale=21.0), [12, 6, 14], L, (X, Y, Z), -1.0433325279452803),
}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _gen(d, mode_no, period, len_scale, seed):
    return Fourier(Gaussian(dim=d, var=1.3, len_scale=len_scale), period=period, mode_no=mode_no, seed=seed)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_known_values_and_periodicity(dim):
    _say(f"fourier known values dim={dim}")
    model, mode_no, period, axes, want = KNOWN[dim]
    srf = SRF(model(), generator="Fourier", mode_no=mode_no, period=period, seed=SEED)
    mesh = srf(generate_grid(axes)).reshape([a.size for a in axes])  # meshgrid route
    grid = srf(axes, mesh_type="structured")                         # per-axis FP64 GEMMs
    direct = srf.generator.structured(axes, method="direct").cpu().numpy()  # grid kernel
    assert grid.shape == mesh.shape
    assert mesh.flat[0] == pytest.approx(want, abs=1e-12)
    assert grid.flat[0] == pytest.approx(want, abs=1e-12)
    assert np.array_equal(direct, mesh)
    assert np.max(np.abs(grid - mesh)) <= 1e-12 * np.max(np.abs(mesh))
    if dim == 1:
        assert grid[0] == pytest.approx(grid[-1], abs=1e-7)
    elif dim == 2:
        assert grid[0, len(Y) // 2] == pytest.approx(grid[-1, len(Y) // 2], abs=1e-7)
        assert grid[len(X) // 2, 0] == pytest.approx(grid[len(X) // 2, -1], abs=1e-7)
    else:
        assert grid[0, len(Y) // 2, 0] == pytest.approx(grid[-1, len(Y) // 2, 0], abs=1e-7)
        assert grid[0, 0, 0] == pytest.approx(grid[0, -1, 0], abs=1e-7)
        assert grid[len(X) // 2, len(Y) // 2, 0] == pytest.approx(grid[len(X) // 2, len(Y) // 2, -1], abs=1e-7)


@pytest.mark.parametrize("d,mode_no,p", [
    (1, [64], 1001),
    (2, [32, 48], 5003),
    (3, [16, 12, 20], 3001),   # 3840 modes: 15 LDS chunks
    (3, [2, 2, 2], 1),
])
def test_unstructured_fp64_fp32_match_oracle(d, mode_no, p):
    _say(f"fourier unstructured d={d} modes={mode_no} P={p}")
    gen = _gen(d, mode_no, [100.0, 70.0, 50.0][:d], 6.0, 11 + d)
    rs = np.random.RandomState(d)
    pos = rs.uniform(-300.0, 300.0, size=(d, p))
    args = (gen._spectrum_factor, gen._modes, gen._z_1, gen._z_2)
    ref = oracle.summate_fourier_c(*args, pos)
    got = summate_fourier(*args, pos)
    assert got.dtype == np.float64 and got.shape == (p,)
    assert oracle.rel_errors(got, ref)[0] <= 1e-12
    assert np.array_equal(gen(pos, add_nugget=False), got)
    f32 = summate_fourier_device(*args, pos, precision="fp32").cpu().numpy()
    rel_max, rel_rms = oracle.rel_errors(f32, ref)
    assert rel_max <= 1e-5 and rel_rms <= 1e-5, (rel_max, rel_rms)


@pytest.mark.parametrize("shape,mode_no", [
    ((37, 40, 23), [8, 6, 10]),
    ((64, 300), [24, 20]),
    ((130, 17, 33), [4, 4, 4]),
])
@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3", "fp16x3"])
def test_grid_paths_match_oracle(shape, mode_no, precision):
    _say(f"fourier grid {precision} shape={shape} modes={mode_no}")
    d = len(shape)
    period = [float(s) * 0.7 for s in shape]
    gen = _gen(d, mode_no, period, 3.0, 5 + d)
    axes = [np.linspace(0.0, 0.7 * s, s) for s in shape]
    args = (gen._spectrum_factor, gen._modes, gen._z_1, gen._z_2)
    ref = oracle.summate_fourier_c(*args, generate_grid(axes))
    got = summate_fourier_grid_device(*args, axes, precision).cpu().numpy().astype(np.float64)
    rel_max, rel_rms = oracle.rel_errors(got, ref)
    _say(f"  rel_max={rel_max:.3e} rel_rms={rel_rms:.3e}")
    if precision == "fp64":
        assert rel_max <= 1e-12
    elif precision in ("fp32", "fp16x3"):
        assert rel_max <= 1e-5 and rel_rms <= 1e-5
    else:
        assert rel_rms <= 1e-5 and rel_max <= 5e-5
    st = gen.structured(axes, precision=precision, add_nugget=False, method="direct").cpu().numpy()
    assert st.shape == shape
    assert np.array_equal(st.astype(np.float64).ravel(), got)


@pytest.mark.parametrize("shape,mode_no", [
    ((37,), [10]),                 # 1D: the last stage on the coefficients
    ((64, 300), [24, 20]),         # 2D: first and last stage
    ((200, 129), [130, 6]),        # more modes than one K stage, ragged tiles
    ((37, 40, 23), [8, 6, 10]),    # 3D: first, batched middle and last stage
    ((130, 17, 33), [4, 4, 4]),
    ((5, 3, 300), [2, 30, 18]),
])
def test_tensor_path_matches_oracle(shape, mode_no):
    """The modes are a tensor grid, so the sum is one FP64 GEMM contraction per axis:
    within 1e-12 of the oracle's per-point sum (it skips only the rounding of the phase
    sum), through the generator and the device API; the output dtype follows precision."""
    from gstools_amd import summate_fourier_tensor_device

    _say(f"fourier tensor shape={shape} modes={mode_no}")
    d = len(shape)
    period = [float(s) * 0.7 for s in shape]
    gen = _gen(d, mode_no, period, 3.0, 11 + d)
    axes = [np.linspace(-2.0, 0.7 * s, s) for s in shape]
    ref = oracle.summate_fourier_c(gen._spectrum_factor, gen._modes, gen._z_1, gen._z_2, generate_grid(axes))
    got = gen.structured(axes, add_nugget=False).cpu().numpy()
    assert got.shape == shape and got.dtype == np.float64
    assert oracle.rel_errors(got.ravel(), ref)[0] <= 1e-12
    dev = summate_fourier_tensor_device(gen._spectrum_factor, gen._kaxes, gen._z_1, gen._z_2, axes)
    assert np.array_equal(dev.cpu().numpy(), got.ravel())
    lo = gen.structured(axes, precision="bf16x3", add_nugget=False)
    assert lo.dtype == torch.float32 and np.array_equal(lo.cpu().numpy(), got.astype(np.float32))


def test_numpy_seam_runs_meshgrids_as_contractions():
    """The drop-in ``summate_fourier`` (numpy in, numpy out) sees the generator's modes
    and a structured mesh as two 'ij' meshgrids and sums per axis: bitwise the tensor
    path, within 1e-12 of the oracle; scattered points take the per-point sum."""
    from gstools_amd import summate_fourier_tensor_device

    _say("fourier numpy seam: meshgrid -> contraction")
    gen = _gen(2, [20, 14], [50.0, 40.0], 4.0, 21)
    axes = [np.linspace(0.0, 50.0, 81), np.linspace(-3.0, 40.0, 67)]  # 5427 points, 280 modes
    pos = generate_grid(axes)
    args = (gen._spectrum_factor, gen._modes, gen._z_1, gen._z_2)
    got = summate_fourier(*args, pos)
    ten = summate_fourier_tensor_device(gen._spectrum_factor, gen._kaxes, gen._z_1, gen._z_2, axes)
    assert np.array_equal(got, ten.cpu().numpy())
    assert oracle.rel_errors(got, oracle.summate_fourier_c(*args, pos))[0] <= 1e-12
    scattered = np.random.RandomState(4).uniform(0, 50, size=(2, 300))
    per_point = summate_fourier_device(*args, scattered, "fp64").cpu().numpy()
    assert np.array_equal(summate_fourier(*args, scattered), per_point)
    # opting out keeps the meshgrid on the per-point sum: bitwise its result
    direct = summate_fourier_device(*args, pos, "fp64").cpu().numpy()
    assert np.array_equal(summate_fourier(*args, pos, tensor=False), direct)
    import importlib
    smod = importlib.import_module("gstools_amd.summate")
    saved = smod.FOURIER_TENSOR
    try:
        smod.FOURIER_TENSOR = False
        assert np.array_equal(summate_fourier(*args, pos), direct)
    finally:
        smod.FOURIER_TENSOR = saved


def test_tensor_path_errors_and_empty():
    from gstools_amd import summate_fourier_tensor_device

    k = [np.arange(4.0), np.arange(3.0)]
    z = np.ones(12)
    with pytest.raises(ValueError):
        summate_fourier_tensor_device(z, k, z, z, [np.arange(5.0)])          # axes / wavenumber axes
    with pytest.raises(ValueError):
        summate_fourier_tensor_device(z[:11], k, z, z, [np.arange(5.0)] * 2)  # mode count
    zero = summate_fourier_tensor_device(np.zeros(0), [np.zeros(0), np.arange(3.0)], np.zeros(0), np.zeros(0),
                                         [np.arange(5.0)] * 2)
    assert np.array_equal(zero.cpu().numpy(), np.zeros(25))
    empty = summate_fourier_tensor_device(z, k, z, z, [np.zeros(0), np.arange(5.0)])
    assert empty.numel() == 0


def test_empty_and_errors():
    _say("fourier empty and error statuses")
    gen = _gen(2, [4, 4], [10.0, 10.0], 2.0, 1)
    args = (gen._spectrum_factor, gen._modes, gen._z_1, gen._z_2)
    assert summate_fourier(*args, np.zeros((2, 0))).shape == (0,)
    empty = summate_fourier(np.zeros(0), np.zeros((2, 0)), np.zeros(0), np.zeros(0), np.ones((2, 5)))
    assert np.array_equal(empty, np.zeros(5))
    with pytest.raises(ValueError):
        summate_fourier(args[0][:-1], *args[1:], np.zeros((2, 3)))
    with pytest.raises(ValueError):
        summate_fourier(args[0].astype(np.float32), *args[1:], np.zeros((2, 3)))
    lib = _native.lib()
    dev = torch.device("cuda")
    k = torch.zeros((2, 16), dtype=torch.float64, device=dev)
    z = torch.zeros(16, dtype=torch.float64, device=dev)
    pos = torch.zeros((2, 10), dtype=torch.float64, device=dev)
    out = torch.empty(10, dtype=torch.float64, device=dev)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    assert lib.gsx_summate_fourier_f64(2, 16, 10, None, p(k), p(z), p(z), p(pos), p(out), None) == -1
    shape = (ctypes.c_int64 * 3)(4, 4, 1)
    ax = torch.zeros(8, dtype=torch.float64, device=dev)
    assert lib.gsx_summate_fourier_grid_f64(2, 16, shape, p(ax), None, p(k), p(z), p(z), p(out), None) == -1
    assert lib.gsx_summate_fourier_grid_mfma(1, 16, shape, p(ax), p(z), p(k), p(z), p(z), 0, 1.0, p(out),
                                             p(out), 0, None) == -5
    torch.cuda.synchronize()
