"""GPU parity of the matrix-core grid path and of the adjoint.

Matrix-core path (split bf16 / fp16 products, FP32 accumulation): compared
with the oracle's FP64 sum on the materialised grid; rel_max and rel_rms are
printed and bounded (fp16x3 <= 1e-5 for both, bf16x3 rel_rms <= 1e-5).

A uniformly spaced last axis of a 2D grid (``MFMA_UNIFORM_LAST``) is summed in its
factorised form, to the same bars; the flag's device check turns a non-uniform axis
into an all-NaN output.

Adjoint: FP64 kernel against the oracle's C adjoint (<= 1e-10 relative to the
gradient's scale), against central finite differences of the GPU forward in
FP64 (h = 1e-4), and against torch float64 autograd of the same formula; the
FP32 kernel within 1e-5 of the FP64 one.  The autograd wrapper's len_scale
gradient is checked against finite differences.
"""

import ctypes
import sys

import numpy as np
import pytest
import torch

import oracle
from gstools_amd import (Gaussian, Matern, RandMeth, _native, summate_adjoint_device, summate_device,
                         summate_grid_device)
from gstools_amd.autograd import RandMethSum, randmeth_field
from gstools_amd.geometry import generate_grid
from gstools_amd.summate import uniform_axis

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _modes(d, n, seed, kscale=1.0):
    rs = np.random.RandomState(seed)
    return rs.normal(size=(d, n)) * kscale, rs.normal(size=n), rs.normal(size=n)


@pytest.mark.parametrize("shape,n", [
    ((37, 40, 23), 300),     # ragged everywhere: rows, columns and modes padded
    ((130, 17, 33), 64),     # two row tiles, column tiles with clamped i1
    ((64, 300), 500),        # 2D grid (one-point middle axis)
    ((5, 16, 16), 16),       # a single K step
    ((40, 8, 300), 200),     # 3D, short middle axis: flat columns (generic kernel)
    ((64, 300, 7), 256),     # 3D, short last axis: flat columns across i1 rows
    ((33, 5, 5), 64),        # both column axes short: one ragged column tile
    ((20, 300, 40), 200),    # short axis 0: axes 0 and 1 swapped (rows = axis 1)
    ((5, 200, 17), 100),     # swapped, and the swapped column axes need the flat kernel
])
@pytest.mark.parametrize("precision", ["fp16x3", "bf16x3"])
def test_mfma_grid_matches_fp64(shape, n, precision):
    _say(f"mfma {precision} shape={shape} N={n}")
    d = len(shape)
    cov, z1, z2 = _modes(d, n, 77 + n, kscale=0.6)
    axes = [np.linspace(-3.0 * (i + 1), 11.0 + 2 * i, s) for i, s in enumerate(shape)]
    ref = oracle.summate_c(cov, z1, z2, generate_grid(axes))
    got = summate_grid_device(cov, z1, z2, axes, precision).cpu().numpy().astype(np.float64)
    rel_max, rel_rms = oracle.rel_errors(got, ref)
    _say(f"  rel_max={rel_max:.3e} rel_rms={rel_rms:.3e}")
    if precision == "fp16x3":
        assert rel_max <= 1e-5 and rel_rms <= 1e-5
    else:
        assert rel_rms <= 1e-5 and rel_max <= 5e-5


def test_mfma_alpha_and_empty_modes():
    _say("mfma alpha / N=0")
    cov, z1, z2 = _modes(3, 40, 5)
    axes = [np.arange(20.0), np.arange(18.0), np.arange(17.0)]
    a = summate_grid_device(cov, z1, z2, axes, "fp16x3", alpha=0.25).cpu().numpy()
    b = summate_grid_device(cov, z1, z2, axes, "fp16x3").cpu().numpy()
    assert np.allclose(a, 0.25 * b, rtol=1e-6, atol=1e-7)
    cov0, z10, z20 = _modes(3, 0, 5)
    z = summate_grid_device(cov0, z10, z20, axes, "bf16x3").cpu().numpy()
    assert np.array_equal(z, np.zeros(20 * 18 * 17, dtype=np.float32))


def _grid_mfma_raw(cov, z1, z2, axes, variant):
    """gsx_summate_grid_mfma through the C ABI with an explicit variant word."""
    lib = _native.lib()
    dev = torch.device("cuda")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    d, n = cov.shape
    shape = [len(a) for a in axes]
    shape_c = (ctypes.c_int64 * 3)(*(shape + [1] * (3 - d)))
    ws_bytes = int(lib.gsx_summate_grid_mfma_workspace(d, n, shape_c))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    out = torch.empty(int(np.prod(shape)), dtype=torch.float32, device=dev)
    flat, tc, t1, t2 = t(np.concatenate(axes)), t(cov), t(z1), t(z2)
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    st = lib.gsx_summate_grid_mfma(d, n, shape_c, p(flat), p(tc), p(t1), p(t2), variant, 1.0, p(out), p(ws),
                                   ws_bytes, None)
    assert st == 0, st
    return out.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("shape,n", [
    ((20, 256), 100),    # one row tile, 16 virtual rows: K split over blocks
    ((37, 1000), 700),   # 1000 = 62 * 16 + 8: the last virtual row is half outside
    ((130, 4100), 64),   # two row tiles, 257 virtual rows
])
def test_mfma_uniform_last_axis(shape, n):
    """A uniformly spaced last axis of a 2D grid (>= 256 points) is factorised into
    y_(16a) and y_b - y_0 and summed by the 3D matrix-core kernel: within the same bars
    of the oracle's FP64 sum as the plain 2D form, and close to that form."""
    _say(f"mfma uniform last axis shape={shape} N={n}")
    cov, z1, z2 = _modes(2, n, 91 + n, kscale=0.6)
    axes = [np.linspace(-4.0, 9.0, shape[0]), 2.5 + 0.37 * np.arange(shape[1])]
    assert uniform_axis(axes[1])
    ref = oracle.summate_c(cov, z1, z2, generate_grid(axes))
    for precision, variant in (("fp16x3", _native.MFMA_FP16X3), ("bf16x3", _native.MFMA_BF16X3)):
        got = summate_grid_device(cov, z1, z2, axes, precision).cpu().numpy().astype(np.float64)
        split = _grid_mfma_raw(cov, z1, z2, axes, variant | _native.MFMA_UNIFORM_LAST)
        plain = _grid_mfma_raw(cov, z1, z2, axes, variant)
        assert np.array_equal(got, split)  # the host sets the flag for this axis
        for name, v in (("split", split), ("plain", plain)):
            rel_max, rel_rms = oracle.rel_errors(v, ref)
            _say(f"  {precision} {name} rel_max={rel_max:.3e} rel_rms={rel_rms:.3e}")
            if precision == "fp16x3":
                assert rel_max <= 1e-5 and rel_rms <= 1e-5
            else:
                assert rel_rms <= 1e-5 and rel_max <= 5e-5
        assert not np.array_equal(split, plain)  # the two forms round differently


def test_mfma_uniform_flag_checked_and_ignored():
    """The flag on a non-uniform axis gives an all-NaN output (the device check); on a 3D
    grid or an axis shorter than 256 points it is ignored (same bits as without it)."""
    _say("mfma uniform flag: non-uniform axis / ignored cases")
    cov, z1, z2 = _modes(2, 200, 3, kscale=0.6)
    y = np.arange(300.0) * 0.5
    y[123] += 1e-9
    assert not uniform_axis(y)
    out = _grid_mfma_raw(cov, z1, z2, [np.arange(40.0), y], _native.MFMA_BF16X3 | _native.MFMA_UNIFORM_LAST)
    assert np.isnan(out).all()
    short = [np.arange(40.0), np.arange(255.0)]
    a = _grid_mfma_raw(cov, z1, z2, short, _native.MFMA_BF16X3 | _native.MFMA_UNIFORM_LAST)
    assert np.array_equal(a, _grid_mfma_raw(cov, z1, z2, short, _native.MFMA_BF16X3))
    cov3, z13, z23 = _modes(3, 100, 4, kscale=0.6)
    ax3 = [np.arange(20.0), np.arange(17.0), np.arange(300.0)]
    a = _grid_mfma_raw(cov3, z13, z23, ax3, _native.MFMA_FP16X3 | _native.MFMA_UNIFORM_LAST)
    assert np.array_equal(a, _grid_mfma_raw(cov3, z13, z23, ax3, _native.MFMA_FP16X3))


@pytest.mark.parametrize("shape,n", [
    ((300,), 500),            # 1D: one column (the coefficients)
    ((64, 300), 333),         # 2D
    ((37, 40, 23), 300),      # 3D, ragged tiles
    ((8, 100, 100), 4096),    # 10000 columns of 2 x 4096 doubles: two 512 MB column chunks
])
def test_fp64gemm_grid_matches_oracle(shape, n):
    """precision "fp64gemm": the grid sum as an FP64 GEMM over column chunks of the phase
    products; it skips only the rounding of the phase sum, so with phases of at most a
    few hundred radians it stays within 1e-12 of the oracle."""
    _say(f"fp64gemm shape={shape} N={n}")
    d = len(shape)
    cov, z1, z2 = _modes(d, n, 5 + n, kscale=0.6)
    axes = [np.linspace(-3.0 * (i + 1), 11.0 + 2 * i, s) for i, s in enumerate(shape)]
    ref = oracle.summate_c(cov, z1, z2, generate_grid(axes))
    got = summate_grid_device(cov, z1, z2, axes, "fp64gemm")
    assert got.dtype == torch.float64
    rel_max = oracle.rel_errors(got.cpu().numpy(), ref)[0]
    _say(f"  rel_max={rel_max:.3e}")
    assert rel_max <= 1e-12


def test_fp64gemm_generators():
    """RandMeth.structured and the Fourier method's direct route take "fp64gemm"; the
    incompressible grid entry refuses it."""
    from gstools_amd import Fourier, IncomprRandMeth, summate_fourier_grid_device, summate_incompr_grid_device

    _say("fp64gemm generators")
    rm = RandMeth(Gaussian(dim=2, var=1.5, len_scale=6.0), mode_no=700, seed=9)
    axes = [np.linspace(0, 40, 50), np.linspace(0, 30, 61)]
    a = rm.structured(axes, precision="fp64gemm", add_nugget=False).cpu().numpy()
    b = rm.structured(axes, precision="fp64", add_nugget=False).cpu().numpy()
    assert a.shape == b.shape and oracle.rel_errors(a.ravel(), b.ravel())[0] <= 1e-12
    gen = Fourier(Gaussian(dim=2, var=1.0, len_scale=5.0), period=[40.0, 30.0], mode_no=[16, 12], seed=4)
    args = (gen._spectrum_factor, gen._modes, gen._z_1, gen._z_2)
    f = summate_fourier_grid_device(*args, axes, "fp64gemm").cpu().numpy()
    ref = oracle.summate_fourier_c(*args, generate_grid(axes))
    assert oracle.rel_errors(f, ref)[0] <= 1e-12
    with pytest.raises(ValueError):
        summate_incompr_grid_device(rm._cov_sample, rm._z_1, rm._z_2, axes, "fp64gemm")


def test_mfma_matern_config_sample():
    """Matern(nu=1.5, len 8), N 4096 (MCMC modes): 4 z-planes of the 512^3 grid."""
    _say("mfma matern sample")
    model = Matern(dim=3, var=1.0, len_scale=8.0, nu=1.5)
    rm = RandMeth(model, mode_no=4096, seed=20170519)
    ax = np.arange(512) * 0.5
    axes = [ax[:64], ax[:48], ax[100:104]]
    ref = oracle.summate_c(rm._cov_sample, rm._z_1, rm._z_2, generate_grid(axes))
    for precision in ("fp16x3", "bf16x3"):
        got = summate_grid_device(rm._cov_sample, rm._z_1, rm._z_2, axes, precision).cpu().numpy()
        rel_max, rel_rms = oracle.rel_errors(got.astype(np.float64), ref)
        _say(f"  {precision} rel_max={rel_max:.3e} rel_rms={rel_rms:.3e}")
        assert rel_rms <= 1e-5


@pytest.mark.parametrize("d", [1, 2, 3])
def test_adjoint_fp64_matches_oracle(d):
    _say(f"adjoint fp64 d={d}")
    cov, z1, z2 = _modes(d, 700, 10 + d, kscale=0.5)
    rs = np.random.RandomState(20 + d)
    pos = rs.uniform(-50, 50, size=(d, 5001))
    g = rs.normal(size=5001)
    r1, r2, rk = oracle.summate_adjoint_c(cov, z1, z2, pos, g)
    a1, a2, ak = (t.cpu().numpy() for t in summate_adjoint_device(cov, z1, z2, pos, g, "fp64"))
    scale = np.sum(np.abs(g)) * np.max(np.abs(pos))
    assert np.max(np.abs(a1 - r1)) <= 1e-12 * np.sum(np.abs(g))
    assert np.max(np.abs(a2 - r2)) <= 1e-12 * np.sum(np.abs(g))
    assert np.max(np.abs(ak - rk)) <= 1e-12 * scale * 4
    b1, b2, bk = (t.cpu().numpy() for t in summate_adjoint_device(cov, z1, z2, pos, g, "fp32"))
    assert np.max(np.abs(b1 - r1)) <= 1e-5 * np.sum(np.abs(g))
    assert np.max(np.abs(bk - rk)) <= 1e-5 * scale * 4


def test_adjoint_empty_points():
    cov, z1, z2 = _modes(2, 33, 3)
    a1, a2, ak = summate_adjoint_device(cov, z1, z2, np.zeros((2, 0)), np.zeros(0), "fp64")
    assert torch.count_nonzero(a1) == 0 and torch.count_nonzero(ak) == 0


def test_adjoint_vs_finite_differences_and_torch_autograd():
    """Stable-like 2D setting at oracle-friendly size: FD (h = 1e-4) in FP64."""
    _say("adjoint FD / autograd")
    dev = torch.device("cuda")
    cov_np, z1_np, z2_np = _modes(2, 64, 31, kscale=1.0 / 15)
    rs = np.random.RandomState(32)
    pos = torch.tensor(rs.uniform(0, 500, size=(2, 3000)), device=dev)
    obs = torch.tensor(rs.normal(size=3000), device=dev)
    z1 = torch.tensor(z1_np, device=dev, requires_grad=True)
    z2 = torch.tensor(z2_np, device=dev, requires_grad=True)
    cov = torch.tensor(cov_np, device=dev, requires_grad=True)
    loss = torch.mean((RandMethSum.apply(z1, z2, cov, pos, "fp64") - obs) ** 2)
    loss.backward()
    # torch float64 autograd of the same formula (dense phase matrix)
    z1r, z2r, covr = (t.detach().clone().requires_grad_(True) for t in (z1, z2, cov))
    ph = covr.T @ pos
    ur = (z1r[:, None] * torch.cos(ph) + z2r[:, None] * torch.sin(ph)).sum(0)
    torch.mean((ur - obs) ** 2).backward()
    for a, b in ((z1.grad, z1r.grad), (z2.grad, z2r.grad), (cov.grad, covr.grad)):
        assert torch.allclose(a, b, rtol=1e-9, atol=1e-9 * b.abs().max().item())

    def f(z1v, z2v, covv):
        return torch.mean((summate_device(covv, z1v, z2v, pos, "fp64") - obs) ** 2).item()

    h = 1e-4
    base = [t.detach().clone() for t in (z1, z2, cov)]
    for which, grad in ((0, z1.grad), (1, z2.grad)):
        for j in (0, 17, 63):
            up = [t.clone() for t in base]
            dn = [t.clone() for t in base]
            up[which][j] += h
            dn[which][j] -= h
            fd = (f(*up) - f(*dn)) / (2 * h)
            assert abs(fd - grad[j].item()) <= 1e-6 * max(1.0, abs(fd))
    hk = 1e-6  # phases move by |x| hk <= 5e-4 rad: truncation error ~4e-8 relative
    for dd, j in ((0, 3), (1, 40)):
        up = [t.clone() for t in base]
        dn = [t.clone() for t in base]
        up[2][dd, j] += hk
        dn[2][dd, j] -= hk
        fd = (f(*up) - f(*dn)) / (2 * hk)
        assert abs(fd - cov.grad[dd, j].item()) <= 1e-5 * max(1.0, abs(fd))


def test_len_scale_gradient():
    _say("len_scale gradient")
    dev = torch.device("cuda")
    model = Gaussian(dim=2, var=1.0, len_scale=1.0)
    rm = RandMeth(model, mode_no=256, seed=20170519)
    k_unit = torch.tensor(rm._cov_sample, device=dev)
    z1 = torch.tensor(rm._z_1, device=dev)
    z2 = torch.tensor(rm._z_2, device=dev)
    pos = torch.tensor(np.random.RandomState(1).uniform(0, 100, size=(2, 2000)), device=dev)
    obs = torch.tensor(np.random.RandomState(2).normal(size=2000), device=dev)
    length = torch.tensor(15.0, dtype=torch.float64, device=dev, requires_grad=True)
    loss = torch.mean((randmeth_field(z1, z2, k_unit, pos, 1.0, length) - obs) ** 2)
    loss.backward()
    h = 1e-4

    def f(lv):
        with torch.no_grad():
            return torch.mean((randmeth_field(z1, z2, k_unit, pos, 1.0, lv) - obs) ** 2).item()

    fd = (f(15.0 + h) - f(15.0 - h)) / (2 * h)
    assert abs(fd - length.grad.item()) <= 1e-6 * max(1.0, abs(fd))
