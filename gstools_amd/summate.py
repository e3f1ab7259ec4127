"""Drop-in summation functions backed by the gfx950 kernels.

``summate``, ``summate_incompr`` and ``summate_fourier`` keep the signature
and semantics of the functions GSTools imports from its native backend
(``src/gstools/field/generator.py:22-24``, dispatched at ``:42-75``):
float64 numpy inputs ``cov_samples (d, N)``, ``z_1 (N,)``, ``z_2 (N,)``,
``pos (d, P)``, a fresh float64 result, ``ValueError`` on bad shapes or
dtypes, and a ``num_threads`` argument that the GPU path accepts and ignores.
``summate_fourier`` takes the per-mode ``spectrum_factor (N,)`` first and the
regular wavenumber grid ``modes (d, N)`` in place of ``cov_samples``.

The ``*_device`` functions take and return torch tensors that stay on the
GPU (numpy inputs are copied over); they add the FP32, structured-grid,
matrix-core and adjoint paths.  Nothing here falls back to the CPU: without a
HIP device or without ``libgsx.so`` every call raises ``RuntimeError``.
"""

import ctypes
import os

import numpy as np
import torch

from . import _native

__all__ = [
    "summate", "summate_incompr", "summate_device", "summate_incompr_device",
    "summate_grid_device", "summate_adjoint_device", "summate_fourier", "summate_fourier_device",
    "summate_fourier_grid_device", "summate_fourier_tensor_device", "tensor_grid_axes", "PRECISIONS",
    "GRID_PRECISIONS", "INCOMPR_GRID_PRECISIONS",
]

PRECISIONS = ("fp64", "fp32")
GRID_PRECISIONS = ("fp64", "fp32", "bf16x3", "fp16x3", "fp64gemm")
# incompressible grids: no FP64 GEMM form
INCOMPR_GRID_PRECISIONS = ("fp64", "fp32", "bf16x3", "fp16x3")


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("gstools_amd: no HIP device is visible (the summation runs only on the GPU)")
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else ctypes.c_void_p(0)


def host_array(t):
    """Device tensor -> fresh numpy array (a blocking copy into numpy-allocated
    memory: on MI355X torch's ``.cpu()`` took 7.9 ms for a 32 MB float64 field,
    this copy 1.8 ms)."""
    a = np.empty(tuple(t.shape), dtype=torch.empty(0, dtype=t.dtype).numpy().dtype)
    if a.size:
        torch.from_numpy(a).copy_(t)
    return a


def _as_f64(x, name, dev):
    if isinstance(x, torch.Tensor):
        if x.dtype != torch.float64:
            raise ValueError(f"{name}: expected float64, got {x.dtype}")
        return x.to(dev).contiguous()
    arr = np.asarray(x)
    if arr.dtype != np.float64:
        if not np.issubdtype(arr.dtype, np.number):
            raise ValueError(f"{name}: expected a float64 array, got {arr.dtype}")
        arr = arr.astype(np.float64)
    return torch.from_numpy(np.ascontiguousarray(arr)).to(dev)


def _modes(cov_samples, z_1, z_2, dev):
    cov = _as_f64(cov_samples, "cov_samples", dev)
    if cov.dim() == 1:
        cov = cov.reshape(1, -1)
    if cov.dim() != 2:
        raise ValueError(f"cov_samples: expected shape (d, N), got {tuple(cov.shape)}")
    z1 = _as_f64(z_1, "z_1", dev).reshape(-1)
    z2 = _as_f64(z_2, "z_2", dev).reshape(-1)
    d, n = cov.shape
    if z1.numel() != n or z2.numel() != n:
        raise ValueError(f"z_1/z_2 must have N={n} entries, got {z1.numel()} and {z2.numel()}")
    if not 1 <= d <= 3:
        raise ValueError(f"only 1, 2 or 3 dimensions are supported, got d={d}")
    return cov, z1, z2


# points on a 2D grid's last axis from which the matrix-core path factorises a
# uniformly spaced axis (n / 16 >= 16 virtual points: one column tile)
SPLIT_MIN = 256


def uniform_axis(axis):
    """True when the 1-D ``axis`` is uniformly spaced in the sense the matrix-core
    path's ``MFMA_UNIFORM_LAST`` flag requires: ``y[j] == y[16 (j // 16)] + (y[j % 16]
    - y[0])`` to 4 ulp of ``max|y|`` for every j (the same FP64 operations as the
    library's device check), with at least ``SPLIT_MIN`` points.  NaN -> False."""
    if isinstance(axis, torch.Tensor):
        y = axis.detach().to("cpu", torch.float64).numpy().reshape(-1)
    else:
        y = np.asarray(axis, dtype=np.float64).reshape(-1)
    n = y.size
    if n < SPLIT_MIN:
        return False
    j = np.arange(n)
    r = y[j & ~15] + (y[j & 15] - y[0])
    tol = 2.0 ** -50 * np.max(np.abs(y))
    return bool(np.all(np.abs(y - r) <= tol))


def _factor(spectrum_factor, n, dev):
    if spectrum_factor is None:
        return None
    f = _as_f64(spectrum_factor, "spectrum_factor", dev).reshape(-1)
    if f.numel() != n:
        raise ValueError(f"spectrum_factor must have N={n} entries, got {f.numel()}")
    return f


def _points(pos, d, dev):
    p = _as_f64(pos, "pos", dev)
    if p.dim() == 1:
        p = p.reshape(1, -1)
    if p.dim() != 2 or p.shape[0] != d:
        raise ValueError(f"pos: expected shape ({d}, P), got {tuple(p.shape)}")
    return p


def summate_device(cov_samples, z_1, z_2, pos, precision="fp64", out=None, spectrum_factor=None):
    """Unscaled sum on the GPU -> torch tensor (P,) float64 or float32.

    With ``spectrum_factor`` (N,) every term is multiplied by its factor (the
    Fourier method's sum).
    """
    dev = _device()
    cov, z1, z2 = _modes(cov_samples, z_1, z_2, dev)
    d, n = cov.shape
    fac = _factor(spectrum_factor, n, dev)
    p = _points(pos, d, dev)
    npts = p.shape[1]
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}")
    dtype = torch.float64 if precision == "fp64" else torch.float32
    if out is None:
        out = torch.empty(npts, dtype=dtype, device=dev)
    elif out.dtype != dtype or out.numel() != npts or not out.is_contiguous():
        raise ValueError("out: wrong dtype, size or layout")
    lib = _native.lib()
    if fac is not None:
        fn = lib.gsx_summate_fourier_f64 if precision == "fp64" else lib.gsx_summate_fourier_f32
        _native.check(fn(d, n, npts, _ptr(fac), _ptr(cov), _ptr(z1), _ptr(z2), _ptr(p), _ptr(out), _stream()),
                      "summate_fourier")
        return out
    fn = lib.gsx_summate_f64 if precision == "fp64" else lib.gsx_summate_f32
    _native.check(fn(d, n, npts, _ptr(cov), _ptr(z1), _ptr(z2), _ptr(p), _ptr(out), _stream()), "summate")
    return out


def summate_incompr_device(cov_samples, z_1, z_2, pos, precision="fp64", out=None):
    """Unscaled incompressible sum on the GPU -> torch tensor (d, P)."""
    dev = _device()
    cov, z1, z2 = _modes(cov_samples, z_1, z_2, dev)
    d, n = cov.shape
    if d < 2:
        raise ValueError(f"summate_incompr: incompressible sums need d = 2 or 3, got d = {d}")
    p = _points(pos, d, dev)
    npts = p.shape[1]
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}")
    dtype = torch.float64 if precision == "fp64" else torch.float32
    if out is None:
        out = torch.empty((d, npts), dtype=dtype, device=dev)
    elif out.dtype != dtype or tuple(out.shape) != (d, npts) or not out.is_contiguous():
        raise ValueError("out: wrong dtype, shape or layout")
    fn = (_native.lib().gsx_summate_incompr_f64 if precision == "fp64"
          else _native.lib().gsx_summate_incompr_f32)
    _native.check(fn(d, n, npts, _ptr(cov), _ptr(z1), _ptr(z2), _ptr(p), _ptr(out), _stream()),
                  "summate_incompr")
    return out


def summate_grid_device(cov_samples, z_1, z_2, axes, precision="fp64", alpha=1.0, out=None,
                        spectrum_factor=None, tune=False):
    """Sum on the 'ij' meshgrid of ``axes`` without materialising it.

    ``precision``: "fp64" / "fp32" (direct sums; "fp64" reproduces the phases
    of the materialised grid bit for bit), "bf16x3" / "fp16x3" (matrix-core
    GEMM over per-axis phase factors, d = 2 or 3) or "fp64gemm" (FP64 GEMM on
    the matrix cores, float64 out; the phase sum is not rounded, about
    |k.x| 2^-52 per term).  Returns a flat tensor of prod(shape) values in C
    order, multiplied by ``alpha`` (split-precision matrix-core path only; the
    other paths return the unscaled sum and require alpha == 1).
    ``spectrum_factor`` (N,) weights each term (the Fourier method's sum).
    ``tune``: on the matrix cores, run the plan the autotuner chose for this
    grid class (timed on first use, then cached; :mod:`.autotune`).
    """
    dev = _device()
    cov, z1, z2 = _modes(cov_samples, z_1, z_2, dev)
    d, n = cov.shape
    fac = _factor(spectrum_factor, n, dev)
    uniform_last = len(axes) == 2 and precision in ("bf16x3", "fp16x3") and uniform_axis(axes[1])
    axes = [_as_f64(a, "axes", dev).reshape(-1) for a in axes]
    if len(axes) != d:
        raise ValueError(f"expected {d} axes, got {len(axes)}")
    shape = [int(a.numel()) for a in axes]
    npts = int(np.prod(shape))
    flat = torch.cat(axes) if axes else torch.empty(0, dtype=torch.float64, device=dev)
    shape_c = (ctypes.c_int64 * 3)(*(shape + [1] * (3 - d)))
    if precision not in GRID_PRECISIONS:
        raise ValueError(f"precision must be one of {GRID_PRECISIONS}")
    dtype = torch.float64 if precision in ("fp64", "fp64gemm") else torch.float32
    if out is None:
        out = torch.empty(npts, dtype=dtype, device=dev)
    elif out.dtype != dtype or out.numel() != npts or not out.is_contiguous():
        raise ValueError("out: wrong dtype, size or layout")
    lib = _native.lib()
    if precision == "fp64gemm":
        # FP64 GEMM on the matrix cores: rows axis 0, columns the other axes (their
        # complex phase products built in bounded column chunks); the phase sum is not
        # rounded, so it differs from the per-point sum by about |k.x| 2^-52 per term
        if alpha != 1.0:
            raise ValueError("alpha applies to the split-precision matrix-core path only")
        ws_bytes = int(lib.gsx_summate_grid_gemm_workspace(d, n, shape_c))
        if ws_bytes < 0:
            _native.check(ws_bytes, "summate_grid_gemm")
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        _native.check(lib.gsx_summate_grid_gemm_f64(d, n, shape_c, _ptr(flat), _ptr(cov), _ptr(z1), _ptr(z2),
                                                    _ptr(fac), _ptr(out), _ptr(ws), ws_bytes, _stream()),
                      "summate_grid_gemm")
        return out
    if precision in ("fp64", "fp32"):
        if alpha != 1.0:
            raise ValueError("alpha applies to the matrix-core path only")
        if fac is not None:
            fn = lib.gsx_summate_fourier_grid_f64 if precision == "fp64" else lib.gsx_summate_fourier_grid_f32
            _native.check(fn(d, n, shape_c, _ptr(flat), _ptr(fac), _ptr(cov), _ptr(z1), _ptr(z2), _ptr(out),
                             _stream()), "summate_fourier_grid")
            return out
        fn = lib.gsx_summate_grid_f64 if precision == "fp64" else lib.gsx_summate_grid_f32
        _native.check(fn(d, n, shape_c, _ptr(flat), _ptr(cov), _ptr(z1), _ptr(z2), _ptr(out), _stream()),
                      "summate_grid")
        return out
    variant = _native.MFMA_BF16X3 if precision == "bf16x3" else _native.MFMA_FP16X3
    if uniform_last:  # factorised last axis (no n x N phase table)
        variant |= _native.MFMA_UNIFORM_LAST
    if tune and fac is None:
        from .autotune import mfma_variant

        variant = mfma_variant(d, n, shape, variant)
    ws_bytes = int(lib.gsx_summate_grid_mfma_workspace(d, n, shape_c))
    if ws_bytes < 0:
        _native.check(int(ws_bytes), "summate_grid_mfma")
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    if fac is not None:
        _native.check(lib.gsx_summate_fourier_grid_mfma(d, n, shape_c, _ptr(flat), _ptr(fac), _ptr(cov), _ptr(z1),
                                                        _ptr(z2), variant, float(alpha), _ptr(out), _ptr(ws),
                                                        ws_bytes, _stream()), "summate_fourier_grid_mfma")
        return out
    _native.check(lib.gsx_summate_grid_mfma(d, n, shape_c, _ptr(flat), _ptr(cov), _ptr(z1), _ptr(z2),
                                            variant, float(alpha), _ptr(out), _ptr(ws), ws_bytes,
                                            _stream()), "summate_grid_mfma")
    return out


def summate_incompr_grid_device(cov_samples, z_1, z_2, axes, precision="fp64", alpha=1.0, out=None, tune=False):
    """Incompressible sum on the 'ij' meshgrid of ``axes`` -> tensor (d, prod(shape)).

    "fp64" / "fp32": the direct sum, one phase per (point, mode) for all d
    components ("fp64" gives the phases of the materialised grid bit for bit;
    the result is unscaled and ``alpha`` must be 1).  "bf16x3" / "fp16x3": one
    matrix-core GEMM per component in one launch, with the projector weight
    per mode, multiplied by ``alpha``; ``tune`` as in :func:`summate_grid_device`.
    """
    dev = _device()
    cov, z1, z2 = _modes(cov_samples, z_1, z_2, dev)
    d, n = cov.shape
    if d < 2:
        raise ValueError(f"summate_incompr: incompressible sums need d = 2 or 3, got d = {d}")
    axes = [_as_f64(a, "axes", dev).reshape(-1) for a in axes]
    if len(axes) != d:
        raise ValueError(f"expected {d} axes, got {len(axes)}")
    shape = [int(a.numel()) for a in axes]
    npts = int(np.prod(shape))
    flat = torch.cat(axes)
    shape_c = (ctypes.c_int64 * 3)(*(shape + [1] * (3 - d)))
    if precision not in INCOMPR_GRID_PRECISIONS:
        raise ValueError(f"precision must be one of {INCOMPR_GRID_PRECISIONS}")
    dtype = torch.float64 if precision == "fp64" else torch.float32
    if out is None:
        out = torch.empty((d, npts), dtype=dtype, device=dev)
    elif out.dtype != dtype or tuple(out.shape) != (d, npts) or not out.is_contiguous():
        raise ValueError("out: wrong dtype, shape or layout")
    lib = _native.lib()
    if precision in ("fp64", "fp32"):
        if alpha != 1.0:
            raise ValueError("alpha applies to the matrix-core path only")
        fn = lib.gsx_summate_incompr_grid_f64 if precision == "fp64" else lib.gsx_summate_incompr_grid_f32
        _native.check(fn(d, n, shape_c, _ptr(flat), _ptr(cov), _ptr(z1), _ptr(z2), _ptr(out), _stream()),
                      "summate_incompr_grid")
        return out
    variant = _native.MFMA_BF16X3 if precision == "bf16x3" else _native.MFMA_FP16X3
    if tune:
        from .autotune import mfma_variant

        variant = mfma_variant(d, n, shape, variant, comps=d)
    ws_bytes = int(lib.gsx_summate_incompr_grid_mfma_workspace(d, n, shape_c))
    if ws_bytes < 0:
        _native.check(int(ws_bytes), "summate_incompr_grid_mfma")
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    _native.check(lib.gsx_summate_incompr_grid_mfma(d, n, shape_c, _ptr(flat), _ptr(cov), _ptr(z1), _ptr(z2),
                                                    variant, float(alpha), _ptr(out), _ptr(ws), ws_bytes,
                                                    _stream()), "summate_incompr_grid_mfma")
    return out


def summate_fourier_device(spectrum_factor, modes, z_1, z_2, pos, precision="fp64", out=None):
    """Fourier-method sum on the GPU -> torch tensor (P,).

    ``sum_j f_j (z1_j cos(k_j.x) + z2_j sin(k_j.x))`` with ``f`` =
    ``spectrum_factor`` (``generator.py:681-688``); unscaled, like the
    reference's ``summate_fourier``.
    """
    if spectrum_factor is None:
        raise ValueError("spectrum_factor is required")
    return summate_device(modes, z_1, z_2, pos, precision, out, spectrum_factor=spectrum_factor)


def summate_fourier_grid_device(spectrum_factor, modes, z_1, z_2, axes, precision="fp64", alpha=1.0, out=None):
    """Fourier-method sum on the 'ij' meshgrid of ``axes`` (flat, C order)."""
    if spectrum_factor is None:
        raise ValueError("spectrum_factor is required")
    return summate_grid_device(modes, z_1, z_2, axes, precision, alpha, out, spectrum_factor=spectrum_factor)


def summate_fourier_tensor_device(spectrum_factor, kaxes, z_1, z_2, axes, out=None):
    """Fourier-method sum on the 'ij' meshgrid of ``axes`` for modes on a tensor grid:
    ``kaxes`` holds one wavenumber array per axis and the modes are their 'ij' meshgrid
    (``spectrum_factor``, ``z_1``, ``z_2`` over all of them, C order), as the Fourier
    generator builds them (``generator.py:838-862``).  One FP64 GEMM contraction per
    axis on the matrix cores (``gsx_summate_fourier_tensor_f64``); returns the flat
    float64 sum in C order."""
    dev = _device()
    kax = [_as_f64(k, "kaxes", dev).reshape(-1) for k in kaxes]
    axs = [_as_f64(a, "axes", dev).reshape(-1) for a in axes]
    d = len(axs)
    if not 1 <= d <= 3 or len(kax) != d:
        raise ValueError("one wavenumber axis per grid axis, 1 to 3 axes")
    mode_no = [int(k.numel()) for k in kax]
    n = int(np.prod(mode_no))
    fac = _as_f64(spectrum_factor, "spectrum_factor", dev).reshape(-1)
    z1 = _as_f64(z_1, "z_1", dev).reshape(-1)
    z2 = _as_f64(z_2, "z_2", dev).reshape(-1)
    if fac.numel() != n or z1.numel() != n or z2.numel() != n:
        raise ValueError(f"spectrum_factor, z_1, z_2 must have prod(mode_no) = {n} entries")
    shape = [int(a.numel()) for a in axs]
    npts = int(np.prod(shape))
    if out is None:
        out = torch.empty(npts, dtype=torch.float64, device=dev)
    elif out.dtype != torch.float64 or out.numel() != npts or not out.is_contiguous():
        raise ValueError("out: wrong dtype, size or layout")
    m_c = (ctypes.c_int64 * 3)(*(mode_no + [1] * (3 - d)))
    s_c = (ctypes.c_int64 * 3)(*(shape + [1] * (3 - d)))
    lib = _native.lib()
    ws_bytes = int(lib.gsx_summate_fourier_tensor_workspace(d, m_c, s_c))
    if ws_bytes < 0:
        _native.check(ws_bytes, "summate_fourier_tensor")
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    flat_k = torch.cat(kax) if n else torch.zeros(1, dtype=torch.float64, device=dev)
    flat_x = torch.cat(axs)
    _native.check(lib.gsx_summate_fourier_tensor_f64(d, m_c, _ptr(flat_k), s_c, _ptr(flat_x), _ptr(fac), _ptr(z1),
                                                      _ptr(z2), _ptr(out), _ptr(ws), ws_bytes, _stream()),
                  "summate_fourier_tensor")
    return out


def summate_adjoint_device(cov_samples, z_1, z_2, pos, g, precision="fp64"):
    """Unscaled adjoint: (gz1 (N,), gz2 (N,), gk (d, N)) float64 tensors."""
    dev = _device()
    cov, z1, z2 = _modes(cov_samples, z_1, z_2, dev)
    d, n = cov.shape
    p = _points(pos, d, dev)
    npts = p.shape[1]
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}")
    gdt = torch.float64 if precision == "fp64" else torch.float32
    if isinstance(g, torch.Tensor):
        gg = g.to(device=dev, dtype=gdt).contiguous().reshape(-1)
    else:
        gg = torch.from_numpy(np.ascontiguousarray(g, dtype=np.float64)).to(dev, gdt).reshape(-1)
    if gg.numel() != npts:
        raise ValueError(f"g: expected {npts} values, got {gg.numel()}")
    gz1 = torch.empty(n, dtype=torch.float64, device=dev)
    gz2 = torch.empty(n, dtype=torch.float64, device=dev)
    gk = torch.empty((d, n), dtype=torch.float64, device=dev)
    lib = _native.lib()
    ws_bytes = int(lib.gsx_summate_adjoint_workspace(d, n, npts))
    if ws_bytes < 0:
        _native.check(int(ws_bytes), "summate_adjoint")
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    fn = lib.gsx_summate_adjoint_f64 if precision == "fp64" else lib.gsx_summate_adjoint_f32
    _native.check(fn(d, n, npts, _ptr(cov), _ptr(z1), _ptr(z2), _ptr(p), _ptr(gg), _ptr(gz1), _ptr(gz2),
                     _ptr(gk), _ptr(ws), ws_bytes, _stream()), "summate_adjoint")
    return gz1, gz2, gk


def _check_numpy(cov_samples, z_1, z_2, pos):
    for name, arr in (("cov_samples", cov_samples), ("z_1", z_1), ("z_2", z_2), ("pos", pos)):
        if isinstance(arr, np.ndarray) and arr.dtype != np.float64:
            raise ValueError(f"Buffer dtype mismatch for {name}: expected 'double' but got '{arr.dtype}'")


def summate(cov_samples, z_1, z_2, pos, num_threads=None):
    """Drop-in for the backend ``summate``: FP64 sum on the GPU -> numpy (P,)."""
    _check_numpy(cov_samples, z_1, z_2, pos)
    return host_array(summate_device(cov_samples, z_1, z_2, pos, "fp64"))


# the numpy Fourier seam takes the per-axis contraction for meshgrids this large
TENSOR_MIN_POINTS = 4096
TENSOR_MIN_MODES = 64
# The numpy seam ``summate_fourier`` runs meshgrids of modes and points as
# per-axis FP64 contractions (within ~1e-14 of the per-point sum, not bit for
# bit).  False, or GSX_FOURIER_TENSOR=0 in the environment, keeps every call
# on the per-point sum, whose phases follow the reference loop bit for bit.
FOURIER_TENSOR = os.environ.get("GSX_FOURIER_TENSOR", "1") != "0"


def tensor_grid_axes(pts):
    """The axes of ``pts`` (d, P) if it is exactly the flattened 'ij' meshgrid of one
    array per dimension (as ``generate_grid`` and the Fourier generator's modes are),
    else None.  Candidate sizes come from where the outer coordinates first change;
    the candidate is then checked entry by entry, so repeated axis values can only
    make it answer None."""
    pts = np.asarray(pts)
    if pts.ndim != 2 or pts.shape[1] == 0:
        return None
    d, n = pts.shape
    sizes = [0] * d
    inner = 1
    for a in range(d - 1, 0, -1):  # run of constant outer coordinates = product of inner sizes
        change = np.flatnonzero((pts[:a] != pts[:a, :1]).any(axis=0))
        run = int(change[0]) if change.size else n
        if run % inner:
            return None
        sizes[a] = run // inner
        inner = run
    if n % inner:
        return None
    sizes[0] = n // inner
    axes = []
    for a in range(d):
        step = int(np.prod(sizes[a + 1:], dtype=np.int64))
        ax = pts[a, ::step][:sizes[a]]
        shape = [1] * d
        shape[a] = sizes[a]
        if not np.array_equal(pts[a].reshape(sizes), np.broadcast_to(ax.reshape(shape), sizes)):
            return None
        axes.append(np.ascontiguousarray(ax))
    return axes


def summate_fourier(spectrum_factor, modes, z_1, z_2, pos, num_threads=None, tensor=None):
    """Drop-in for the backend ``summate_fourier``: FP64 sum on the GPU -> numpy (P,).

    When ``modes`` and ``pos`` are both flattened 'ij' meshgrids in 2D or 3D (the
    Fourier generator on a structured mesh) of at least ``TENSOR_MIN_POINTS`` points
    and ``TENSOR_MIN_MODES`` modes, the sum runs as one FP64 GEMM contraction per
    axis (``summate_fourier_tensor_device``; within ~1e-14 of the per-point sum, not
    bit for bit); otherwise as the per-point sum.  ``tensor=False`` (or
    ``GSX_FOURIER_TENSOR=0`` in the environment, read at import) keeps every call on
    the per-point sum, whose phases follow the reference loop bit for bit."""
    use_tensor = FOURIER_TENSOR if tensor is None else bool(tensor)
    _check_numpy(modes, z_1, z_2, pos)
    if isinstance(spectrum_factor, np.ndarray) and spectrum_factor.dtype != np.float64:
        raise ValueError(f"Buffer dtype mismatch for spectrum_factor: expected 'double' but got "
                         f"'{spectrum_factor.dtype}'")
    if use_tensor and isinstance(modes, np.ndarray) and isinstance(pos, np.ndarray) and modes.ndim == 2 \
            and pos.ndim == 2 \
            and 2 <= modes.shape[0] == pos.shape[0] <= 3 and pos.shape[1] >= TENSOR_MIN_POINTS \
            and modes.shape[1] >= TENSOR_MIN_MODES and np.size(spectrum_factor) == modes.shape[1] \
            and np.size(z_1) == modes.shape[1] and np.size(z_2) == modes.shape[1]:
        kax = tensor_grid_axes(modes)
        xax = tensor_grid_axes(pos) if kax is not None else None
        if xax is not None:
            return host_array(summate_fourier_tensor_device(spectrum_factor, kax, z_1, z_2, xax))
    return host_array(summate_fourier_device(spectrum_factor, modes, z_1, z_2, pos, "fp64"))


def summate_incompr(cov_samples, z_1, z_2, pos, num_threads=None):
    """Drop-in for the backend ``summate_incompr``: FP64 sum -> numpy (d, P)."""
    _check_numpy(cov_samples, z_1, z_2, pos)
    return host_array(summate_incompr_device(cov_samples, z_1, z_2, pos, "fp64"))
