"""ctypes binding of ``libgsx.so`` (declarations in ``include/gsx.h``).

There is no CPU fallback: when the library is missing or cannot be loaded
every summation entry point raises ``RuntimeError``.
"""

import ctypes
import os
import threading

__all__ = ["lib", "check", "GsxError", "LIB_PATH", "EXPORTS", "MFMA_BF16X3", "MFMA_FP16X3", "MFMA_UNIFORM_LAST"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgsx.so")

GSX_OK = 0
GSX_E_INVALID = -1
GSX_E_LAUNCH = -2
GSX_E_NO_DEVICE = -3
GSX_E_WORKSPACE = -4
GSX_E_UNSUPPORTED = -5
GSX_E_SCRATCH = -6
GSX_PROBE_BF16 = 0
GSX_PROBE_F64 = 1
NOALLOC_SCRATCH_BYTES = 32 << 20  # GSX_NOALLOC_SCRATCH_BYTES
MFMA_BF16X3 = 0
MFMA_FP16X3 = 1
MFMA_UNIFORM_LAST = 16  # flag: the last axis of a 2D grid is uniformly spaced
MFMA_FILL_SHIFT = 8     # tuning: log2 of the K-split fill target (8..11) in bits 8-11
MFMA_NO_SWAP = 0x1000   # tuning: never swap axes 0 and 1 of a 3D grid

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_f64 = ctypes.c_double
_i64p = ctypes.POINTER(ctypes.c_int64)


class StretchModel(ctypes.Structure):
    """Mirror of ``gsx_stretch_model``."""

    _fields_ = [("kind", ctypes.c_int32), ("dim", ctypes.c_int32), ("len_rescaled", ctypes.c_double),
                ("c", ctypes.c_double * 6)]


STRETCH_GAUSSIAN, STRETCH_EXPONENTIAL, STRETCH_MATERN = 0, 1, 2


class BatchItem(ctypes.Structure):
    """Mirror of ``gsx_batch_item``."""

    _fields_ = [
        ("d", ctypes.c_int32),
        ("kind", ctypes.c_int32),
        ("N", ctypes.c_int64),
        ("P", ctypes.c_int64),
        ("shape", ctypes.c_int64 * 3),
        ("cov", _vp),
        ("z1", _vp),
        ("z2", _vp),
        ("pos", _vp),
        ("axes", _vp),
        ("out", _vp),
    ]


_FWD = [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]
_GRID = [_i64, _i64, _i64p, _vp, _vp, _vp, _vp, _vp, _vp]
_FOUR = [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
_FOUR_GRID = [_i64, _i64, _i64p, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
_ADJ = [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp]

# name -> (restype, argtypes); every function declared in include/gsx.h
EXPORTS = {
    "gsx_status_string": (ctypes.c_char_p, [_i32]),
    "gsx_version": (_i32, []),
    "gsx_summate_f64": (_i32, _FWD),
    "gsx_summate_f32": (_i32, _FWD),
    "gsx_summate_incompr_f64": (_i32, _FWD),
    "gsx_summate_incompr_f32": (_i32, _FWD),
    "gsx_summate_grid_f64": (_i32, _GRID),
    "gsx_summate_grid_f32": (_i32, _GRID),
    "gsx_summate_grid_mfma_workspace": (_i64, [_i64, _i64, _i64p]),
    "gsx_summate_grid_mfma_plan": (_i32, [_i64, _i64, _i64p, _i64, _i32, _i64p]),
    "gsx_summate_grid_mfma": (_i32, [_i64, _i64, _i64p, _vp, _vp, _vp, _vp, _i32, _f64, _vp, _vp, _i64, _vp]),
    "gsx_summate_incompr_grid_f64": (_i32, _GRID),
    "gsx_summate_incompr_grid_f32": (_i32, _GRID),
    "gsx_summate_incompr_grid_mfma_workspace": (_i64, [_i64, _i64, _i64p]),
    "gsx_summate_incompr_grid_mfma": (_i32, [_i64, _i64, _i64p, _vp, _vp, _vp, _vp, _i32, _f64, _vp, _vp, _i64,
                                             _vp]),
    "gsx_summate_fourier_f64": (_i32, _FOUR),
    "gsx_summate_fourier_f32": (_i32, _FOUR),
    "gsx_summate_fourier_grid_f64": (_i32, _FOUR_GRID),
    "gsx_summate_fourier_grid_f32": (_i32, _FOUR_GRID),
    "gsx_summate_grid_gemm_workspace": (_i64, [_i64, _i64, _i64p]),
    "gsx_summate_grid_gemm_f64": (_i32, [_i64, _i64, _i64p, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "gsx_summate_fourier_tensor_workspace": (_i64, [_i64, _i64p, _i64p]),
    "gsx_summate_fourier_tensor_f64": (_i32, [_i64, _i64p, _vp, _i64p, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "gsx_summate_fourier_grid_mfma": (_i32, [_i64, _i64, _i64p, _vp, _vp, _vp, _vp, _vp, _i32, _f64, _vp, _vp, _i64,
                                             _vp]),
    "gsx_summate_adjoint_workspace": (_i64, [_i64, _i64, _i64]),
    "gsx_summate_adjoint_f64": (_i32, _ADJ),
    "gsx_summate_adjoint_f32": (_i32, _ADJ),
    "gsx_summate_adjoint_f32_noalloc": (_i32, [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gsx_krige_workspace": (_i64, [_i64, _i64, _i32]),
    "gsx_krige_cov_f64": (_i32, [_i32, _f64, _f64, _f64, _f64, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp]),
    "gsx_krige_f64": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "gsx_vario_workspace": (_i64, [_i64, _i64, _i64, _i64, _i64]),
    "gsx_vario_unstructured_f64": (_i32, [_i64, _i64, _i64, _i64, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _i64, _vp]),
    "gsx_vario_unstructured_haversine_f64": (_i32, [_i64, _i64, _i64, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _i64, _vp]),
    "gsx_vario_directional_f64": (_i32, [_i64, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _f64, _f64, _i32, _i32,
                                         _vp, _vp, _vp, _i64, _vp]),
    "gsx_vario_structured_f64": (_i32, [_i64, _i64, _vp, _vp, _i32, _vp, _vp, _vp, _i64, _vp]),
    "gsx_stretch_draws": (_i32, [_i64, _vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp, _i32]),
    "gsx_stretch_propose": (_i32, [_i64, _i32, _vp, _i64, _i64, _i64, _i32, _f64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gsx_stretch_accept": (_i32, [_i64, _i32, _vp, _i64, _i64, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64]),
    "gsx_mode_draws": (_i32, [_i64, _vp, _vp, _i32, _i32, _i32, _f64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                              _i32]),
    "gsx_stretch_run_device": (_i32, [_i64, _i32, _vp, _i64, _vp, _i64, _f64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gsx_stretch_pick_device": (_i32, [_i64, _vp, _vp, _i64, _i32, _vp, _vp, _vp]),
    "gsx_batch_plan_create": (_i32, [ctypes.POINTER(BatchItem), _i64, _i64, _i64, ctypes.POINTER(_vp)]),
    "gsx_batch_plan_launch": (_i32, [_vp, _vp]),
    "gsx_batch_plan_launch_graph": (_i32, [_vp, _vp]),
    "gsx_batch_plan_tiles": (_i64, [_vp]),
    "gsx_batch_plan_destroy": (None, [_vp]),
    "gsx_probe_mfma": (_i32, [_i32, _i32, _i64, _i64, _vp, _vp]),
}

_lock = threading.Lock()
_lib = None


class GsxError(RuntimeError):
    """A gsx C-ABI call returned a failure status."""


def lib():
    """Load ``libgsx.so`` once; raise ``RuntimeError`` when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"gstools_amd: native library {LIB_PATH} is missing; build it with "
                    "`python -m gstools_amd.build` (there is no CPU fallback)")
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in EXPORTS.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def check(status, what):
    """Map a gsx status to ``ValueError`` (invalid) or ``GsxError``."""
    if status == GSX_OK:
        return
    msg = lib().gsx_status_string(status).decode()
    if status == GSX_E_INVALID:
        raise ValueError(f"{what}: {msg}")
    raise GsxError(f"{what}: {msg} (status {status})")
