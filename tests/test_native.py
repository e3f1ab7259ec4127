"""CPU checks of the native boundary: the library loads, exports every symbol
``include/gsx.h`` declares, rejects bad arguments before touching a device,
and its device math (built for the host) matches mpmath."""

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from gstools_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gsx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gsx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_native.EXPORTS), "ctypes table and header disagree"


def test_status_strings_and_version():
    lib = _native.lib()
    assert lib.gsx_version() >= 100
    assert lib.gsx_status_string(0) == b"ok"
    assert b"invalid" in lib.gsx_status_string(-1)


def test_invalid_arguments_are_rejected_on_the_host():
    lib = _native.lib()
    null = ctypes.c_void_p(0)
    # d out of range, negative sizes: rejected before any HIP call
    assert lib.gsx_summate_f64(4, 10, 10, null, null, null, null, null, null) == -1
    assert lib.gsx_summate_f32(0, 10, 10, null, null, null, null, null, null) == -1
    assert lib.gsx_summate_f64(2, -1, 10, null, null, null, null, null, null) == -1
    assert lib.gsx_summate_incompr_f64(1, 10, 10, null, null, null, null, null, null) == -1
    # nothing to do: P == 0 returns ok without launching
    assert lib.gsx_summate_f64(2, 10, 0, null, null, null, null, null, null) == 0
    # null pointers with work to do
    assert lib.gsx_summate_f64(2, 10, 10, null, null, null, null, null, null) == -1
    assert lib.gsx_summate_adjoint_workspace(4, 1, 1) == -1
    assert lib.gsx_summate_adjoint_workspace(2, 0, 5) == 0
    assert lib.gsx_summate_adjoint_workspace(2, 2048, 10**6) > 0
    shape = (ctypes.c_int64 * 3)(10, 10, 1)
    assert lib.gsx_summate_grid_f64(2, 5, shape, null, null, null, null, null, null) == -1
    # matrix-core probe: unknown kind, no iterations or blocks, null output
    assert lib.gsx_probe_mfma(2, 0, 10, 8, null, null) == -1
    assert lib.gsx_probe_mfma(0, 0, 0, 8, null, null) == -1
    assert lib.gsx_probe_mfma(1, 1, 10, 0, null, null) == -1
    assert lib.gsx_probe_mfma(1, 0, 10, 8, null, null) == -1
    # allocation-free adjoint: bad d, and more modes than its static array holds
    args = [null] * 8 + [None]
    assert lib.gsx_summate_adjoint_f32_noalloc(4, 10, 10, *args) == -1
    too_many = _native.NOALLOC_SCRATCH_BYTES // (8 * 8) + 1
    assert lib.gsx_summate_adjoint_f32_noalloc(3, too_many, 10, *args) == _native.GSX_E_SCRATCH
    assert lib.gsx_summate_adjoint_f32_noalloc(1, 2 * too_many, 10, *args) == _native.GSX_E_SCRATCH
    assert b"scratch" in lib.gsx_status_string(_native.GSX_E_SCRATCH)


def test_python_boundary_validates_like_the_reference():
    from gstools_amd import summate

    cov = np.zeros((2, 3))
    z = np.zeros(3)
    with pytest.raises(ValueError):
        summate(cov.astype(np.float32), z, z, np.zeros((2, 4)))
    with pytest.raises((ValueError, RuntimeError)):
        summate(cov, z, z, np.zeros((3, 4)))


@pytest.fixture(scope="module")
def devmath(tmp_path_factory):
    out = tmp_path_factory.mktemp("devmath") / "devmath.so"
    src = os.path.join(ROOT, "tests", "native", "devmath_host.cpp")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-o", str(out), src], check=True)
    lib = ctypes.CDLL(str(out))
    dp = ctypes.POINTER(ctypes.c_double)
    lib.devmath_cos_shifted.argtypes = [dp, dp, dp, ctypes.c_long]
    lib.devmath_sincos.argtypes = [dp, dp, dp, ctypes.c_long]
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def test_device_trig_matches_mpmath(devmath):
    import mpmath as mp

    rs = np.random.RandomState(11)
    n = 3000
    x = np.concatenate([rs.uniform(-10, 10, n // 3), rs.uniform(-3e5, 3e5, n // 3),
                        rs.uniform(-2**46, 2**46, n - 2 * (n // 3))])
    x[:40] = (np.arange(40) - 20) * (np.pi / 2)  # near the reduction boundaries
    phi = rs.uniform(-np.pi, np.pi, n)
    cs = np.empty(n)
    s = np.empty(n)
    c = np.empty(n)
    devmath.devmath_cos_shifted(_p(x), _p(phi), _p(cs), n)
    devmath.devmath_sincos(_p(x), _p(s), _p(c), n)
    with mp.workdps(40):
        ref_cs = np.array([float(mp.cos(mp.mpf(a) - mp.mpf(b))) for a, b in zip(x, phi)])
        ref_s = np.array([float(mp.sin(mp.mpf(a))) for a in x])
        ref_c = np.array([float(mp.cos(mp.mpf(a))) for a in x])
    assert np.max(np.abs(cs - ref_cs)) < 1e-15
    assert np.max(np.abs(s - ref_s)) < 3e-16
    assert np.max(np.abs(c - ref_c)) < 3e-16


@pytest.mark.parametrize("fn", ["gsx_summate_grid_mfma_workspace", "gsx_summate_incompr_grid_mfma_workspace"])
def test_mfma_workspace_queries(fn):
    """Workspace sizes are host arithmetic: no modes, empty axes, small grids (whose K
    steps are split over blocks, adding a partial-sum buffer) and large grids."""
    q = getattr(_native.lib(), fn)
    shp = lambda *s: (ctypes.c_int64 * 3)(*(list(s) + [1] * (3 - len(s))))
    assert q(3, 0, shp(8, 8, 8)) >= 0  # N = 0: no K steps
    assert q(3, 64, shp(0, 8, 8)) >= 0 and q(2, 64, shp(16, 0)) >= 0  # empty axes
    small = q(3, 4096, shp(40, 48, 48))
    tables = 2 * (128 * 4096) * 2 * 2 + (48 * 4096 + (3 if "incompr" in fn else 1) * 48 * 4096) * 8
    assert small > tables  # the partial sums come after the tables
    big = q(3, 4096, shp(512, 512, 512))
    assert big == 2 * (512 * 4096) * 2 * 2 + (512 * 4096 + (3 if "incompr" in fn else 1) * 512 * 4096) * 8
    assert q(3, -1, shp(8, 8, 8)) < 0 and q(1, 8, shp(8)) < 0


def test_uniform_axis_check():
    """The host side of ``MFMA_UNIFORM_LAST``: uniform spacing to 4 ulp of max|y| with at
    least 256 points; a perturbed point, a NaN, a short axis or a geometric axis fail."""
    from gstools_amd.summate import SPLIT_MIN, uniform_axis

    assert SPLIT_MIN == 256
    assert uniform_axis(np.arange(4096.0))
    assert uniform_axis(np.linspace(-5.0, 5.0, 257))
    assert uniform_axis(2.5 + 0.37 * np.arange(4100))
    assert uniform_axis(1e6 + 0.1 * np.arange(1000))
    assert not uniform_axis(np.arange(255.0))
    y = np.arange(300.0) * 0.5
    y[123] += 1e-9
    assert not uniform_axis(y)
    y[123] = np.nan
    assert not uniform_axis(y)
    assert not uniform_axis(np.geomspace(1.0, 2.0, 400))
    # the 2D workspace query also covers the factorised form (256 points and more)
    q = _native.lib().gsx_summate_grid_mfma_workspace
    shp = lambda *s: (ctypes.c_int64 * 3)(*(list(s) + [1] * (3 - len(s))))
    assert q(2, 100, shp(20, 256)) >= q(3, 100, shp(20, 16, 16))
    assert q(2, 4096, shp(4096, 4096)) >= q(3, 4096, shp(4096, 256, 16))


def _mfma_plan(d, N, shape, comps=1, variant=0):
    info = (ctypes.c_int64 * 8)()
    shp = (ctypes.c_int64 * 3)(*(list(shape) + [1] * (3 - len(shape))))
    assert _native.lib().gsx_summate_grid_mfma_plan(d, N, shp, comps, variant, info) == 0
    keys = ("swapped", "split", "row_tiles", "col_tiles", "ksplit", "blocks", "ws", "kernel")
    return dict(zip(keys, list(info)))


def test_mfma_plan_decisions():
    """The matrix-core grid plan is host arithmetic: a 3D grid with a short axis 0
    runs with axes 0 and 1 swapped (fewer tiles; here a K split with partial
    sums), a cubic grid keeps its order, a thin slab keeps axis 0 as rows, and
    the workspace of each plan is the tables plus the partial sums."""
    a256 = lambda b: -(-b // 256) * 256

    def tables(rows, n1, n2, N=4096, comps=1):
        rows_p = -(-rows // 128) * 128
        return 2 * (rows_p * 2 * N) * 2 + (n1 * N + comps * n2 * N) * 8

    # (32, 512, 512): swapped -> rows 512, column axes (32, 512): 4 x 64 = 256 blocks,
    # K split in 2 -> tables + 2 partial sums of P floats
    P = 32 * 512 * 512
    p = _mfma_plan(3, 4096, (32, 512, 512))
    assert p["swapped"] == 1 and p["row_tiles"] == 4 and p["col_tiles"] == 64 and p["ksplit"] == 2
    assert p["ws"] == a256(tables(512, 32, 512)) + 2 * P * 4 and p["kernel"] == 3
    # unswapped: rows 32 (one 128-row tile) x 1024 column tiles, no K split
    q = _mfma_plan(3, 4096, (32, 512, 512), variant=_native.MFMA_NO_SWAP)
    assert q["swapped"] == 0 and q["row_tiles"] == 1 and q["col_tiles"] == 1024 and q["ksplit"] == 1
    assert q["ws"] == tables(32, 512, 512)
    # (512, 512, 512) and a thin slab (512, 512, 8) keep axis 0 as rows
    cube = _mfma_plan(3, 4096, (512, 512, 512))
    assert cube["swapped"] == 0 and cube["ksplit"] == 1 and cube["ws"] == tables(512, 512, 512)
    slab = _mfma_plan(3, 4096, (512, 512, 8))
    ks = 512 // (4 * 16)  # 4 row tiles x 16 flat column tiles of 256 -> 64 blocks
    assert slab["kernel"] == 2 and slab["ksplit"] == ks
    assert slab["ws"] == a256(tables(512, 512, 8)) + ks * 512 * 512 * 8 * 4


def test_mfma_fill_targets_and_workspace_bound():
    """The K-split fill target scales the split (at least 128 modes per split);
    the generic workspace query covers every tuning choice."""
    wsq = _native.lib().gsx_summate_grid_mfma_workspace
    for shape in ((32, 512, 512), (512, 512, 8), (64, 64), (300, 200), (2048, 2048), (16, 16, 16)):
        d = len(shape)
        shp = (ctypes.c_int64 * 3)(*(list(shape) + [1] * (3 - d)))
        bound = wsq(d, 4096, shp)
        splits = []
        for f in (8, 9, 10, 11):
            for extra in (0, _native.MFMA_NO_SWAP):
                p = _mfma_plan(d, 4096, shape, variant=(f << _native.MFMA_FILL_SHIFT) | extra)
                assert p["ws"] <= bound
                assert 4096 // 16 // p["ksplit"] >= 8 or p["ksplit"] == 1
                if not extra:
                    splits.append(p["ksplit"])
        assert splits == sorted(splits)  # more resident blocks -> no smaller split
        assert _mfma_plan(d, 4096, shape) == _mfma_plan(d, 4096, shape, variant=9 << _native.MFMA_FILL_SHIFT)
    info = (ctypes.c_int64 * 8)()
    bad = (ctypes.c_int64 * 3)(64, 64, 1)
    assert _native.lib().gsx_summate_grid_mfma_plan(2, 4096, bad, 1, 12 << _native.MFMA_FILL_SHIFT, info) == -1


def test_fourier_tensor_workspace():
    """Workspace of the per-axis Fourier contraction: coefficients, rotation tables,
    intermediate and last-axis table, each 256-byte aligned (host arithmetic)."""
    q = _native.lib().gsx_summate_fourier_tensor_workspace
    c3 = lambda *s: (ctypes.c_int64 * 3)(*(list(s) + [1] * (3 - len(s))))
    al = lambda n: -(-n // 32) * 32
    # d = 2, modes (24, 20), grid (64, 300): c 2*480, E1 (2*64 x 2*24), T1 2*64*20, last 2*20*300
    assert q(2, c3(24, 20), c3(64, 300)) == 8 * (al(960) + al(6144) + al(2560) + al(12000))
    # d = 3, modes (8, 6, 10), grid (37, 40, 23)
    want = al(2 * 480) + al(2 * 37 * 2 * 8) + al(2 * 40 * 2 * 6) + al(2 * 37 * 60) + al(37 * 2 * 40 * 10) + al(2 * 10 * 23)
    assert q(3, c3(8, 6, 10), c3(37, 40, 23)) == 8 * want
    assert q(1, c3(10), c3(37)) == 8 * (al(20) + al(2 * 10 * 37))
    assert q(4, c3(1, 1, 1), c3(1, 1, 1)) < 0 and q(2, c3(-1, 2), c3(3, 3)) < 0
    assert q(2, c3(0, 3), c3(5, 5)) >= 0 and q(3, c3(4, 0, 2), c3(5, 0, 5)) >= 0  # no modes / points


def test_grid_gemm_workspace():
    """Workspace of the FP64 GEMM grid form: the A table, the phase tables and one
    column chunk of B (at most 512 MB, a multiple of 128 columns), 256-byte aligned."""
    q = _native.lib().gsx_summate_grid_gemm_workspace
    c3 = lambda *s: (ctypes.c_int64 * 3)(*(list(s) + [1] * (3 - len(s))))
    al = lambda n: -(-n // 32) * 32
    # 2D (64, 300), N 333: 300 columns fit one chunk
    assert q(2, 333, c3(64, 300)) == 8 * (al(64 * 2 * 333) + al(2 * 300 * 333) + al(2 * 333 * 300))
    # 3D (8, 100, 100), N 4096: chunks of 2^29 / (2 * 4096 * 8) = 8192 columns
    want = al(8 * 2 * 4096) + al(2 * 100 * 4096) + al(2 * 100 * 4096) + al(2 * 4096 * 8192)
    assert q(3, 4096, c3(8, 100, 100)) == 8 * want
    assert q(1, 10, c3(7)) == 8 * (al(7 * 20) + al(20) + al(20))
    assert q(0, 10, c3(7)) < 0 and q(2, -1, c3(3, 3)) < 0
