"""Allocation-free adjoint entry point against the workspace form (C ABI).

``gsx_summate_adjoint_f32_noalloc`` keeps its partial sums in a static
device array; whenever the workspace form's partial sums fit that array it
runs the same plan, so both must agree bit for bit.  Two calls in a row, the
second on a smaller input, show that the array's leftover contents from the
first call do not leak into the second.
"""

import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from gstools_amd import _native  # noqa: E402

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _case(d, n, p, seed):
    rs = np.random.RandomState(seed)
    cov = rs.normal(size=(d, n)) * 0.3
    z1 = rs.normal(size=n)
    z2 = rs.normal(size=n)
    pos = rs.uniform(0.0, 200.0, size=(d, p))
    g = rs.normal(size=p)
    dev = torch.device("cuda", 0)
    return tuple(torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (cov, z1, z2, pos)) + (
        torch.from_numpy(g).to(dev, torch.float32),)


def _outputs(d, n):
    dev = torch.device("cuda", 0)
    # filled with NaN so that an entry the call does not write shows up
    return (torch.full((n,), float("nan"), dtype=torch.float64, device=dev),
            torch.full((n,), float("nan"), dtype=torch.float64, device=dev),
            torch.full((d, n), float("nan"), dtype=torch.float64, device=dev))


def _with_workspace(lib, d, n, p, cov, z1, z2, pos, g):
    gz1, gz2, gk = _outputs(d, n)
    ws_bytes = int(lib.gsx_summate_adjoint_workspace(d, n, p))
    assert ws_bytes >= 0
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=cov.device)
    st = lib.gsx_summate_adjoint_f32(d, n, p, _ptr(cov), _ptr(z1), _ptr(z2), _ptr(pos), _ptr(g), _ptr(gz1),
                                     _ptr(gz2), _ptr(gk), _ptr(ws), ws_bytes, None)
    assert st == 0, lib.gsx_status_string(st)
    return gz1, gz2, gk


def _no_alloc(lib, d, n, p, cov, z1, z2, pos, g):
    gz1, gz2, gk = _outputs(d, n)
    st = lib.gsx_summate_adjoint_f32_noalloc(d, n, p, _ptr(cov), _ptr(z1), _ptr(z2), _ptr(pos), _ptr(g),
                                             _ptr(gz1), _ptr(gz2), _ptr(gk), None)
    assert st == 0, lib.gsx_status_string(st)
    return gz1, gz2, gk


@pytest.mark.parametrize("d", [1, 2, 3])
def test_no_alloc_adjoint_matches_workspace_form(d):
    lib = _native.lib()
    # second call smaller in both modes and points
    for n, p, seed in ((2048, 300_000, 11 + d), (700, 5_003, 29 + d)):
        print(f"no-alloc adjoint: d={d} N={n} P={p} seed={seed}", flush=True)
        args = _case(d, n, p, seed)
        ref = _with_workspace(lib, d, n, p, *args)
        got = _no_alloc(lib, d, n, p, *args)
        torch.cuda.synchronize()
        for r, o in zip(ref, got):
            r, o = r.cpu().numpy(), o.cpu().numpy()
            assert np.all(np.isfinite(o))
            np.testing.assert_array_equal(o, r)


def test_no_alloc_adjoint_capped_plan_close():
    """A call whose workspace-form partial sums exceed the array uses fewer
    point chunks: same gradients up to rounding (FP64 sums of FP32 partials)."""
    lib = _native.lib()
    d, n, p = 3, 300_000, 4096
    need = int(lib.gsx_summate_adjoint_workspace(d, n, p))
    assert need > _native.NOALLOC_SCRATCH_BYTES  # the capped plan is exercised
    print(f"no-alloc adjoint capped: d={d} N={n} P={p}", flush=True)
    args = _case(d, n, p, 7)
    ref = _with_workspace(lib, d, n, p, *args)
    got = _no_alloc(lib, d, n, p, *args)
    torch.cuda.synchronize()
    for r, o in zip(ref, got):
        r, o = r.cpu().numpy(), o.cpu().numpy()
        scale = np.max(np.abs(r))
        assert np.max(np.abs(o - r)) <= 1e-12 * scale * np.sqrt(p)


def test_no_alloc_adjoint_rejects_too_many_modes():
    """Sizes beyond the static array are refused on the host before any launch."""
    lib = _native.lib()
    d = 3
    n = _native.NOALLOC_SCRATCH_BYTES // ((2 + 2 * d) * 8) + 1
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda")
    st = lib.gsx_summate_adjoint_f32_noalloc(d, n, 10, _ptr(dummy), _ptr(dummy), _ptr(dummy), _ptr(dummy),
                                             _ptr(dummy), _ptr(dummy), _ptr(dummy), _ptr(dummy), None)
    assert st == _native.GSX_E_SCRATCH
    assert b"scratch" in lib.gsx_status_string(st)
