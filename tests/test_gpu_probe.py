"""The matrix-core ceiling probe (``gsx_probe_mfma``): it runs the MFMAs it
counts.  With near-constant operands every entry of an accumulator gains the
same closed-form amount per instruction, so each thread's checksum is known:
f64 (16x16x4): 4 a_q b_q' per instruction, summed over the 16 accumulators
(q = i mod 4, q' = i / 4 mod 4) and their 4 entries = 16 iters (sum a)(sum b);
bf16 (16x16x32): 4 sum_e a_e b_e per instruction (the 8 K elements of a lane
group pair with the same e in A and B)."""
import ctypes

import numpy as np
import pytest
import torch

from gstools_amd import _native

pytestmark = pytest.mark.gpu


def _run(kind, rnd, iters, blocks, dtype):
    out = torch.full((blocks * 256,), float("nan"), dtype=dtype, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _native.check(_native.lib().gsx_probe_mfma(kind, rnd, iters, blocks, ctypes.c_void_p(out.data_ptr()), stream),
                  "probe_mfma")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_probe_f64_constant_checksum():
    iters = 7
    got = _run(_native.GSX_PROBE_F64, 0, iters, 3, torch.float64)
    a = np.array([1.0 + 1e-3 * q for q in range(4)])
    b = np.array([1.0 - 1e-3 * q for q in range(4)])
    want = 16 * iters * a.sum() * b.sum()
    assert np.allclose(got, want, rtol=1e-13, atol=0)


def test_probe_bf16_constant_checksum():
    iters = 5
    got = _run(_native.GSX_PROBE_BF16, 0, iters, 3, torch.float32)
    bf = lambda x: torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).item()
    ab = sum(bf(1.0 + 1e-3 * e) * bf(1.0 - 1e-3 * e) for e in range(8))
    want = 16 * 4 * iters * 4 * ab
    assert np.allclose(got, want, rtol=1e-5, atol=0)


@pytest.mark.parametrize("kind", [_native.GSX_PROBE_BF16, _native.GSX_PROBE_F64])
def test_probe_random_operands_finite(kind):
    dtype = torch.float32 if kind == _native.GSX_PROBE_BF16 else torch.float64
    got = _run(kind, 1, 20, 4, dtype)
    assert np.isfinite(got).all() and np.abs(got).max() > 0
