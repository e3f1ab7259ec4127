"""CPU checks of the FP32 kernels' phase reduction by magic addition.

The FP32 sum starts its FP64 phase chain (revolutions) at
M + R - phi/2pi, M = 1.5 * 2^20, R = 2^-24, so that for |t| < 2^19 the result
lies in [2^20, 2^21), where the FP64 ulp is 2^-32; the low 32 bits are then
frac(t + R) in fixed point, and one v_alignbit_b32(0x7F, lo, 9) makes the
float 1 + f with f = frac(t) rounded to 2^-23, whose cosine in revolutions is
cos(2 pi t).  These checks restate that bit layout with numpy integer
operations (the same shifts and masks, no kernel) and test it against what
the mathematics fixes: the phase error bound, the absence of bias, the range
limit (one binade past it the layout breaks), and the guard that sends a
mode to the fract form.
"""

import numpy as np
import pytest

M = 1.5 * 2.0**20  # kPhaseMagic
R = 2.0**-24       # kPhaseRound
GUARD = 2.0**18    # kPhaseMagicMax


def _magic_unit(t):
    """t (revolutions, FP64) -> the float32 the kernels hand to v_cos_f32."""
    rev = (M + R) + t  # the chain's last rounding, at ulp 2^-32 inside the binade
    lo = rev.view(np.uint64).astype(np.uint32)
    bits = np.uint32(0x7F << 23) | (lo >> np.uint32(9))  # v_alignbit_b32(0x7F, lo, 9)
    return bits.view(np.float32)


def _wrap(d):
    return d - np.round(d)


def test_magic_unit_is_one_plus_the_fraction():
    rs = np.random.RandomState(5)
    t = np.concatenate([rs.uniform(-2.0**19 + 1, 2.0**19 - 1, 200_000), rs.uniform(-4, 4, 50_000),
                        np.arange(-64, 65, dtype=np.float64)])
    f = _magic_unit(t).astype(np.float64)
    assert np.all((f >= 1.0) & (f < 2.0))
    err = _wrap((f - 1.0) - t)  # phase error in revolutions, modulo 1
    # rounding to 2^-23 plus the chain's 2^-33 rounding
    assert np.abs(err).max() <= 2.0**-24 + 2.0**-32
    assert abs(err.mean()) < 2e-9  # round to nearest: no bias
    # the cosine it feeds equals cos(2 pi t) to the phase error
    assert np.abs(np.cos(2 * np.pi * f) - np.cos(2 * np.pi * t)).max() <= 2 * np.pi * (2.0**-24 + 2.0**-32)


def test_integer_phases_land_on_one():
    t = np.arange(-1000.0, 1001.0)
    f = _magic_unit(t)
    assert np.all(f == np.float32(1.0))  # frac 0 -> exactly 1 + 0
    half = _magic_unit(t + 0.5)
    assert np.all(half == np.float32(1.5))


def test_layout_breaks_one_binade_out():
    """Past |t| = 2^19 the sum leaves [2^20, 2^21) and the low word no longer
    holds the fraction: this is what the 2^18 guard keeps away."""
    t = np.array([2.0**19 + 0.25, 2.0**20 + 0.25, -(2.0**19) - 0.25])
    f = _magic_unit(t).astype(np.float64)
    assert np.all(np.abs(_wrap((f - 1.0) - t)) > 1e-3)
    inside = np.array([2.0**19 - 0.75, -(2.0**19) + 0.75])
    g = _magic_unit(inside).astype(np.float64)
    assert np.all(np.abs(_wrap((g - 1.0) - inside)) <= 2.0**-24 + 2.0**-32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_guard_bounds_every_phase(seed):
    """A mode passes the guard when |k/2pi|_1 times the segment's largest
    |x|_inf is within 2^18; every phase k.x/2pi - phi/2pi of such a mode
    over the segment then stays inside the magic binade."""
    rs = np.random.RandomState(seed)
    x = rs.uniform(-1.2e4, 1.2e4, size=(3, 2048))
    k = rs.standard_cauchy(size=(3, 512)) * 2.0  # heavy tail: some modes fail the guard
    kr = k / (2 * np.pi)
    xseg = np.abs(x).max()
    ok = np.abs(kr).sum(axis=0) * xseg <= GUARD
    assert 0 < ok.sum() < ok.size
    t = kr[:, ok].T @ x - rs.uniform(-0.5, 0.5, size=(int(ok.sum()), 1))
    assert np.abs(t).max() < 2.0**19 - 1
