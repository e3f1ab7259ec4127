// Double-precision trigonometric kernels for the randomisation-method sum.
//
// Written from the mathematics, not from any vendor math library:
//  * argument reduction: two-term Cody-Waite with fused multiply-add,
//      n = rint(x / c), r = fma(-n, c_hi, x); r = fma(-n, c_lo, r),
//    where c_hi + c_lo represents c (pi or pi/2) to ~2^-107 relative.  Each
//    fma rounds the exact value once, so |r - (x - n c)| <= 2 * 2^-53 * |r| +
//    n * 2^-107 c: below 3e-16 absolute for |x| <= 2^46.  n and its low bits
//    (quadrant / parity) come from one add of 1.5 * 2^52.
//  * polynomials: Chebyshev least-squares fits (mpmath.chebyfit, 60 digits)
//    of cos(sqrt(s)) and sin(sqrt(s))/sqrt(s) in s = r^2; the fit errors are
//    printed next to each table and sit far below double rounding.
//
// The header compiles both for gfx950 (hipcc) and for the host (g++), so
// the unit tests check exactly this code on the CPU against mpmath.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define GSX_HD __host__ __device__ __forceinline__
#else
#define GSX_HD static inline
#endif

namespace gsx {

// pi = kPiHi + kPiLo, pi/2 = kPio2Hi + kPio2Lo (hi = nearest double).
constexpr double kPiHi = 0x1.921fb54442d18p+1;
constexpr double kPiLo = 0x1.1a62633145c07p-53;
constexpr double kPio2Hi = 0x1.921fb54442d18p+0;
constexpr double kPio2Lo = 0x1.1a62633145c07p-54;
constexpr double kInvPi = 0x1.45f306dc9c883p-2;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kInv2Pi = 0x1.45f306dc9c883p-3;
// Adding kRound = 1.5 * 2^52 rounds |t| < 2^51 to the nearest integer (ties to
// even) and leaves that integer's low bits in the low word of the sum.
constexpr double kRound = 0x1.8p+52;
// Fast-path bound: for |x| <= 2^46 the rounding error of x / c moves the reduced
// argument at most 0.4% of c beyond the polynomial ranges' slack (1-2%), and
// the dropped term n * (c - c_hi - c_lo) stays below 1e-18.
constexpr double kFastPhaseMax = 0x1.0p+46;

// cos(r) on |r| <= 1.5865043 (pi/2 plus 1%); degree 8 in s = r^2,
// fit error 4.69e-18.
constexpr double kCosHalfPi[9] = {
    0x1.0000000000000p+0,  -0x1.ffffffffffffbp-2, 0x1.5555555555388p-5,
    -0x1.6c16c16c080f6p-10, 0x1.a01a0198970dbp-16, -0x1.27e4f94f36716p-22,
    0x1.1eecd48195b3bp-29, -0x1.934dce64c0cd3p-37, 0x1.9ed61e7ca43a8p-45,
};
// cos(r) on |r| <= 0.80110613 (pi/4 plus 2%); degree 7, fit error 4.16e-20.
constexpr double kCosQuarterPi[8] = {
    0x1.0000000000000p+0,  -0x1.0000000000000p-1, 0x1.555555555552ep-5,
    -0x1.6c16c16c12f13p-10, 0x1.a01a019a43961p-16, -0x1.27e4f69d11024p-22,
    0x1.1ee916166af1ap-29, -0x1.8f4ba81ba52f1p-37,
};
// sin(r)/r on |r| <= 0.80110613; degree 6, fit error 4.15e-18.
constexpr double kSinQuarterPi[7] = {
    0x1.0000000000000p+0,  -0x1.555555555553ep-3, 0x1.111111110ed79p-7,
    -0x1.a01a019ae6a05p-13, 0x1.71de348af77c0p-19, -0x1.ae5db7352c71fp-26,
    0x1.5d623b237043fp-33,
};

GSX_HD double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <int NC>
GSX_HD double horner(const double (&c)[NC], double s) {
    double p = c[NC - 1];
#pragma unroll
    for (int i = NC - 2; i >= 0; --i) p = fma_(p, s, c[i]);
    return p;
}

GSX_HD uint32_t lo_word(double v) { return (uint32_t)__builtin_bit_cast(uint64_t, v); }

// Flip the sign of v when bit 0 of `odd` is set (one 32-bit XOR of the high word).
GSX_HD double flip_sign_if(double v, uint32_t odd) {
    uint64_t b = __builtin_bit_cast(uint64_t, v);
    b ^= (uint64_t)(odd << 31) << 32;
    return __builtin_bit_cast(double, b);
}

// cos(theta - phi) for |theta| <= kFastPhaseMax, |phi| <= pi, with
// phi_over_pi = phi / pi precomputed:  theta - phi = n pi + r,
// |r| <= pi/2 (+ slack), cos(theta - phi) = (-1)^n cos(r).
GSX_HD double cos_shifted(double theta, double phi, double phi_over_pi) {
    const double tm = fma_(theta, kInvPi, -phi_over_pi) + kRound;
    const double n = tm - kRound;
    double r = fma_(-n, kPiHi, theta);
    r = fma_(-n, kPiLo, r);
    r = r - phi;
    const double c = horner(kCosHalfPi, r * r);
    return flip_sign_if(c, lo_word(tm));
}

// sin(x) and cos(x) for |x| <= kFastPhaseMax.
GSX_HD void sincos_fast(double x, double* s_out, double* c_out) {
    const double tm = fma_(x, kTwoOverPi, kRound);
    const double n = tm - kRound;
    double r = fma_(-n, kPio2Hi, x);
    r = fma_(-n, kPio2Lo, r);
    const double s2 = r * r;
    const double sn = r * horner(kSinQuarterPi, s2);
    const double cs = horner(kCosQuarterPi, s2);
    const uint32_t q = lo_word(tm);
    const double sv = (q & 1u) ? cs : sn;
    const double cv = (q & 1u) ? sn : cs;
    *s_out = flip_sign_if(sv, q >> 1);
    *c_out = flip_sign_if(cv, (q + 1u) >> 1);
}

}  // namespace gsx
