// Double-precision trigonometric kernels for the randomisation-method sum.
//
// Written from the mathematics, not from any vendor math library:
//  * argument reduction: two-term Cody-Waite with fused multiply-add,
//      n = rint(x / c), r = fma(-n, c_hi, x); r = fma(-n, c_lo, r),
//    where c_hi + c_lo represents c (pi or pi/2) to ~2^-107 relative.  Each
//    fma rounds the exact value once, so |r - (x - n c)| <= 2 * 2^-53 * |r| +
//    n * 2^-107 c: below 3e-16 absolute for |x| < 2^30.
//  * polynomials: Chebyshev least-squares fits (mpmath.chebyfit, 60 digits)
//    of cos(sqrt(s)) and sin(sqrt(s))/sqrt(s) in s = r^2; the fit errors are
//    printed next to each table and sit far below double rounding.
//
// The header compiles both for gfx950 (hipcc) and for the host (g++), so
// the unit tests check exactly this code on the CPU against mpmath.
#pragma once
#include <stdint.h>
#include <string.h>

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
// Fast-path bound for the reduced phase (|x| below this keeps n < 2^29).
constexpr double kFastPhaseMax = 0x1.0p+30;

// cos(r) on |r| <= 1.5715817 (pi/2 plus slack); degree 8 in s = r^2,
// fit error 3.96e-18.
constexpr double kCosHalfPi[9] = {
    0x1.0000000000000p+0,  -0x1.ffffffffffffbp-2, 0x1.55555555553c1p-5,
    -0x1.6c16c16c09a0ep-10, 0x1.a01a019943860p-16, -0x1.27e4f9772ac4fp-22,
    0x1.1eecdeb37c5fbp-29, -0x1.9350806cfca59p-37, 0x1.9f1fe88a8d46cp-45,
};
// cos(r) on |r| <= 0.78618356 (pi/4 plus slack); degree 7, fit error 3.08e-20.
constexpr double kCosQuarterPi[8] = {
    0x1.0000000000000p+0,  -0x1.0000000000000p-1, 0x1.5555555555536p-5,
    -0x1.6c16c16c13984p-10, 0x1.a01a019b13ebbp-16, -0x1.27e4f72151b96p-22,
    0x1.1ee968c8051d5p-29, -0x1.8f740c23980c3p-37,
};
// sin(r)/r on |r| <= 0.78618356; degree 6, fit error 3.19e-18.
constexpr double kSinQuarterPi[7] = {
    0x1.0000000000000p+0,  -0x1.5555555555543p-3, 0x1.111111110f391p-7,
    -0x1.a01a019ba03b3p-13, 0x1.71de3528c352cp-19, -0x1.ae5e31a6de003p-26,
    0x1.5d859215ca078p-33,
};

GSX_HD double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <int NC>
GSX_HD double horner(const double (&c)[NC], double s) {
    double p = c[NC - 1];
#pragma unroll
    for (int i = NC - 2; i >= 0; --i) p = fma_(p, s, c[i]);
    return p;
}

GSX_HD double flip_sign_if(double v, uint32_t odd) {
    uint64_t b;
    memcpy(&b, &v, 8);
    b ^= (uint64_t)(odd & 1u) << 63;
    memcpy(&v, &b, 8);
    return v;
}

// cos(theta - phi) for |theta| < kFastPhaseMax, |phi| <= pi, with
// phi_over_pi = phi / pi precomputed:  theta - phi = n pi + r,
// |r| <= pi/2 (+ slack), cos(theta - phi) = (-1)^n cos(r).
GSX_HD double cos_shifted(double theta, double phi, double phi_over_pi) {
    const double n = __builtin_rint(fma_(theta, kInvPi, -phi_over_pi));
    double r = fma_(-n, kPiHi, theta);
    r = fma_(-n, kPiLo, r);
    r = r - phi;
    const double c = horner(kCosHalfPi, r * r);
    return flip_sign_if(c, (uint32_t)(int32_t)n);
}

// sin(x) and cos(x) for |x| < kFastPhaseMax.
GSX_HD void sincos_fast(double x, double* s_out, double* c_out) {
    const double n = __builtin_rint(x * kTwoOverPi);
    double r = fma_(-n, kPio2Hi, x);
    r = fma_(-n, kPio2Lo, r);
    const double s2 = r * r;
    const double sn = r * horner(kSinQuarterPi, s2);
    const double cs = horner(kCosQuarterPi, s2);
    const uint32_t q = (uint32_t)(int32_t)n;
    const double sv = (q & 1u) ? cs : sn;
    const double cv = (q & 1u) ? sn : cs;
    *s_out = flip_sign_if(sv, (q >> 1) & 1u);
    *c_out = flip_sign_if(cv, ((q + 1u) >> 1) & 1u);
}

}  // namespace gsx
