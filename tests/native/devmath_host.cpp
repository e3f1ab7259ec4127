// Host build of the device trigonometric kernels (gsx_math.h) so that the
// CPU tests can check the exact reduction + polynomial code against mpmath.
#include "../../gstools_amd/csrc/gsx_math.h"

extern "C" {
void devmath_cos_shifted(const double* theta, const double* phi, double* out, long n) {
    for (long i = 0; i < n; ++i) out[i] = gsx::cos_shifted(theta[i], phi[i], phi[i] * gsx::kInvPi);
}
void devmath_sincos(const double* x, double* s, double* c, long n) {
    for (long i = 0; i < n; ++i) gsx::sincos_fast(x[i], &s[i], &c[i]);
}
}
