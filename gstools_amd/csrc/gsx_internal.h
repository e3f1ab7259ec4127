// Shared host/device plumbing for the gsx kernels (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gsx.h"
#include "gsx_math.h"

namespace gsx {

constexpr int kThreads = 256;  // threads per block for the point-parallel kernels

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Check the launch that was just enqueued.
inline int launch_status() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GSX_OK : GSX_E_LAUNCH;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Points of an unstructured (d, P) row-major array.
template <int D>
struct UnstructuredPoints {
    const double* __restrict__ pos;
    int64_t P;
    __device__ __forceinline__ void get(int64_t i, double* x) const {
#pragma unroll
        for (int d = 0; d < D; ++d) x[d] = pos[d * P + i];
    }
};

// Points of the 'ij' meshgrid of D axes; point i = (i0 * n1 + i1) * n2 + i2.
// Requires prod(shape) < 2^32 (checked on the host).
template <int D>
struct GridPoints {
    const double* __restrict__ ax0;
    const double* __restrict__ ax1;
    const double* __restrict__ ax2;
    uint32_t n1, n2;
    __device__ __forceinline__ void get(int64_t i, double* x) const {
        const uint32_t u = (uint32_t)i;
        if constexpr (D == 1) {
            x[0] = ax0[u];
        } else if constexpr (D == 2) {
            const uint32_t i0 = u / n1;
            x[0] = ax0[i0];
            x[1] = ax1[u - i0 * n1];
        } else {
            const uint32_t q = u / n2;
            const uint32_t i0 = q / n1;
            x[0] = ax0[i0];
            x[1] = ax1[q - i0 * n1];
            x[2] = ax2[u - q * n2];
        }
    }
};

// FP32 phase reduction by magic addition.  A phase chain started at
// kPhaseMagic + kPhaseRound - phi/2pi and summed by FMAs keeps a phase t (in
// revolutions, |t| < 2^19) inside [2^20, 2^21), where the FP64 ulp is 2^-32:
// the low 32 bits of the sum hold frac(t + 2^-24) in 32-bit fixed point.  Their
// top 23 bits under the exponent of 1.0 (one v_alignbit_b32) are the float
// 1 + frac(t) rounded to 2^-23, and since v_cos_f32 / v_sin_f32 take
// revolutions, cos(2 pi (1 + f)) = cos(2 pi t).  This replaces v_fract_f64 and
// v_cvt_f32_f64 by one integer op; the phase error is at most 6e-8 revolutions
// (3e-8 for the fract-and-convert form).  Callers check |k/2pi|_1 |x|_inf <=
// kPhaseMagicMax per chunk of modes and otherwise use the fract form.
constexpr double kPhaseMagic = 0x1.8p+20;
constexpr double kPhaseRound = 0x1.0p-24;
constexpr double kPhaseMagicMax = 0x1.0p+18;

__device__ __forceinline__ float magic_rev_to_unit(double rev) {
    const uint32_t lo = (uint32_t)__double_as_longlong(rev);
    return __uint_as_float(__builtin_amdgcn_alignbit(0x7Fu, lo, 9u));
}

// Running bound max(xm, |v|), infinite once v is NaN or infinite.
__device__ __forceinline__ double coord_bound(double xm, double v) {
    const double a = fabs(v);
    return a <= 0x1.fffffffffffffp+1023 ? fmax(xm, a) : __builtin_inf();
}

// Phase k.x accumulated in dimension order with every product rounded
// before the add, as a plain C loop without fused multiply-add computes it.
template <int D>
__device__ __forceinline__ double phase_exact(const double* k, const double* x) {
#pragma clang fp contract(off)
    double t = k[0] * x[0];
#pragma unroll
    for (int d = 1; d < D; ++d) t = t + k[d] * x[d];
    return t;
}

}  // namespace gsx
