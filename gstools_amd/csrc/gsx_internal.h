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
