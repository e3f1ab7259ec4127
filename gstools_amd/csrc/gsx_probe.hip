// Matrix-core ceiling probes: back-to-back independent MFMAs with no memory
// traffic in the loop, i.e. the rate the chip attains for a given operand
// pattern.  Constant operands leave the multipliers' inputs still; random
// ones toggle them as real data does, which draws more power and lowers the
// clock.  A measurement aid for the roofline figures of the GEMM-shaped
// kernels (the split-bf16 grid sums, the FP64 kriging and Fourier GEMMs); no
// part of any summation path.
#include "gsx_internal.h"

namespace gsx {

namespace {

typedef float probe_f32x4 __attribute__((ext_vector_type(4)));
typedef double probe_f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 probe_bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kProbeThreads = 256;  // four waves, one per SIMD; two blocks per CU give two waves per SIMD
constexpr int kProbeAcc = 16;       // independent accumulators per wave

__device__ __forceinline__ uint32_t probe_next(uint32_t& h) {
    h = h * 1664525u + 1013904223u;
    return h >> 8;
}

// v_mfma_f32_16x16x32_bf16: 16 x 16 x 32 x 2 = 16384 flops per instruction
template <bool RND>
__global__ __launch_bounds__(kProbeThreads, 2) void k_probe_bf16(int64_t iters, float* __restrict__ out) {
    probe_f32x4 acc[kProbeAcc];
#pragma unroll
    for (int i = 0; i < kProbeAcc; ++i) acc[i] = probe_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    probe_bf16x8 a[4], b[4];
    uint32_t h = (uint32_t)(blockIdx.x * kProbeThreads + threadIdx.x) * 2654435761u + 12345u;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float ra = RND ? (float)probe_next(h) / 16777216.0f - 0.5f : 1.0f + 1e-3f * e;
            const float rb = RND ? (float)probe_next(h) / 16777216.0f - 0.5f : 1.0f - 1e-3f * e;
            a[q][e] = (__bf16)ra;
            b[q][e] = (__bf16)rb;
        }
    for (int64_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < kProbeAcc; ++i)
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i & 3], b[(i >> 2) & 3], acc[i], 0, 0, 0);
    }
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < kProbeAcc; ++i) s += (acc[i][0] + acc[i][1]) + (acc[i][2] + acc[i][3]);
    out[(int64_t)blockIdx.x * kProbeThreads + threadIdx.x] = s;
}

// v_mfma_f64_16x16x4_f64: 16 x 16 x 4 x 2 = 2048 flops per instruction
template <bool RND>
__global__ __launch_bounds__(kProbeThreads, 2) void k_probe_f64(int64_t iters, double* __restrict__ out) {
    probe_f64x4 acc[kProbeAcc];
#pragma unroll
    for (int i = 0; i < kProbeAcc; ++i) acc[i] = probe_f64x4{0.0, 0.0, 0.0, 0.0};
    double a[4], b[4];
    uint32_t h = (uint32_t)(blockIdx.x * kProbeThreads + threadIdx.x) * 2654435761u + 777u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        a[q] = RND ? (double)probe_next(h) / 16777216.0 + (double)probe_next(h) * 5.9604644775390625e-8 / 16777216.0
                   : 1.0 + 1e-3 * q;
        b[q] = RND ? (double)probe_next(h) / 16777216.0 - (double)probe_next(h) * 5.9604644775390625e-8 / 16777216.0
                   : 1.0 - 1e-3 * q;
    }
    for (int64_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < kProbeAcc; ++i)
            acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i & 3], b[(i >> 2) & 3], acc[i], 0, 0, 0);
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kProbeAcc; ++i) s += (acc[i][0] + acc[i][1]) + (acc[i][2] + acc[i][3]);
    out[(int64_t)blockIdx.x * kProbeThreads + threadIdx.x] = s;
}

}  // namespace

}  // namespace gsx

extern "C" {

int gsx_probe_mfma(int kind, int random_operands, int64_t iters, int64_t blocks, void* out, void* stream) {
    if ((kind != GSX_PROBE_BF16 && kind != GSX_PROBE_F64) || iters < 1 || blocks < 1 || blocks > 0x7fffffffLL ||
        !out)
        return GSX_E_INVALID;
    hipStream_t s = gsx::as_stream(stream);
    const dim3 grid((unsigned)blocks), block(gsx::kProbeThreads);
    if (kind == GSX_PROBE_BF16) {
        float* o = static_cast<float*>(out);
        if (random_operands)
            hipLaunchKernelGGL(gsx::k_probe_bf16<true>, grid, block, 0, s, iters, o);
        else
            hipLaunchKernelGGL(gsx::k_probe_bf16<false>, grid, block, 0, s, iters, o);
    } else {
        double* o = static_cast<double*>(out);
        if (random_operands)
            hipLaunchKernelGGL(gsx::k_probe_f64<true>, grid, block, 0, s, iters, o);
        else
            hipLaunchKernelGGL(gsx::k_probe_f64<false>, grid, block, 0, s, iters, o);
    }
    return gsx::launch_status();
}

}  // extern "C"
