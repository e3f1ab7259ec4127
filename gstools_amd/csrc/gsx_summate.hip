// Point-parallel randomisation-method summation on gfx950.
//
//   u(x_i) = sum_j z1_j cos(k_j.x_i) + z2_j sin(k_j.x_i)        (summate)
//   u_c(x_i) = sum_j p_c(k_j) (z1_j cos + z2_j sin)             (summate_incompr)
//
// Reference contract: src/gstools/field/generator.py:193-199 (scalar) and
// :479-494 (incompressible), called at :266-268 and :554-560.
//
// Layout and schedule:
//  * one thread owns PPT points (coalesced loads of the (d, P) position rows,
//    points interleaved by kThreads so every load/store is unit-stride across
//    a wavefront);
//  * modes are staged through LDS in chunks of kChunk: each block turns its
//    chunk of (cov, z1, z2) into a per-mode table -- wavenumbers, amplitude
//    a_j = |z1_j - i z2_j| and phase phi_j = atan2(z2_j, z1_j), so that
//    z1 cos(t) + z2 sin(t) = a cos(t - phi) needs ONE cosine per term;
//  * every lane reads the same table entry (LDS broadcast, no bank conflict).
//
// The Fourier method (generator.py:659-691) is the same sum with a per-mode
// spectrum factor f_j multiplying each term; it is folded into the amplitude
// when the table is built (fac == nullptr for the randomisation method).
//
// FP64 path: the phase is the reference's exact product-then-add sequence;
// cos(t - phi) uses the two-term Cody-Waite reduction of gsx_math.h with phi
// folded into the reduction (no rounding of t - phi at large |t|).  A wave
// whose |x|_inf * max_j |k_j|_1 may exceed 2^45 takes a per-term guarded path
// that falls back to the library sincos for |t| > 2^46.
//
// FP32 path (gsx_fp32sum.h): the phase is computed in revolutions in FP64
// (t = k/2pi . x - phi/2pi, three FMAs), reduced by magic addition (the low
// word of the sum is the fraction in fixed point, one v_alignbit_b32 makes it
// a float in [1, 2)) for the modes whose |k/2pi|_1 times the largest |x|_inf
// of the point's 2048-point segment stays within 2^18 revolutions, else
// exactly with v_fract_f64 and a conversion, and fed to the hardware cosine
// (v_cos_f32 takes revolutions).  The mode sum
// runs in FP32 for blocks of kSub modes and is then added into an FP64
// accumulator (blocked compensation: FP32 rounding error grows only over
// kSub terms).
#include "gsx_fp32sum.h"
#include "gsx_internal.h"

namespace gsx {

constexpr int kChunk = 256;  // modes per LDS chunk (== kThreads: one mode per thread to build)
constexpr double kSafeK1X = 0x1.0p+45;  // |k|_1 |x|_inf bound of the FP64 fast path

// cos(theta - phi) at any |theta|: the fast reduction where it is exact, else the
// library sincos.  Kept out of line so the fast loop's register budget does not
// pay for the wide-range code.
__device__ __noinline__ double cos_shifted_any(double th, double phi, double phipi) {
    if (fabs(th) <= kFastPhaseMax) return cos_shifted(th, phi, phipi);
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(phi, &sp, &cp);
    return ct * cp + st * sp;
}

__device__ __forceinline__ void lds_max_u64(unsigned long long* p, double v) {
    // non-negative doubles order like their bit patterns
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// -------------------------------------------------------------- FP64 --------
template <int D, bool INCOMPR>
struct TabF64 {
    // scalar:  k[D], a, phi, phi/pi ; incompr: k[D], phi, phi/pi, w[D] = p(k) * a
    static constexpr int kW = INCOMPR ? 2 * D + 2 : D + 3;
    static constexpr int kStride = (kW + 1) & ~1;  // 16-byte aligned rows
};

template <int D, int PPT, bool INCOMPR, class Pts>
__global__ __launch_bounds__(kThreads) void k_forward_f64(Pts pts, int64_t N, int64_t P,
                                                          const double* __restrict__ cov,
                                                          const double* __restrict__ z1,
                                                          const double* __restrict__ z2,
                                                          const double* __restrict__ fac,
                                                          double* __restrict__ out) {
    using T = TabF64<D, INCOMPR>;
    constexpr int NOUT = INCOMPR ? D : 1;
    __shared__ __align__(16) double tab[kChunk * T::kStride];
    __shared__ unsigned long long s_kmax;

    double x[PPT][D];
    double acc[PPT][NOUT];
    double xmax[PPT];
    const int64_t base = (int64_t)blockIdx.x * (kThreads * PPT) + threadIdx.x;
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
        const int64_t i = base + (int64_t)p * kThreads;
        if (i < P) {
            pts.get(i, x[p]);
        } else {
#pragma unroll
            for (int d = 0; d < D; ++d) x[p][d] = 0.0;
        }
        xmax[p] = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) xmax[p] = fmax(xmax[p], fabs(x[p][d]));
#pragma unroll
        for (int c = 0; c < NOUT; ++c) acc[p][c] = 0.0;
    }

    for (int64_t j0 = 0; j0 < N; j0 += kChunk) {
        const int cnt = (int)(N - j0 < kChunk ? N - j0 : kChunk);
        __syncthreads();
        if (threadIdx.x == 0) s_kmax = 0ull;
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            const int64_t j = j0 + threadIdx.x;
            double* t = tab + threadIdx.x * T::kStride;
            double k[D];
            double k1 = 0.0, k2 = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                k[d] = cov[d * N + j];
                t[d] = k[d];
                k1 += fabs(k[d]);
                k2 += k[d] * k[d];
            }
            const double u1 = z1[j], u2 = z2[j];
            const double phi = atan2(u2, u1);
            const double a = fac ? fac[j] * hypot(u1, u2) : hypot(u1, u2);
            if constexpr (INCOMPR) {
                t[D] = phi;
                t[D + 1] = phi * kInvPi;
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    const double proj = (c == 0 ? 1.0 : 0.0) - k[c] * k[0] / k2;
                    t[D + 2 + c] = proj * a;
                }
            } else {
                t[D] = a;
                t[D + 1] = phi;
                t[D + 2] = phi * kInvPi;
            }
            lds_max_u64(&s_kmax, k1);
        }
        __syncthreads();
        const double kmax = __longlong_as_double((long long)s_kmax);
        bool safe = true;
#pragma unroll
        for (int p = 0; p < PPT; ++p) safe = safe && (xmax[p] * kmax <= kSafeK1X);

        if (__all(safe)) {
            for (int m = 0; m < cnt; ++m) {
                const double* t = tab + m * T::kStride;
                double k[D];
#pragma unroll
                for (int d = 0; d < D; ++d) k[d] = t[d];
                if constexpr (INCOMPR) {
                    const double phi = t[D], phipi = t[D + 1];
                    double w[D];
#pragma unroll
                    for (int c = 0; c < D; ++c) w[c] = t[D + 2 + c];
#pragma unroll
                    for (int p = 0; p < PPT; ++p) {
                        const double cs = cos_shifted(phase_exact<D>(k, x[p]), phi, phipi);
#pragma unroll
                        for (int c = 0; c < D; ++c) acc[p][c] = fma_(w[c], cs, acc[p][c]);
                    }
                } else {
                    const double a = t[D], phi = t[D + 1], phipi = t[D + 2];
#pragma unroll
                    for (int p = 0; p < PPT; ++p)
                        acc[p][0] = fma_(a, cos_shifted(phase_exact<D>(k, x[p]), phi, phipi), acc[p][0]);
                }
            }
        } else {
            // Guarded path: library sincos for phases beyond the fast reduction.
            for (int m = 0; m < cnt; ++m) {
                const double* t = tab + m * T::kStride;
                double k[D];
#pragma unroll
                for (int d = 0; d < D; ++d) k[d] = t[d];
                const double phi = INCOMPR ? t[D] : t[D + 1];
                const double phipi = INCOMPR ? t[D + 1] : t[D + 2];
                for (int p = 0; p < PPT; ++p) {
                    const double cs = cos_shifted_any(phase_exact<D>(k, x[p]), phi, phipi);
                    if constexpr (INCOMPR) {
                        for (int c = 0; c < D; ++c) acc[p][c] = fma_(t[D + 2 + c], cs, acc[p][c]);
                    } else {
                        acc[p][0] = fma_(t[D], cs, acc[p][0]);
                    }
                }
            }
        }
    }

#pragma unroll
    for (int p = 0; p < PPT; ++p) {
        const int64_t i = base + (int64_t)p * kThreads;
        if (i < P) {
#pragma unroll
            for (int c = 0; c < NOUT; ++c) out[c * P + i] = acc[p][c];
        }
    }
}

// -------------------------------------------------------------- FP32 --------
// One block: kThreads * PPT points (a divisor of the kSeg-point segment), all
// modes through fp32_item_sum (gsx_fp32sum.h).  At least 3 waves per SIMD: the
// two inlined mode loops would otherwise take the 3D sum to 198 VGPRs (2
// waves); bounded, it fits in 168.
template <int D, int PPT, bool INCOMPR, class Pts>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_forward_f32(
    Pts pts, int64_t N, int64_t P, const double* __restrict__ cov, const double* __restrict__ z1,
    const double* __restrict__ z2, const double* __restrict__ fac, float* __restrict__ out) {
    constexpr int NOUT = INCOMPR ? D : 1;
    __shared__ double tabd[kChunk * (D + 1)];
    __shared__ float tabf[kChunk * NOUT];  // a (scalar) or p(k) * a (incompr)
    __shared__ double s_wmax[kWaves];
    __shared__ int s_nmagic[kWaves];

    double x[PPT][D];
    double acc[PPT][NOUT];
    double own = 0.0;
    const int64_t p0 = (int64_t)blockIdx.x * (kThreads * PPT);
    const int64_t base = p0 + threadIdx.x;
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
        const int64_t i = base + (int64_t)p * kThreads;
        if (i < P) {
            pts.get(i, x[p]);
#pragma unroll
            for (int d = 0; d < D; ++d) own = coord_bound(own, x[p][d]);
        } else {
#pragma unroll
            for (int d = 0; d < D; ++d) x[p][d] = 0.0;
        }
#pragma unroll
        for (int c = 0; c < NOUT; ++c) acc[p][c] = 0.0;
    }
    const double xseg = segment_xmax<D, PPT>(pts, P, p0, own, s_wmax);
    fp32_item_sum<D, PPT, INCOMPR, kChunk>(N, cov, z1, z2, fac, xseg, x, acc, tabd, tabf, s_nmagic);

#pragma unroll
    for (int p = 0; p < PPT; ++p) {
        const int64_t i = base + (int64_t)p * kThreads;
        if (i < P) {
#pragma unroll
            for (int c = 0; c < NOUT; ++c) out[c * P + i] = (float)acc[p][c];
        }
    }
}

// ------------------------------------------------------------ launchers -----
constexpr int kPptF64 = 2;
// FP32 points per thread: 8 for the scalar sum (the per-block mode tables and
// LDS reads spread over twice the terms: 1.5-2% faster than 4 on the 4e6-point
// 3D sum, at 152 VGPRs); 4 for the incompressible sum, whose d accumulators
// per point would take 8 points to 228 VGPRs (2 waves per SIMD)
template <bool INCOMPR>
constexpr int ppt_f32() { return INCOMPR ? 4 : 8; }

template <int D, bool INCOMPR, class Pts, class OutT>
int launch_forward(Pts pts, int64_t N, int64_t P, const double* cov, const double* z1,
                   const double* z2, const double* fac, OutT* out, hipStream_t s) {
    if constexpr (sizeof(OutT) == 8) {
        const int64_t blocks = ceil_div(P, (int64_t)kThreads * kPptF64);
        hipLaunchKernelGGL((k_forward_f64<D, kPptF64, INCOMPR, Pts>), dim3((unsigned)blocks),
                           dim3(kThreads), 0, s, pts, N, P, cov, z1, z2, fac, out);
    } else {
        constexpr int ppt = ppt_f32<INCOMPR>();
        const int64_t blocks = ceil_div(P, (int64_t)kThreads * ppt);
        hipLaunchKernelGGL((k_forward_f32<D, ppt, INCOMPR, Pts>), dim3((unsigned)blocks),
                           dim3(kThreads), 0, s, pts, N, P, cov, z1, z2, fac, out);
    }
    return launch_status();
}

template <bool INCOMPR, class OutT>
int forward_unstructured(int64_t d, int64_t N, int64_t P, const double* cov, const double* z1,
                         const double* z2, const double* fac, const double* pos, OutT* out,
                         void* stream) {
    if (d < (INCOMPR ? 2 : 1) || d > 3 || N < 0 || P < 0) return GSX_E_INVALID;
    if (P == 0) return GSX_OK;
    if (!pos || !out || (N > 0 && (!cov || !z1 || !z2))) return GSX_E_INVALID;
    if (P > (int64_t)0x7fffffff * kThreads) return GSX_E_INVALID;
    hipStream_t s = as_stream(stream);
    switch (d) {
        case 1: return launch_forward<1, INCOMPR>(UnstructuredPoints<1>{pos, P}, N, P, cov, z1, z2, fac, out, s);
        case 2: return launch_forward<2, INCOMPR>(UnstructuredPoints<2>{pos, P}, N, P, cov, z1, z2, fac, out, s);
        default: return launch_forward<3, INCOMPR>(UnstructuredPoints<3>{pos, P}, N, P, cov, z1, z2, fac, out, s);
    }
}

template <class OutT, bool INCOMPR = false>
int forward_grid(int64_t d, int64_t N, const int64_t* shape, const double* axes, const double* cov,
                 const double* z1, const double* z2, const double* fac, OutT* out, void* stream) {
    if (d < (INCOMPR ? 2 : 1) || d > 3 || N < 0 || !shape) return GSX_E_INVALID;
    int64_t P = 1;
    for (int i = 0; i < d; ++i) {
        if (shape[i] < 0) return GSX_E_INVALID;
        P *= shape[i];
    }
    if (P == 0) return GSX_OK;
    if (P >= ((int64_t)1 << 32)) return GSX_E_INVALID;
    if (!axes || !out || (N > 0 && (!cov || !z1 || !z2))) return GSX_E_INVALID;
    hipStream_t s = as_stream(stream);
    const double* a0 = axes;
    const double* a1 = d > 1 ? a0 + shape[0] : nullptr;
    const double* a2 = d > 2 ? a1 + shape[1] : nullptr;
    const uint32_t n1 = d > 1 ? (uint32_t)shape[1] : 1u;
    const uint32_t n2 = d > 2 ? (uint32_t)shape[2] : 1u;
    if constexpr (INCOMPR) {
        if (d == 2) return launch_forward<2, true>(GridPoints<2>{a0, a1, a2, n1, n2}, N, P, cov, z1, z2, fac, out, s);
        return launch_forward<3, true>(GridPoints<3>{a0, a1, a2, n1, n2}, N, P, cov, z1, z2, fac, out, s);
    } else {
        switch (d) {
            case 1:
                return launch_forward<1, false>(GridPoints<1>{a0, a1, a2, n1, n2}, N, P, cov, z1, z2, fac, out, s);
            case 2:
                return launch_forward<2, false>(GridPoints<2>{a0, a1, a2, n1, n2}, N, P, cov, z1, z2, fac, out, s);
            default:
                return launch_forward<3, false>(GridPoints<3>{a0, a1, a2, n1, n2}, N, P, cov, z1, z2, fac, out, s);
        }
    }
}

}  // namespace gsx

extern "C" {

int gsx_summate_f64(int64_t d, int64_t N, int64_t P, const double* cov, const double* z1,
                    const double* z2, const double* pos, double* out, void* stream) {
    return gsx::forward_unstructured<false>(d, N, P, cov, z1, z2, nullptr, pos, out, stream);
}

int gsx_summate_f32(int64_t d, int64_t N, int64_t P, const double* cov, const double* z1,
                    const double* z2, const double* pos, float* out, void* stream) {
    return gsx::forward_unstructured<false>(d, N, P, cov, z1, z2, nullptr, pos, out, stream);
}

int gsx_summate_incompr_f64(int64_t d, int64_t N, int64_t P, const double* cov, const double* z1,
                            const double* z2, const double* pos, double* out, void* stream) {
    return gsx::forward_unstructured<true>(d, N, P, cov, z1, z2, nullptr, pos, out, stream);
}

int gsx_summate_incompr_f32(int64_t d, int64_t N, int64_t P, const double* cov, const double* z1,
                            const double* z2, const double* pos, float* out, void* stream) {
    return gsx::forward_unstructured<true>(d, N, P, cov, z1, z2, nullptr, pos, out, stream);
}

int gsx_summate_grid_f64(int64_t d, int64_t N, const int64_t* shape, const double* axes,
                         const double* cov, const double* z1, const double* z2, double* out,
                         void* stream) {
    return gsx::forward_grid(d, N, shape, axes, cov, z1, z2, nullptr, out, stream);
}

int gsx_summate_grid_f32(int64_t d, int64_t N, const int64_t* shape, const double* axes,
                         const double* cov, const double* z1, const double* z2, float* out,
                         void* stream) {
    return gsx::forward_grid(d, N, shape, axes, cov, z1, z2, nullptr, out, stream);
}

int gsx_summate_incompr_grid_f64(int64_t d, int64_t N, const int64_t* shape, const double* axes,
                                 const double* cov, const double* z1, const double* z2, double* out,
                                 void* stream) {
    return gsx::forward_grid<double, true>(d, N, shape, axes, cov, z1, z2, nullptr, out, stream);
}

int gsx_summate_incompr_grid_f32(int64_t d, int64_t N, const int64_t* shape, const double* axes,
                                 const double* cov, const double* z1, const double* z2, float* out,
                                 void* stream) {
    return gsx::forward_grid<float, true>(d, N, shape, axes, cov, z1, z2, nullptr, out, stream);
}

// Fourier method: the randomisation sum with per-mode spectrum factors.
int gsx_summate_fourier_f64(int64_t d, int64_t N, int64_t P, const double* spectrum_factor,
                            const double* modes, const double* z1, const double* z2, const double* pos,
                            double* out, void* stream) {
    if (N > 0 && !spectrum_factor) return GSX_E_INVALID;
    return gsx::forward_unstructured<false>(d, N, P, modes, z1, z2, spectrum_factor, pos, out, stream);
}

int gsx_summate_fourier_f32(int64_t d, int64_t N, int64_t P, const double* spectrum_factor,
                            const double* modes, const double* z1, const double* z2, const double* pos,
                            float* out, void* stream) {
    if (N > 0 && !spectrum_factor) return GSX_E_INVALID;
    return gsx::forward_unstructured<false>(d, N, P, modes, z1, z2, spectrum_factor, pos, out, stream);
}

int gsx_summate_fourier_grid_f64(int64_t d, int64_t N, const int64_t* shape, const double* axes,
                                 const double* spectrum_factor, const double* modes, const double* z1,
                                 const double* z2, double* out, void* stream) {
    if (N > 0 && !spectrum_factor) return GSX_E_INVALID;
    return gsx::forward_grid(d, N, shape, axes, modes, z1, z2, spectrum_factor, out, stream);
}

int gsx_summate_fourier_grid_f32(int64_t d, int64_t N, const int64_t* shape, const double* axes,
                                 const double* spectrum_factor, const double* modes, const double* z1,
                                 const double* z2, float* out, void* stream) {
    if (N > 0 && !spectrum_factor) return GSX_E_INVALID;
    return gsx::forward_grid(d, N, shape, axes, modes, z1, z2, spectrum_factor, out, stream);
}

}  // extern "C"
