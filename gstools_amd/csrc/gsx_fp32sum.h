// The FP32 mode sum shared by the single-launch kernel (gsx_summate.hip) and
// the persistent batch kernel (gsx_batch.hip): one thread, PPT points, every
// mode of one item, FP64 phase in revolutions, hardware cosine, blocked FP32
// partial sums added into FP64 accumulators.
//
// Two phase reductions (gsx_internal.h): magic addition for the modes whose
// |k/2pi|_1 times the largest |x|_inf of the point's segment stays within
// kPhaseMagicMax revolutions, fract-and-convert for the rest.  Which modes take
// which, and the order they are summed in, depend on two units that do not
// depend on the kernel's block or tile shape, so both kernels produce the same
// bits for a point:
//  * the segment: the aligned run of kSeg points of the item that contains the
//    point (a block or tile covers kThreads * PPT points, a divisor of kSeg);
//  * the mode group: kGroup consecutive modes; within a group the magic-path
//    modes are summed first, then the others, each in mode order, FP32 partial
//    sums restarting at the start of each of the two runs and every kSub modes.
// A coordinate that is NaN or infinite makes its segment's bound infinite, so
// every mode takes the fract form there and NaN propagates as in the plain sum.
#pragma once
#include "gsx_internal.h"

namespace gsx {

constexpr int kSub = 64;      // FP32 terms per FP32 partial sum
constexpr int kSeg = 2048;    // points per segment
constexpr int kGroup = 128;   // modes per group
constexpr int kWaves = kThreads / 64;

// Modes [mb, me) of the LDS table into the FP64 accumulators, FP32 over kSub
// modes at a time.  MAGIC: the phase chain starts at the table's kPhaseMagic
// offset and is reduced by magic_rev_to_unit; else the offset is taken off
// again (exact: both lie within a factor 2) and the phase reduced by
// v_fract_f64 and converted, valid at any |t|.
template <int D, int PPT, int NOUT, bool MAGIC>
__device__ __forceinline__ void sum_modes_f32(const double* tabd, const float* tabf, int mb, int me,
                                              const double (&x)[PPT][D], double (&acc)[PPT][NOUT]) {
    constexpr int kWd = D + 1;
    for (int m0 = mb; m0 < me; m0 += kSub) {
        const int m1 = (m0 + kSub < me) ? m0 + kSub : me;
        float part[PPT][NOUT];
#pragma unroll
        for (int p = 0; p < PPT; ++p)
#pragma unroll
            for (int c = 0; c < NOUT; ++c) part[p][c] = 0.0f;
#pragma unroll 2
        for (int m = m0; m < m1; ++m) {
            const double* t = tabd + m * kWd;
            double kr[D];
#pragma unroll
            for (int d = 0; d < D; ++d) kr[d] = t[d];
            const double c0 = MAGIC ? t[D] : t[D] - (kPhaseMagic + kPhaseRound);
            float w[NOUT];
#pragma unroll
            for (int c = 0; c < NOUT; ++c) w[c] = tabf[m * NOUT + c];
#pragma unroll
            for (int p = 0; p < PPT; ++p) {
                double rev = c0;
#pragma unroll
                for (int d = 0; d < D; ++d) rev = fma_(kr[d], x[p][d], rev);
                const float f = MAGIC ? magic_rev_to_unit(rev) : (float)__builtin_amdgcn_fract(rev);
                const float cs = __builtin_amdgcn_cosf(f);
#pragma unroll
                for (int c = 0; c < NOUT; ++c) part[p][c] = __builtin_fmaf(w[c], cs, part[p][c]);
            }
        }
#pragma unroll
        for (int p = 0; p < PPT; ++p)
#pragma unroll
            for (int c = 0; c < NOUT; ++c) acc[p][c] += (double)part[p][c];
    }
}

// Largest |x|_inf over the segment holding points [p0, p0 + kThreads * PPT)
// (p0 a multiple of kThreads * PPT), infinite if any coordinate is not finite.
// `own` is this thread's value over its own PPT points; when the block covers
// the whole segment it is all that is needed.  Uniform over the block; ends
// with a barrier after which s_wmax may be reused.
template <int D, int PPT, class Pts>
__device__ __forceinline__ double segment_xmax(const Pts& pts, int64_t P, int64_t p0, double own,
                                               double* s_wmax) {
    double xm = own;
    if constexpr (PPT * kThreads != kSeg) {
        static_assert(kSeg % (PPT * kThreads) == 0, "a block covers a divisor of the segment");
        const int64_t s0 = p0 - p0 % kSeg;
        xm = 0.0;
#pragma unroll
        for (int q = 0; q < kSeg / kThreads; ++q) {
            const int64_t i = s0 + threadIdx.x + (int64_t)q * kThreads;
            if (i < P) {
                double y[D];
                pts.get(i, y);
#pragma unroll
                for (int d = 0; d < D; ++d) xm = coord_bound(xm, y[d]);
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) xm = fmax(xm, __shfl_xor(xm, o));
    __syncthreads();  // earlier readers of s_wmax are done
    if ((threadIdx.x & 63) == 0) s_wmax[threadIdx.x >> 6] = xm;
    __syncthreads();
    double r = s_wmax[0];
#pragma unroll
    for (int v = 1; v < kWaves; ++v) r = fmax(r, s_wmax[v]);
    return r;
}

// All N modes of one item into acc, for this thread's PPT points x, with the
// segment bound xseg.  Tables: tabd CHUNK * (D + 1) doubles, tabf CHUNK * NOUT
// floats, s_nmagic kWaves ints, all in LDS.  CHUNK is a multiple of kGroup and
// at most kThreads.  Every thread of the block calls it with the same N.
template <int D, int PPT, bool INCOMPR, int CHUNK>
__device__ __forceinline__ void fp32_item_sum(int64_t N, const double* __restrict__ cov,
                                              const double* __restrict__ z1, const double* __restrict__ z2,
                                              const double* __restrict__ fac, double xseg,
                                              const double (&x)[PPT][D],
                                              double (&acc)[PPT][INCOMPR ? D : 1], double* tabd,
                                              float* tabf, int* s_nmagic) {
    static_assert(CHUNK % kGroup == 0 && CHUNK <= kThreads, "chunk of whole groups, one mode per thread");
    constexpr int NOUT = INCOMPR ? D : 1;
    constexpr int kWd = D + 1;  // k/2pi [D], kPhaseMagic + kPhaseRound - phi/2pi
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6, grp = tid / kGroup;
    for (int64_t j0 = 0; j0 < N; j0 += CHUNK) {
        const int cnt = (int)(N - j0 < CHUNK ? N - j0 : CHUNK);
        const bool live = tid < cnt;
        double e[kWd];
        float w[NOUT];
        bool magic = false;
        if (live) {
            const int64_t j = j0 + tid;
            double k[D];
            double k2 = 0.0, k1 = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                k[d] = cov[d * N + j];
                e[d] = k[d] * kInv2Pi;
                k1 += fabs(e[d]);
                k2 += k[d] * k[d];
            }
            const double u1 = z1[j], u2 = z2[j];
            e[D] = (kPhaseMagic + kPhaseRound) - atan2(u2, u1) * kInv2Pi;
            magic = k1 * xseg <= kPhaseMagicMax;  // false for NaN or infinite bounds
            const double a = fac ? fac[j] * hypot(u1, u2) : hypot(u1, u2);
            if constexpr (INCOMPR) {
#pragma unroll
                for (int c = 0; c < D; ++c) w[c] = (float)(((c == 0 ? 1.0 : 0.0) - k[c] * k[0] / k2) * a);
            } else {
                w[0] = (float)a;
            }
        }
        const unsigned long long bal = __ballot(magic);
        __syncthreads();  // the previous chunk's table is no longer read
        if (lane == 0) s_nmagic[wave] = __popcll(bal);
        __syncthreads();
        if (live) {
            // a group is kGroup / 64 waves; magic modes of the group before this one
            constexpr int kGw = kGroup / 64;
            int before = 0, ngm = 0;
#pragma unroll
            for (int v = 0; v < kGw; ++v) {
                const int c = s_nmagic[grp * kGw + v];
                ngm += c;
                before += (grp * kGw + v) < wave ? c : 0;
            }
            const int mb = before + __popcll(bal & ((1ull << lane) - 1ull));
            const int g0 = grp * kGroup;  // every thread of the group below this one is live
            const int dst = magic ? g0 + mb : g0 + ngm + (tid - g0) - mb;
#pragma unroll
            for (int d = 0; d < kWd; ++d) tabd[dst * kWd + d] = e[d];
#pragma unroll
            for (int c = 0; c < NOUT; ++c) tabf[dst * NOUT + c] = w[c];
        }
        __syncthreads();
        for (int g0 = 0; g0 < cnt; g0 += kGroup) {
            const int ge = (g0 + kGroup < cnt) ? g0 + kGroup : cnt;
            int ngm = 0;
#pragma unroll
            for (int v = 0; v < kGroup / 64; ++v) ngm += s_nmagic[g0 / 64 + v];
            sum_modes_f32<D, PPT, NOUT, true>(tabd, tabf, g0, g0 + ngm, x, acc);
            sum_modes_f32<D, PPT, NOUT, false>(tabd, tabf, g0 + ngm, ge, x, acc);
        }
    }
}

}  // namespace gsx
