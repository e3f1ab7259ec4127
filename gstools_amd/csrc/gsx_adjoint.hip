// Adjoint of the randomisation-method sum (gradients for calibration).
//
// Forward (src/gstools/field/generator.py:193-199):
//   u_i = sum_j z1_j cos(t_ij) + z2_j sin(t_ij),   t_ij = k_j . x_i
// For an upstream gradient g_i = dL/du_i:
//   dL/dz1_j   = sum_i g_i cos t_ij
//   dL/dz2_j   = sum_i g_i sin t_ij
//   dL/dk_jd   = sum_i g_i x_id (z2_j cos t_ij - z1_j sin t_ij)
//              = z2_j Cx_jd - z1_j Sx_jd,  Cx_jd = sum_i (g_i x_id) cos t_ij, ...
//
// Schedule: a reduction over points per mode.  Stage 1 runs a grid of
// (mode tile of kThreads modes) x (point chunk); each thread owns one mode and
// walks the chunk's points, which are staged through LDS with g_i x_id
// precomputed (every lane reads the same point: LDS broadcast).  It writes
// the 2 + 2D partial sums of its (chunk, mode) to the workspace.  Stage 2
// sums the chunks of each mode in a fixed order, so results are bitwise
// reproducible run to run (no atomics).
//
// FP64: exact reference phase and sincos_fast (two-term reduction) with a
// library fallback for waves whose phases may exceed 2^46.  FP32: phase in
// revolutions in FP64, reduced by magic addition (gsx_internal.h) where the
// wave's |k/2pi|_1 times the stage's largest |x|_inf stays within 2^18
// revolutions, else by v_fract_f64, hardware v_sin_f32 / v_cos_f32, FP32
// partial sums per LDS chunk folded into FP64 accumulators.  The FP32 path
// sums over the shifted phase psi_ij = t_ij - phi_j, z1 + i z2 = a e^{i phi}
// (the forward kernel's amplitude-phase form), which needs 2 + D sums per
// mode instead of 2 + 2D:
//   C_j = sum_i g_i cos psi_ij,  S_j = sum_i g_i sin psi_ij,
//   Sx_jd = sum_i g_i x_id sin psi_ij,
//   dL/dz1_j = cos phi C - sin phi S,   dL/dz2_j = sin phi C + cos phi S,
//   dL/dk_jd = z2 Cx - z1 Sx = -a_j Sx_jd   (z2 cos t - z1 sin t = -a sin psi).
#include <mutex>

#include "gsx_internal.h"

namespace gsx {

constexpr int kAdjPts = 256;         // points per LDS stage
constexpr int64_t kAdjTargetBlocks = 2048;
constexpr int kAdjModesF32 = 2;        // modes per thread of the FP32 stage 1

struct AdjPlan {
    int64_t mode_tiles, chunks, chunk_pts;
};

// Point chunks of the stage-1 grid.  max_chunks_cap (> 0) bounds the number
// of chunks, and with it the partial sums (chunks x (2 + 2D) x N doubles).
inline AdjPlan adj_plan(int64_t N, int64_t P, int64_t max_chunks_cap = 0) {
    AdjPlan pl;
    pl.mode_tiles = ceil_div(N, kThreads);
    const int64_t max_chunks = ceil_div(P, kAdjPts);
    int64_t chunks = ceil_div(kAdjTargetBlocks, pl.mode_tiles);
    if (chunks > max_chunks) chunks = max_chunks;
    if (max_chunks_cap > 0 && chunks > max_chunks_cap) chunks = max_chunks_cap;
    if (chunks < 1) chunks = 1;
    pl.chunk_pts = ceil_div(ceil_div(P, chunks), kAdjPts) * kAdjPts;
    pl.chunks = ceil_div(P, pl.chunk_pts);
    return pl;
}

// sin / cos at any |x|: the fast reduction where it is exact, else the library.
__device__ __noinline__ void sincos_any(double x, double* sn, double* cs) {
    if (fabs(x) <= kFastPhaseMax) {
        sincos_fast(x, sn, cs);
    } else {
        sincos(x, sn, cs);
    }
}

// partial sums per (chunk, mode): FP64 2 + 2D (unshifted phase), FP32 2 + D (shifted)
template <int D, bool F64>
constexpr int adj_sums() {
    return F64 ? 2 + 2 * D : 2 + D;
}

// MPT modes per thread (threads t, t + kThreads, ... of the block's
// kThreads * MPT modes): every staged point's LDS reads and the loop's
// bookkeeping serve MPT modes; each mode sums its points in the same order
// whatever MPT is.  The FP64 path takes MPT = 1.
// One LDS stage of cnt points into the FP32 partial sums of this thread's MPT
// modes.  MAGIC: the phase chain starts at kPhaseMagic + kPhaseRound -
// phi/2pi and is reduced by magic_rev_to_unit (the caller has checked the
// stage's bound); else by v_fract_f64 and a conversion.
template <int D, int MPT, int Q, typename GT, bool MAGIC>
__device__ __forceinline__ void adj_stage_f32(int cnt, const double* s_x, const GT* s_g,
                                              const double (&k)[MPT][D], const double (&mphi)[MPT],
                                              float (&pa)[MPT][Q]) {
    double c0[MPT];
#pragma unroll
    for (int u = 0; u < MPT; ++u) c0[u] = MAGIC ? (kPhaseMagic + kPhaseRound) + mphi[u] : mphi[u];
    // four points per iteration: their LDS reads are issued together and one
    // wait serves them (the compiler unrolls only two with both forms inlined)
#pragma unroll 4
    for (int m = 0; m < cnt; ++m) {
        double x[D];
#pragma unroll
        for (int d = 0; d < D; ++d) x[d] = s_x[m * D + d];
        const GT* gg = s_g + m * (D + 1);
        const float g0 = gg[0];
        float gx[D];
#pragma unroll
        for (int d = 0; d < D; ++d) gx[d] = gg[1 + d];
#pragma unroll
        for (int u = 0; u < MPT; ++u) {
            double rev = c0[u];
#pragma unroll
            for (int d = 0; d < D; ++d) rev = fma_(k[u][d], x[d], rev);
            const float f = MAGIC ? magic_rev_to_unit(rev) : (float)__builtin_amdgcn_fract(rev);
            const float cs = __builtin_amdgcn_cosf(f);
            const float sn = __builtin_amdgcn_sinf(f);
            pa[u][0] = __builtin_fmaf(g0, cs, pa[u][0]);
            pa[u][1] = __builtin_fmaf(g0, sn, pa[u][1]);
#pragma unroll
            for (int d = 0; d < D; ++d) pa[u][2 + d] = __builtin_fmaf(gx[d], sn, pa[u][2 + d]);
        }
    }
}

template <int D, typename GT, int MPT>
__global__ __launch_bounds__(kThreads) void k_adjoint_stage1(int64_t N, int64_t P, int64_t chunk_pts,
                                                             const double* __restrict__ cov,
                                                             const double* __restrict__ z1,
                                                             const double* __restrict__ z2,
                                                             const double* __restrict__ pos,
                                                             const GT* __restrict__ g,
                                                             double* __restrict__ part) {
    constexpr bool F64 = sizeof(GT) == 8;
    static_assert(!F64 || MPT == 1, "the FP64 stage takes one mode per thread");
    constexpr int Q = adj_sums<D, F64>();
    __shared__ double s_x[kAdjPts * D];
    __shared__ GT s_g[kAdjPts * (D + 1)];  // g, g x_0 .. g x_{D-1}
    __shared__ unsigned long long s_xmax;

    int64_t j[MPT];
    bool active[MPT];
    double k[MPT][D];
    double mphi[MPT];
    double k1 = 0.0;
#pragma unroll
    for (int u = 0; u < MPT; ++u) {
        j[u] = (int64_t)blockIdx.x * kThreads * MPT + u * kThreads + threadIdx.x;
        active[u] = j[u] < N;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            k[u][d] = active[u] ? cov[d * N + j[u]] : 0.0;
            if (!F64) k[u][d] *= kInv2Pi;
            k1 += fabs(k[u][d]);
        }
        // FP32: -phi / 2pi, the phase shift in revolutions
        mphi[u] = (!F64 && active[u]) ? -atan2(z2[j[u]], z1[j[u]]) * kInv2Pi : 0.0;
    }
    double acc[MPT][Q];
#pragma unroll
    for (int u = 0; u < MPT; ++u)
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[u][q] = 0.0;

    const int64_t p0 = (int64_t)blockIdx.y * chunk_pts;
    const int64_t p1 = (p0 + chunk_pts < P) ? p0 + chunk_pts : P;
    for (int64_t s0 = p0; s0 < p1; s0 += kAdjPts) {
        const int cnt = (int)((p1 - s0) < kAdjPts ? (p1 - s0) : kAdjPts);
        __syncthreads();
        if (threadIdx.x == 0) s_xmax = 0ull;
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            const int64_t i = s0 + threadIdx.x;
            const GT gi = g[i];
            double xm = 0.0;
            s_g[threadIdx.x * (D + 1)] = gi;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const double xv = pos[d * P + i];
                s_x[threadIdx.x * D + d] = xv;
                s_g[threadIdx.x * (D + 1) + 1 + d] = (GT)(gi * xv);
                xm = coord_bound(xm, xv);
            }
            atomicMax(&s_xmax, (unsigned long long)__double_as_longlong(xm));
        }
        __syncthreads();
        if constexpr (F64) {
            const double xmax = __longlong_as_double((long long)s_xmax);
            if (__all(k1 * xmax <= 0x1.0p+45)) {
                for (int m = 0; m < cnt; ++m) {
                    double x[D];
#pragma unroll
                    for (int d = 0; d < D; ++d) x[d] = s_x[m * D + d];
                    double sn, cs;
                    sincos_fast(phase_exact<D>(k[0], x), &sn, &cs);
                    const GT* gg = s_g + m * (D + 1);
                    acc[0][0] = fma_(gg[0], cs, acc[0][0]);
                    acc[0][1] = fma_(gg[0], sn, acc[0][1]);
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        acc[0][2 + d] = fma_(gg[1 + d], cs, acc[0][2 + d]);
                        acc[0][2 + D + d] = fma_(gg[1 + d], sn, acc[0][2 + D + d]);
                    }
                }
            } else {
                for (int m = 0; m < cnt; ++m) {
                    double x[D];
                    for (int d = 0; d < D; ++d) x[d] = s_x[m * D + d];
                    const double th = phase_exact<D>(k[0], x);
                    double sn, cs;
                    sincos_any(th, &sn, &cs);
                    const GT* gg = s_g + m * (D + 1);
                    acc[0][0] = fma_(gg[0], cs, acc[0][0]);
                    acc[0][1] = fma_(gg[0], sn, acc[0][1]);
                    for (int d = 0; d < D; ++d) {
                        acc[0][2 + d] = fma_(gg[1 + d], cs, acc[0][2 + d]);
                        acc[0][2 + D + d] = fma_(gg[1 + d], sn, acc[0][2 + D + d]);
                    }
                }
            }
        } else {
            float pa[MPT][Q];
#pragma unroll
            for (int u = 0; u < MPT; ++u)
#pragma unroll
                for (int q = 0; q < Q; ++q) pa[u][q] = 0.0f;
            // magic-addition reduction when every mode of the wave keeps
            // |k/2pi|_1 |x|_inf within 2^18 revolutions over this stage
            const double xmax = __longlong_as_double((long long)s_xmax);
            if (__all(k1 * xmax <= kPhaseMagicMax)) {
                adj_stage_f32<D, MPT, Q, GT, true>(cnt, s_x, s_g, k, mphi, pa);
            } else {
                adj_stage_f32<D, MPT, Q, GT, false>(cnt, s_x, s_g, k, mphi, pa);
            }
#pragma unroll
            for (int u = 0; u < MPT; ++u)
#pragma unroll
                for (int q = 0; q < Q; ++q) acc[u][q] += (double)pa[u][q];
        }
    }
#pragma unroll
    for (int u = 0; u < MPT; ++u) {
        if (active[u]) {
            double* dst = part + (int64_t)blockIdx.y * Q * N;
#pragma unroll
            for (int q = 0; q < Q; ++q) dst[q * N + j[u]] = acc[u][q];
        }
    }
}

template <int D, bool F64>
__global__ __launch_bounds__(kThreads) void k_adjoint_stage2(int64_t N, int64_t chunks,
                                                             const double* __restrict__ z1,
                                                             const double* __restrict__ z2,
                                                             const double* __restrict__ part,
                                                             double* __restrict__ gz1,
                                                             double* __restrict__ gz2,
                                                             double* __restrict__ gk) {
    constexpr int Q = adj_sums<D, F64>();
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j >= N) return;
    double acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = 0.0;
    for (int64_t c = 0; c < chunks; ++c) {
        const double* src = part + c * Q * N;
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[q] += src[q * N + j];
    }
    const double a1 = z1[j], a2 = z2[j];
    if constexpr (F64) {
        gz1[j] = acc[0];
        gz2[j] = acc[1];
#pragma unroll
        for (int d = 0; d < D; ++d) gk[d * N + j] = a2 * acc[2 + d] - a1 * acc[2 + D + d];
    } else {
        // rotate the shifted sums back by phi (a = 0: phi = 0, nothing to rotate)
        const double a = hypot(a1, a2);
        const double c = a > 0.0 ? a1 / a : 1.0, sn = a > 0.0 ? a2 / a : 0.0;
        gz1[j] = c * acc[0] - sn * acc[1];
        gz2[j] = sn * acc[0] + c * acc[1];
#pragma unroll
        for (int d = 0; d < D; ++d) gk[d * N + j] = -a * acc[2 + d];
    }
}

template <int D, typename GT>
int adjoint_launch(const AdjPlan& pl, int64_t N, int64_t P, const double* cov, const double* z1,
                   const double* z2, const double* pos, const GT* g, double* gz1, double* gz2,
                   double* gk, void* ws, hipStream_t s) {
    double* part = static_cast<double*>(ws);
    constexpr int mpt = sizeof(GT) == 8 ? 1 : kAdjModesF32;
    hipLaunchKernelGGL((k_adjoint_stage1<D, GT, mpt>), dim3((unsigned)ceil_div(N, (int64_t)kThreads * mpt),
                                                          (unsigned)pl.chunks),
                       dim3(kThreads), 0, s, N, P, pl.chunk_pts, cov, z1, z2, pos, g, part);
    int st = launch_status();
    if (st != GSX_OK) return st;
    hipLaunchKernelGGL((k_adjoint_stage2<D, sizeof(GT) == 8>), dim3((unsigned)pl.mode_tiles), dim3(kThreads), 0, s,
                       N, pl.chunks, z1, z2, part, gz1, gz2, gk);
    return launch_status();
}

// Fills the gradients with zeros when there are no points.
__global__ void k_zero(double* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Gradients when there are points; ws holds pl.chunks x (2 + 2d) x N doubles.
template <typename GT>
int adjoint_run(const AdjPlan& pl, int64_t d, int64_t N, int64_t P, const double* cov,
                const double* z1, const double* z2, const double* pos, const GT* g, double* gz1,
                double* gz2, double* gk, void* ws, hipStream_t s) {
    switch (d) {
        case 1: return adjoint_launch<1>(pl, N, P, cov, z1, z2, pos, g, gz1, gz2, gk, ws, s);
        case 2: return adjoint_launch<2>(pl, N, P, cov, z1, z2, pos, g, gz1, gz2, gk, ws, s);
        default: return adjoint_launch<3>(pl, N, P, cov, z1, z2, pos, g, gz1, gz2, gk, ws, s);
    }
}

// Argument checks shared by both entry points; *done is set when nothing
// is left to do (no modes, or no points: the gradients are zero-filled).
template <typename GT>
int adjoint_prologue(int64_t d, int64_t N, int64_t P, const double* cov, const double* z1,
                     const double* z2, const double* pos, const GT* g, double* gz1, double* gz2,
                     double* gk, hipStream_t s, bool* done) {
    *done = true;
    if (d < 1 || d > 3 || N < 0 || P < 0) return GSX_E_INVALID;
    if (N == 0) return GSX_OK;
    if (!cov || !z1 || !z2 || !gz1 || !gz2 || !gk) return GSX_E_INVALID;
    if (P == 0) {
        const unsigned b = (unsigned)ceil_div(N, kThreads);
        hipLaunchKernelGGL(k_zero, dim3(b), dim3(kThreads), 0, s, gz1, N);
        hipLaunchKernelGGL(k_zero, dim3(b), dim3(kThreads), 0, s, gz2, N);
        hipLaunchKernelGGL(k_zero, dim3((unsigned)ceil_div(d * N, kThreads)), dim3(kThreads), 0, s, gk, d * N);
        return launch_status();
    }
    if (!pos || !g) return GSX_E_INVALID;
    *done = false;
    return GSX_OK;
}

template <typename GT>
int adjoint_entry(int64_t d, int64_t N, int64_t P, const double* cov, const double* z1,
                  const double* z2, const double* pos, const GT* g, double* gz1, double* gz2,
                  double* gk, void* ws, int64_t ws_bytes, void* stream) {
    hipStream_t s = as_stream(stream);
    bool done;
    const int st = adjoint_prologue<GT>(d, N, P, cov, z1, z2, pos, g, gz1, gz2, gk, s, &done);
    if (done) return st;
    const int64_t need = gsx_summate_adjoint_workspace(d, N, P);
    if (!ws || ws_bytes < need) return GSX_E_WORKSPACE;
    return adjoint_run<GT>(adj_plan(N, P), d, N, P, cov, z1, z2, pos, g, gz1, gz2, gk, ws, s);
}

// ---------------------------------------------------------------------------
// Allocation-free variant: the partial sums live in a fixed device array.
//
// `scratch` is one array per GPU (each device that loads this code object
// holds its own copy until the process ends).  Its host-side name is only a
// stand-in for the symbol: the device address comes from
// hipGetSymbolAddress, once per device with that device current, and only
// that address is handed to kernels.  Stage 1 writes every partial sum that
// stage 2 reads, so nothing needs clearing between calls; two calls that
// overlap on one GPU share the array and race.
constexpr int64_t kScratchBytes = GSX_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (resolved on first use).
static int scratch_address(void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return GSX_E_NO_DEVICE;
    if (dev < 0 || dev >= kMaxDevices) return GSX_E_NO_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return GSX_E_LAUNCH;
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return GSX_OK;
}

template <typename GT>
int adjoint_noalloc_entry(int64_t d, int64_t N, int64_t P, const double* cov, const double* z1,
                          const double* z2, const double* pos, const GT* g, double* gz1,
                          double* gz2, double* gk, void* stream) {
    if (d >= 1 && d <= 3 && N > 0 && P > 0) {
        // one chunk of partial sums per mode must fit the array
        const int64_t per_chunk = (2 + 2 * d) * N * (int64_t)sizeof(double);
        if (per_chunk > kScratchBytes) return GSX_E_SCRATCH;
    }
    hipStream_t s = as_stream(stream);
    bool done;
    const int st = adjoint_prologue<GT>(d, N, P, cov, z1, z2, pos, g, gz1, gz2, gk, s, &done);
    if (done) return st;
    void* ws = nullptr;
    const int sa = scratch_address(&ws);
    if (sa != GSX_OK) return sa;
    // the workspace entry point's plan whenever its partial sums fit, else
    // as many point chunks as fit (the sums then differ in rounding only)
    const int64_t per_chunk = (2 + 2 * d) * N * (int64_t)sizeof(double);
    const AdjPlan pl = adj_plan(N, P, kScratchBytes / per_chunk);
    return adjoint_run<GT>(pl, d, N, P, cov, z1, z2, pos, g, gz1, gz2, gk, ws, s);
}

}  // namespace gsx

extern "C" {

int64_t gsx_summate_adjoint_workspace(int64_t d, int64_t N, int64_t P) {
    if (d < 1 || d > 3 || N < 0 || P < 0) return GSX_E_INVALID;
    if (N == 0 || P == 0) return 0;
    const gsx::AdjPlan pl = gsx::adj_plan(N, P);
    return pl.chunks * (2 + 2 * d) * N * (int64_t)sizeof(double);
}

int gsx_summate_adjoint_f64(int64_t d, int64_t N, int64_t P, const double* cov, const double* z1,
                            const double* z2, const double* pos, const double* g, double* gz1,
                            double* gz2, double* gk, void* workspace, int64_t workspace_bytes,
                            void* stream) {
    return gsx::adjoint_entry<double>(d, N, P, cov, z1, z2, pos, g, gz1, gz2, gk, workspace,
                                      workspace_bytes, stream);
}

int gsx_summate_adjoint_f32(int64_t d, int64_t N, int64_t P, const double* cov, const double* z1,
                            const double* z2, const double* pos, const float* g, double* gz1,
                            double* gz2, double* gk, void* workspace, int64_t workspace_bytes,
                            void* stream) {
    return gsx::adjoint_entry<float>(d, N, P, cov, z1, z2, pos, g, gz1, gz2, gk, workspace,
                                     workspace_bytes, stream);
}

int gsx_summate_adjoint_f32_noalloc(int64_t d, int64_t N, int64_t P, const double* cov,
                                    const double* z1, const double* z2, const double* pos,
                                    const float* g, double* gz1, double* gz2, double* gk,
                                    void* stream) {
    return gsx::adjoint_noalloc_entry<float>(d, N, P, cov, z1, z2, pos, g, gz1, gz2, gk, stream);
}

}  // extern "C"
