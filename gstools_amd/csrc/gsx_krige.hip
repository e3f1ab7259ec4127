// Kriging sums on gfx950: field and variance term of every target point.
//
//   w_k      = M v_k                      (M = krig_mat (K, K), v_k = column k of krig_vecs (K, P))
//   field_k  = cond . w_k
//   error_k  = v_k . w_k                  (variance form only)
//
// Reference contract: the external calc_field_krige(_and_variance)
// (src/gstools/krige/base.py:42-61, called at :307-317); the caller turns
// error into the kriging variance max(sill - error, 0).
//
// Variance form: W = M V is an FP64 GEMM (2 K^2 P flops) with an epilogue
// that never stores W: every block owns a 64-row x 128-point tile of W,
// accumulates it on the matrix cores (v_mfma_f64_16x16x4, 4 waves in a 2 x 2
// arrangement, each 32 rows x 64 points), and reduces V o W and cond o W over
// its 64 rows into per-row-tile partial sums.  A second kernel adds the
// partials of the row tiles in a fixed order (bitwise reproducible).  Operand
// tiles (64 x 16 of M, 16 x 128 of V) are staged through LDS with the next
// stage prefetched into registers.
//
// Field-only form: field = (M^T cond) . v_k, a weight vector (K) followed by
// an HBM-bound pass over V (no GEMM).
//
// MFMA operand layouts (v_mfma_f64_16x16x4, 64 lanes, g = lane / 16):
//   A: lane holds row (lane % 16), K element g
//   B: lane holds column (lane % 16), K element g
//   C: lane holds column (lane % 16), rows 4g .. 4g+3
#include "gsx_internal.h"

namespace gsx {

typedef __attribute__((ext_vector_type(4))) double f64x4;

constexpr int kKrRows = 64;   // W rows per block
constexpr int kKrPts = 128;   // points per block
constexpr int kKrK = 16;      // reduction depth per LDS stage
constexpr int kKrApad = kKrK + 1;
constexpr int kKrBpad = kKrPts + 2;

__global__ __launch_bounds__(256) void k_krige_gemm(int64_t K, int64_t P, const double* __restrict__ M,
                                                    const double* __restrict__ V,
                                                    const double* __restrict__ cond,
                                                    double* __restrict__ part_f, double* __restrict__ part_e) {
    __shared__ double As[kKrRows * kKrApad];
    __shared__ double Bs[kKrK * kKrBpad];
    __shared__ double red[2][2][kKrPts];  // [row half][field / error][point]

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int g = lane >> 4, l16 = lane & 15;
    const int64_t i0 = (int64_t)blockIdx.y * kKrRows;
    const int64_t p0 = (int64_t)blockIdx.x * kKrPts;

    // staging assignment: A 64 x 16 (4 per thread), B 16 x 128 (8 per thread)
    const int ar = t >> 2, ac = (t & 3) * 4;
    const int br = t >> 4, bc = (t & 15) * 8;
    double pa[4], pb[8];
    auto load = [&](int64_t j0) {
        const int64_t row = i0 + ar;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t col = j0 + ac + u;
            pa[u] = (row < K && col < K) ? M[row * K + col] : 0.0;
        }
        const int64_t j = j0 + br;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int64_t p = p0 + bc + u;
            pb[u] = (j < K && p < P) ? V[j * P + p] : 0.0;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int u = 0; u < 4; ++u) As[ar * kKrApad + ac + u] = pa[u];
#pragma unroll
        for (int u = 0; u < 8; ++u) Bs[br * kKrBpad + bc + u] = pb[u];
    };

    f64x4 acc[2][4];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = f64x4{0.0, 0.0, 0.0, 0.0};

    load(0);
    for (int64_t j0 = 0; j0 < K; j0 += kKrK) {
        __syncthreads();
        store();
        __syncthreads();
        if (j0 + kKrK < K) load(j0 + kKrK);
#pragma unroll
        for (int kk = 0; kk < kKrK; kk += 4) {
            double a[2], b[4];
#pragma unroll
            for (int rt = 0; rt < 2; ++rt) a[rt] = As[(wr * 32 + rt * 16 + l16) * kKrApad + kk + g];
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) b[ct] = Bs[(kk + g) * kKrBpad + wc * 64 + ct * 16 + l16];
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int ct = 0; ct < 4; ++ct)
                    acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt], b[ct], acc[rt][ct], 0, 0, 0);
        }
    }

    // epilogue: per column, sum over this wave's 32 rows of cond o W and V o W
    double sf[4], se[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const int64_t p = p0 + wc * 64 + ct * 16 + l16;
        double f = 0.0, e = 0.0;
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t i = i0 + wr * 32 + rt * 16 + 4 * g + r;
                if (i < K && p < P) {
                    const double w = acc[rt][ct][r];
                    f += cond[i] * w;
                    e += V[i * P + p] * w;
                }
            }
        // the four lane groups holding this column, in a fixed order
        f += __shfl_xor(f, 16);
        e += __shfl_xor(e, 16);
        f += __shfl_xor(f, 32);
        e += __shfl_xor(e, 32);
        sf[ct] = f;
        se[ct] = e;
    }
    if (g == 0) {
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            red[wr][0][wc * 64 + ct * 16 + l16] = sf[ct];
            red[wr][1][wc * 64 + ct * 16 + l16] = se[ct];
        }
    }
    __syncthreads();
    if (t < kKrPts) {
        const int64_t p = p0 + t;
        if (p < P) {
            part_f[(int64_t)blockIdx.y * P + p] = red[0][0][t] + red[1][0][t];
            part_e[(int64_t)blockIdx.y * P + p] = red[0][1][t] + red[1][1][t];
        }
    }
}

__global__ void k_krige_reduce(int64_t tiles, int64_t P, const double* __restrict__ part_f,
                               const double* __restrict__ part_e, double* __restrict__ field,
                               double* __restrict__ error) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double f = 0.0, e = 0.0;
    for (int64_t it = 0; it < tiles; ++it) {
        f += part_f[it * P + p];
        e += part_e[it * P + p];
    }
    field[p] = f;
    error[p] = e;
}

// a_j = sum_i cond_i M[i, j]
__global__ void k_krige_weights(int64_t K, const double* __restrict__ M, const double* __restrict__ cond,
                                double* __restrict__ a) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= K) return;
    double s = 0.0;
    for (int64_t i = 0; i < K; ++i) s += cond[i] * M[i * K + j];
    a[j] = s;
}

// field_p = sum_j a_j V[j, p]; two points per thread
__global__ __launch_bounds__(256) void k_krige_field(int64_t K, int64_t P, const double* __restrict__ a,
                                                     const double* __restrict__ V, double* __restrict__ field) {
    const int64_t p = (int64_t)blockIdx.x * 512 + threadIdx.x;
    const int64_t q = p + 256;
    double f0 = 0.0, f1 = 0.0;
    for (int64_t j = 0; j < K; ++j) {
        const double aj = a[j];
        if (p < P) f0 += aj * V[j * P + p];
        if (q < P) f1 += aj * V[j * P + q];
    }
    if (p < P) field[p] = f0;
    if (q < P) field[q] = f1;
}

__global__ void k_fill_zero_f64(double* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

inline int64_t krige_row_tiles(int64_t K) { return ceil_div(K, (int64_t)kKrRows); }

inline int64_t krige_ws_bytes(int64_t K, int64_t P, bool variance) {
    if (variance) return 2 * krige_row_tiles(K) * P * (int64_t)sizeof(double);
    return K * (int64_t)sizeof(double);
}

}  // namespace gsx

extern "C" {

int64_t gsx_krige_workspace(int64_t K, int64_t P, int variance) {
    if (K < 0 || P < 0) return GSX_E_INVALID;
    return gsx::krige_ws_bytes(K, P, variance != 0);
}

int gsx_krige_f64(int64_t K, int64_t P, const double* krig_mat, const double* krig_vecs, const double* cond,
                  double* field, double* error, void* workspace, int64_t workspace_bytes, void* stream) {
    if (K < 0 || P < 0) return GSX_E_INVALID;
    if (P == 0) return GSX_OK;
    if (!field) return GSX_E_INVALID;
    hipStream_t s = gsx::as_stream(stream);
    if (K == 0) {
        const unsigned blocks = (unsigned)gsx::ceil_div(P, (int64_t)256);
        hipLaunchKernelGGL(gsx::k_fill_zero_f64, dim3(blocks), dim3(256), 0, s, field, P);
        if (error) hipLaunchKernelGGL(gsx::k_fill_zero_f64, dim3(blocks), dim3(256), 0, s, error, P);
        return gsx::launch_status();
    }
    if (!krig_mat || !krig_vecs || !cond) return GSX_E_INVALID;
    const bool variance = error != nullptr;
    if (!workspace || workspace_bytes < gsx::krige_ws_bytes(K, P, variance)) return GSX_E_WORKSPACE;
    double* ws = static_cast<double*>(workspace);
    if (!variance) {
        hipLaunchKernelGGL(gsx::k_krige_weights, dim3((unsigned)gsx::ceil_div(K, (int64_t)256)), dim3(256), 0, s,
                           K, krig_mat, cond, ws);
        int st = gsx::launch_status();
        if (st != GSX_OK) return st;
        hipLaunchKernelGGL(gsx::k_krige_field, dim3((unsigned)gsx::ceil_div(P, (int64_t)512)), dim3(256), 0, s, K,
                           P, ws, krig_vecs, field);
        return gsx::launch_status();
    }
    const int64_t tiles = gsx::krige_row_tiles(K);
    const int64_t col_blocks = gsx::ceil_div(P, (int64_t)gsx::kKrPts);
    if (tiles > 65535 || col_blocks > 0x7fffffffLL) return GSX_E_INVALID;
    double* part_f = ws;
    double* part_e = ws + tiles * P;
    hipLaunchKernelGGL(gsx::k_krige_gemm, dim3((unsigned)col_blocks, (unsigned)tiles), dim3(256), 0, s, K, P,
                       krig_mat, krig_vecs, cond, part_f, part_e);
    int st = gsx::launch_status();
    if (st != GSX_OK) return st;
    hipLaunchKernelGGL(gsx::k_krige_reduce, dim3((unsigned)gsx::ceil_div(P, (int64_t)256)), dim3(256), 0, s, tiles,
                       P, part_f, part_e, field, error);
    return gsx::launch_status();
}

}  // extern "C"
