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
// Field (both forms): field = (M^T cond) . v_k, a weight vector (K) followed
// by an HBM-bound pass over V (no GEMM).
//
// Variance form: error_k = v_k^T M v_k = v_k^T S v_k with S = (M + M^T) / 2
// (the antisymmetric part of M adds nothing to a quadratic form), and with S
// symmetric only the blocks on and right of the diagonal are needed:
//
//   error_k = sum over row tiles r of  v_r^T (M2 V)_r,
//   M2 = S on the diagonal 128 x 128 blocks, 2 S = M + M^T right of them,
//
// so the GEMM W = M2 V runs, for row tile r, over the columns from 128 r on:
// about K^2 P flops instead of 2 K^2 P.  Every block owns a 128-row x
// 128-point tile of W, accumulates it on the matrix cores (v_mfma_f64_16x16x4,
// 4 waves in a 2 x 2 arrangement, each 64 rows x 64 points, 16 MFMA tiles at 2
// waves per SIMD) and reduces V o W over its 128 rows into a per-row-tile
// partial sum (W is never stored).  A second kernel adds the partials of the
// row tiles in a fixed order (bitwise reproducible).  Operand tiles (128 x 16
// of M2, 16 x 128 of V) are staged through LDS with the next stage prefetched
// into registers.
//
// Right-hand sides: k_krige_cov writes the covariance rows of V,
// C(|c_i - x_p|) for conditioning points c_i and targets x_p, for models
// with closed-form correlations (the reference evaluates
// model.cov_nugget / model.covariance on scipy's cdist, krige/base.py:357-388;
// the distance here is rounded the same way: differences, squares and sums
// in dimension order, no fused multiply-add).
//
// MFMA operand layouts (v_mfma_f64_16x16x4, 64 lanes, g = lane / 16):
//   A: lane holds row (lane % 16), K element g
//   B: lane holds column (lane % 16), K element g
//   C: lane holds column (lane % 16), rows g, g+4, g+8, g+12 (register r holds
//      row g + 4r; unlike the 16-bit 16x16 MFMAs, whose lanes hold rows 4g..4g+3)
#include "gsx_internal.h"

namespace gsx {

typedef __attribute__((ext_vector_type(4))) double f64x4;

constexpr int kKrRows = 128;  // W rows per block
constexpr int kKrRF = kKrRows / 32;  // 16-row MFMA tiles per wave (two row waves)
constexpr int kKrPts = 128;   // points per block
constexpr int kKrK = 16;      // reduction depth per LDS stage
constexpr int kKrApad = kKrK + 1;
// B rows padded by 16 doubles: the operand reads of K groups g and g + 1
// (rows 144 doubles apart) fall in disjoint LDS banks
constexpr int kKrBpad = kKrPts + 16;

// M2 of the variance form, padded to Rp = 128 * row tiles rows and Kp (K
// rounded up to a multiple of 16) columns, row stride Kp: row tile r keeps its
// columns from 128 r on, zero beyond K (the GEMM's A loads then need no
// bounds checks); the entries left of the diagonal block are never read and
// not written
inline int64_t krige_kp(int64_t K) { return ceil_div(K, (int64_t)kKrK) * kKrK; }

__global__ __launch_bounds__(256) void k_krige_sym(int64_t K, int64_t Rp, int64_t Kp, const double* __restrict__ M,
                                                   double* __restrict__ M2) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= Rp * Kp) return;
    const int64_t i = idx / Kp, l = idx - i * Kp;
    const int64_t b0 = (i / kKrRows) * kKrRows;
    if (l < b0) return;
    if (i >= K || l >= K) {
        M2[idx] = 0.0;
        return;
    }
    const double s = M[i * K + l] + M[l * K + i];
    M2[idx] = l < b0 + kKrRows ? 0.5 * s : s;
}

// W = M2 V over the columns from the row tile on (K steps in descending
// order); per-row-tile partial sums of V o W.  Row tile r runs about K - 128 r
// K steps, so each block takes the pair (u, R - 1 - u): equal work per block
// (about K + 128 steps), and the long tile first.  The staging loads are
// 16 bytes wide: A (the padded M2, row stride Kp) four per thread with no
// bounds checks; B as point pairs (kPairB: P even and V 16-byte aligned),
// else one point of eight rows per thread.
template <bool kPairB>
__global__ __launch_bounds__(256, 2) void k_krige_gemm(int64_t K, int64_t Kp, int64_t P, int64_t row_tiles,
                                                    int64_t col_tiles, const double* __restrict__ M,
                                                    const double* __restrict__ V,
                                                    double* __restrict__ part_e) {
    typedef __attribute__((ext_vector_type(2))) double f64x2;
    __shared__ double As[kKrRows * kKrApad];
    __shared__ double Bs[kKrK * kKrBpad];
    __shared__ double red[2][kKrPts];  // [row half][point]

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int g = lane >> 4, l16 = lane & 15;
    // XCD-aware order: dispatch puts block b on XCD b % 8; the blocks of one XCD
    // take all row-tile pairs of a column tile in turn, so that column's slice of
    // V (K x 128 doubles) is reused from that XCD's L2
    const int64_t R = row_tiles, U = (R + 1) / 2;
    int64_t u_idx, ct_idx;
    {
        // the grid has U * C8 blocks, C8 = col_tiles rounded up to a multiple of
        // 8: every XCD owns the column tiles ct == xcd (mod 8) with their U pairs
        // (a bijection onto pairs x [0, C8)); padding blocks leave
        const int64_t b = blockIdx.x, xcd = b & 7, slot = b >> 3;
        u_idx = slot % U;
        ct_idx = (slot / U) * 8 + xcd;
    }
    if (ct_idx >= col_tiles) return;  // whole block: before any barrier
    const int64_t p0 = ct_idx * kKrPts;
    const int ntiles = (u_idx == R - 1 - u_idx) ? 1 : 2;
#pragma unroll 1
    for (int ti = 0; ti < ntiles; ++ti) {
        const int64_t rt_idx = ti == 0 ? u_idx : R - 1 - u_idx;
        const int64_t i0 = rt_idx * kKrRows;

        // staging assignment: A kKrRows x kKrK, 8 consecutive doubles of one row per
        // thread (four 16-byte loads from the padded M2); B kKrK x 128, either point
        // pair (t % 64) * 2 of rows t / 64 + 4u (16-byte loads and LDS stores) or
        // point t % 128 of rows t / 128 + 2u (unit-stride, conflict-free stores)
        static_assert(kKrRows == 128 && kKrK == 16, "staging assignment");
        const int ar = t >> 1, ac = (t & 1) * 8;
        const f64x2* arow = reinterpret_cast<const f64x2*>(M + (i0 + ar) * Kp + ac);
        const int br = kPairB ? t >> 6 : t >> 7;
        const int bc = kPairB ? (t & 63) * 2 : t & 127;
        f64x2 pa[4];
        f64x2 pb2[4];    // point pairs: loads from clamped (valid) addresses, zeroed at the store (bit u of mb)
        double pb1[8];   // single points
        unsigned mb = 0;
        auto load = [&](int64_t j0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) pa[v] = arow[(j0 >> 1) + v];
            const int64_t p = p0 + bc;
            if constexpr (kPairB) {
                const int64_t pc = p < P ? p : P - 2;  // P even: a pair is whole or out of range
                mb = 0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int64_t j = j0 + br + 4 * u;
                    pb2[u] = *reinterpret_cast<const f64x2*>(V + (j < K ? j : K - 1) * P + pc);
                    mb |= (j < K && p < P) ? (1u << u) : 0u;
                }
            } else {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int64_t j = j0 + br + 2 * u;
                    pb1[u] = (j < K && p < P) ? V[j * P + p] : 0.0;
                }
            }
        };
        auto store = [&]() {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                As[ar * kKrApad + ac + 2 * v] = pa[v].x;
                As[ar * kKrApad + ac + 2 * v + 1] = pa[v].y;
            }
            if constexpr (kPairB) {
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    *reinterpret_cast<f64x2*>(&Bs[(br + 4 * u) * kKrBpad + bc]) =
                        ((mb >> u) & 1u) ? pb2[u] : f64x2{0.0, 0.0};
            } else {
#pragma unroll
                for (int u = 0; u < 8; ++u) Bs[(br + 2 * u) * kKrBpad + bc] = pb1[u];
            }
        };

        f64x4 acc[kKrRF][4];
#pragma unroll
        for (int rt = 0; rt < kKrRF; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = f64x4{0.0, 0.0, 0.0, 0.0};

        // columns left of the diagonal block carry nothing in M2.  The K steps run
        // from the last one down to the diagonal block: every row tile starts at the
        // same rows of V, so the row tiles of one column tile (resident together on
        // one XCD) read each V row at about the same time and share it through L2.
        const int64_t jlast = ((K - 1) / kKrK) * kKrK;
        load(jlast);
        for (int64_t j0 = jlast; j0 >= i0; j0 -= kKrK) {
            __syncthreads();
            store();
            __syncthreads();
            if (j0 - kKrK >= i0) load(j0 - kKrK);
#pragma unroll
            for (int kk = 0; kk < kKrK; kk += 4) {
                double a[kKrRF], b[4];
#pragma unroll
                for (int rt = 0; rt < kKrRF; ++rt) a[rt] = As[(wr * (kKrRows / 2) + rt * 16 + l16) * kKrApad + kk + g];
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) b[ct] = Bs[(kk + g) * kKrBpad + wc * 64 + ct * 16 + l16];
#pragma unroll
                for (int rt = 0; rt < kKrRF; ++rt)
#pragma unroll
                    for (int ct = 0; ct < 4; ++ct)
                        acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt], b[ct], acc[rt][ct], 0, 0, 0);
            }
        }

        // epilogue: per column, sum over this wave's kKrRows / 2 rows of V o W
        double se[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            const int64_t p = p0 + wc * 64 + ct * 16 + l16;
            double e = 0.0;
#pragma unroll
            for (int rt = 0; rt < kKrRF; ++rt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t i = i0 + wr * (kKrRows / 2) + rt * 16 + g + 4 * r;
                    if (i < K && p < P) e += V[i * P + p] * acc[rt][ct][r];
                }
            // the four lane groups holding this column, in a fixed order
            e += __shfl_xor(e, 16);
            e += __shfl_xor(e, 32);
            se[ct] = e;
        }
        if (g == 0) {
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) red[wr][wc * 64 + ct * 16 + l16] = se[ct];
        }
        __syncthreads();
        if (t < kKrPts) {
            const int64_t p = p0 + t;
            if (p < P) part_e[rt_idx * P + p] = red[0][t] + red[1][t];
        }
        __syncthreads();  // red is rewritten by the next tile's epilogue
    }
}

__global__ void k_krige_reduce(int64_t tiles, int64_t P, const double* __restrict__ part_e,
                               double* __restrict__ error) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double e = 0.0;
    for (int64_t it = 0; it < tiles; ++it) e += part_e[it * P + p];
    error[p] = e;
}

// a_j = sum_i cond_i M[i, j] in two passes: partial sums over kWChunks row
// chunks (block = 64 columns x 4 waves, each wave every 4th row of the chunk,
// the waves added in a fixed order), then the chunks in order
constexpr int kWChunks = 16;

__global__ __launch_bounds__(256) void k_krige_weights_part(int64_t K, const double* __restrict__ M,
                                                            const double* __restrict__ cond,
                                                            double* __restrict__ part) {
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t j = (int64_t)blockIdx.x * 64 + lane;
    const int64_t rows = (K + kWChunks - 1) / kWChunks;
    const int64_t r0 = (int64_t)blockIdx.y * rows;
    const int64_t r1 = r0 + rows < K ? r0 + rows : K;
    double s = 0.0;
    if (j < K)
        for (int64_t i = r0 + wave; i < r1; i += 4) s += cond[i] * M[i * K + j];
    red[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && j < K) part[blockIdx.y * K + j] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

__global__ __launch_bounds__(256) void k_krige_weights_sum(int64_t K, const double* __restrict__ part,
                                                           double* __restrict__ a) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= K) return;
    double s = 0.0;
    for (int c = 0; c < kWChunks; ++c) s += part[c * K + j];
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

// The same sum for even P and a 16-byte aligned V: four consecutive points per
// thread, read as two 16-byte loads per row (2 KB contiguous per wave and row),
// rows unrolled by 4 for loads in flight; each point still adds its terms in
// row order.
__global__ __launch_bounds__(256) void k_krige_field4(int64_t K, int64_t P, const double* __restrict__ a,
                                                      const double* __restrict__ V, double* __restrict__ field) {
    typedef __attribute__((ext_vector_type(2))) double f64x2;
    const int64_t p = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (p >= P) return;
    const bool hi = p + 2 < P;  // P even: the second pair is whole or absent
    f64x2 f0 = {0.0, 0.0}, f1 = {0.0, 0.0};
#pragma unroll 4
    for (int64_t j = 0; j < K; ++j) {
        const double aj = a[j];
        const f64x2* row = reinterpret_cast<const f64x2*>(V + j * P + p);
        const f64x2 v0 = row[0];
        f0.x += aj * v0.x;
        f0.y += aj * v0.y;
        if (hi) {
            const f64x2 v1 = row[1];
            f1.x += aj * v1.x;
            f1.y += aj * v1.y;
        }
    }
    *reinterpret_cast<f64x2*>(field + p) = f0;
    if (hi) *reinterpret_cast<f64x2*>(field + p + 2) = f1;
}

struct CovModelArgs {
    int kind;           // 0 Gaussian exp(-h^2), 1 Exponential exp(-h), 2 Stable exp(-h^alpha),
                        // 3 Spherical 1 - 1.5 h + 0.5 h^3 (h <= 1)
    double var;         // variance
    double len;         // rescaled length scale: h = r / len
    double alpha;       // Stable shape
    double sill_at_0;   // >= 0: value where r <= 1e-8 (cov_nugget); < 0: none
};

__device__ __forceinline__ double cov_of(const CovModelArgs& m, double r) {
    if (m.sill_at_0 >= 0.0 && r <= 1e-8) return m.sill_at_0;
    const double h = r / m.len;
    double c;
    switch (m.kind) {
        case 0: c = exp(-(h * h)); break;
        case 1: c = exp(-h); break;
        case 2: c = exp(-pow(h, m.alpha)); break;
        default: {
            const double g = fmin(h, 1.0);
            c = 1.0 - 1.5 * g + 0.5 * (g * g * g);
        }
    }
    return m.var * c;
}

// out[i * ld + p] = C(|c_i - x_p|), i < K, p < P.  A thread keeps its target in
// registers and writes kCovRows rows (conditioning points i0 .. i0 + kCovRows - 1,
// read as wave-uniform values), so the targets are read once per kCovRows rows.
constexpr int kCovRows = 16;

template <int DIM>
__global__ __launch_bounds__(256) void k_krige_cov(CovModelArgs m, int64_t K, int64_t P,
                                                   const double* __restrict__ cpos,
                                                   const double* __restrict__ pos, double* __restrict__ out,
                                                   int64_t ld) {
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.y * kCovRows;
    if (p >= P) return;
    double x[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) x[d] = pos[d * P + p];
    const int rows = (int)(K - i0 < kCovRows ? K - i0 : kCovRows);
    for (int r = 0; r < rows; ++r) {
        const int64_t i = i0 + r;
        double d2 = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double t = cpos[d * K + i] - x[d];
            d2 += t * t;
        }
        out[i * ld + p] = cov_of(m, sqrt(d2));
    }
}

__global__ void k_fill_zero_f64(double* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

inline int64_t krige_row_tiles(int64_t K) { return ceil_div(K, (int64_t)kKrRows); }

// field: the weight vector (K) and its row-chunk partials (kWChunks x K);
// variance form: also the row-tile partials (tiles x P) and the padded M2
// (128 tiles x Kp), 16-byte aligned after the others (K even or padded by one)
inline int64_t krige_ws_bytes(int64_t K, int64_t P, bool variance) {
    const int64_t field = (1 + kWChunks) * K;
    if (variance) {
        const int64_t head = krige_row_tiles(K) * P + field;
        return ((head + 1) / 2 * 2 + krige_row_tiles(K) * kKrRows * krige_kp(K)) * (int64_t)sizeof(double);
    }
    return field * (int64_t)sizeof(double);
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
    const int64_t tiles = gsx::krige_row_tiles(K);
    const int64_t col_blocks = gsx::ceil_div(P, (int64_t)gsx::kKrPts);
    const int64_t Kp = gsx::krige_kp(K), Rp = tiles * gsx::kKrRows;
    if (variance && (tiles * (col_blocks + 8) > 0x7fffffffLL || gsx::ceil_div(Rp * Kp, (int64_t)256) > 0x7fffffffLL))
        return GSX_E_INVALID;
    double* part_e = ws;
    double* a = variance ? ws + tiles * P : ws;
    double* wpart = a + K;
    // field = (M^T cond) . v
    hipLaunchKernelGGL(gsx::k_krige_weights_part, dim3((unsigned)gsx::ceil_div(K, (int64_t)64), gsx::kWChunks),
                       dim3(256), 0, s, K, krig_mat, cond, wpart);
    int st = gsx::launch_status();
    if (st != GSX_OK) return st;
    hipLaunchKernelGGL(gsx::k_krige_weights_sum, dim3((unsigned)gsx::ceil_div(K, (int64_t)256)), dim3(256), 0, s, K,
                       wpart, a);
    st = gsx::launch_status();
    if (st != GSX_OK) return st;
    const bool vec = P % 2 == 0 && reinterpret_cast<uintptr_t>(krig_vecs) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(field) % 16 == 0;
    if (vec)
        hipLaunchKernelGGL(gsx::k_krige_field4, dim3((unsigned)gsx::ceil_div(P, (int64_t)1024)), dim3(256), 0, s, K,
                           P, a, krig_vecs, field);
    else
        hipLaunchKernelGGL(gsx::k_krige_field, dim3((unsigned)gsx::ceil_div(P, (int64_t)512)), dim3(256), 0, s, K,
                           P, a, krig_vecs, field);
    st = gsx::launch_status();
    if (st != GSX_OK || !variance) return st;
    // error = v^T M2 v over the block upper triangle (M2 16-byte aligned: the
    // workspace is, and its offset is rounded up to an even number of doubles)
    const int64_t head = tiles * P + (1 + gsx::kWChunks) * K;
    double* m2 = ws + (head + 1) / 2 * 2;
    if (reinterpret_cast<uintptr_t>(m2) % 16 != 0) return GSX_E_INVALID;
    hipLaunchKernelGGL(gsx::k_krige_sym, dim3((unsigned)gsx::ceil_div(Rp * Kp, (int64_t)256)), dim3(256), 0, s, K, Rp,
                       Kp, krig_mat, m2);
    st = gsx::launch_status();
    if (st != GSX_OK) return st;
    const int64_t col8 = gsx::ceil_div(col_blocks, (int64_t)8) * 8;
    const dim3 grid((unsigned)(col8 * ((tiles + 1) / 2)));
    const bool pair_b = P % 2 == 0 && reinterpret_cast<uintptr_t>(krig_vecs) % 16 == 0;
    if (pair_b)
        hipLaunchKernelGGL(gsx::k_krige_gemm<true>, grid, dim3(256), 0, s, K, Kp, P, tiles, col_blocks, m2, krig_vecs,
                           part_e);
    else
        hipLaunchKernelGGL(gsx::k_krige_gemm<false>, grid, dim3(256), 0, s, K, Kp, P, tiles, col_blocks, m2,
                           krig_vecs, part_e);
    st = gsx::launch_status();
    if (st != GSX_OK) return st;
    hipLaunchKernelGGL(gsx::k_krige_reduce, dim3((unsigned)gsx::ceil_div(P, (int64_t)256)), dim3(256), 0, s, tiles,
                       P, part_e, error);
    return gsx::launch_status();
}

int gsx_krige_cov_f64(int model_kind, double var, double len_rescaled, double alpha, double sill_at_zero,
                      int64_t dim, int64_t K, int64_t P, const double* cond_pos, const double* pos, double* out,
                      int64_t ld_out, void* stream) {
    if (model_kind < 0 || model_kind > 3 || dim < 1 || dim > 3 || K < 0 || P < 0 || ld_out < P) return GSX_E_INVALID;
    if (!(len_rescaled > 0.0)) return GSX_E_INVALID;
    if (K == 0 || P == 0) return GSX_OK;
    if (K > 65535) return GSX_E_INVALID;
    if (!cond_pos || !pos || !out) return GSX_E_INVALID;
    gsx::CovModelArgs m{model_kind, var, len_rescaled, alpha, sill_at_zero};
    hipStream_t s = gsx::as_stream(stream);
    const dim3 grid((unsigned)gsx::ceil_div(P, (int64_t)256), (unsigned)gsx::ceil_div(K, (int64_t)gsx::kCovRows));
    switch (dim) {
        case 1: hipLaunchKernelGGL((gsx::k_krige_cov<1>), grid, dim3(256), 0, s, m, K, P, cond_pos, pos, out, ld_out); break;
        case 2: hipLaunchKernelGGL((gsx::k_krige_cov<2>), grid, dim3(256), 0, s, m, K, P, cond_pos, pos, out, ld_out); break;
        default: hipLaunchKernelGGL((gsx::k_krige_cov<3>), grid, dim3(256), 0, s, m, K, P, cond_pos, pos, out, ld_out);
    }
    return gsx::launch_status();
}

}  // extern "C"
