// FP64 GEMMs on the matrix cores for structured grids: the Fourier method with
// a tensor grid of modes (one contraction per axis, below) and, further down,
// the randomisation-method grid sum as one FP64 GEMM (precision "fp64gemm").
//
// Fourier method on structured grids with a tensor grid of modes.
//
// The Fourier generator's modes are the 'ij' meshgrid of one wavenumber axis
// per dimension (generator.py:838-862), so on a structured grid the sum
//
//   u(x) = sum_m f_m (z1_m cos(k_m . x) + z2_m sin(k_m . x))
//        = Re sum_{m1..md} c[m1..md] e^{i k1(m1) x1} ... e^{i kd(md) xd},
//   c = f (z1 - i z2),
//
// factorises into one contraction per axis: real FP64 GEMMs over split
// complex operands (a complex product T = E X is the real product
// [[Er, -Ei], [Ei, Er]] [Xr; Xi]).  With n_i grid points and M_i modes per
// axis (d = 3):
//
//   first   T1[x][(part), (m2 m3)] = sum_m1 E1[x][m1] c[m1][(m2 m3)]
//             A rows (x, part) x cols (part', m1), B = [cr; ci] (2 M1 x M2 M3)
//   middle  T2[x][y][(part), m3] = sum_m2 E2[y][m2] T1[x][m2][m3]
//             batched over x: B(x) = the 2 M2 rows of T1 for x (real, imag)
//   last    u[(x y)][z] = Re sum_m3 T2[(x y)][m3] E3[z][m3]
//             A = T2 rows (x y) of [T2r, T2i], B = [E3r^T; -E3i^T]
//
// (d = 2: first and last; d = 1: the last stage on c itself).  The work is
// about 4 (n1 M1 M2 M3 + n1 n2 M2 M3 + n1 n2 n3 M3) flops (d = 3) instead of
// 12 P N for the per-point sum: 2e10 flops against 1.3e16 for a 4096^2 grid
// with 256^2 modes.  Everything is FP64: the phase tables are cos / sin of
// the products k_i(m) x_i (correctly rounded), the GEMMs run on the FP64
// matrix cores (v_mfma_f64_16x16x4).  The result differs from the per-point
// sum only by the rounding of the phase sum k1 x1 + k2 x2 (+ k3 x3), which
// the per-point form rounds and this form does not.
#include "gsx_internal.h"
#include "gsx_math.h"

namespace gsx {

typedef __attribute__((ext_vector_type(4))) double f64x4;

constexpr int kFR = 128;     // C rows per block
constexpr int kFK = 16;      // K per LDS stage
constexpr int kFApad = kFK + 1;

// Batched real FP64 GEMM, row-major: C_b = A_b B_b, A_b (M x K, lda), B_b (K x N,
// ldb), C_b (M x N, ldc), operand b at A + b sA, B + b sB, C + b sC.  Block =
// (batch, row tile, column tile) of 128 x 64 WC: 4 waves in (4 / WC) x WC, each
// (128 WC / 4) rows x 64 columns of 16 x 16 MFMA tiles (WC = 2: 4 x 4; WC = 1,
// for launches with few tiles or N <= 64: 2 x 4); A and B tiles staged through
// LDS with the next stage prefetched into registers; out-of-range operand
// entries are zero.
// v_mfma_f64_16x16x4 layouts (g = lane / 16): A: lane holds row lane % 16, K
// element g; B: column lane % 16, K element g; C: column lane % 16, register r
// holds row g + 4 r.
template <int WC>
__global__ __launch_bounds__(256, 2) void k_fourier_dgemm(int64_t M, int64_t N, int64_t K,
                                                          const double* __restrict__ A, int64_t lda, int64_t sA,
                                                          const double* __restrict__ B, int64_t ldb, int64_t sB,
                                                          double* __restrict__ C, int64_t ldc, int64_t sC,
                                                          int64_t row_tiles, int64_t col_tiles) {
    constexpr int kBC = 64 * WC;            // block columns
    constexpr int kBpad = kBC + 16;
    constexpr int kRF = kFR * WC / 4 / 16;  // 16-row MFMA tiles per wave
    constexpr int kBT = 256 / kBC;          // B rows staged per pass
    __shared__ double As[kFR * kFApad];
    __shared__ double Bs[kFK * kBpad];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wr = wave / WC, wc = wave % WC;
    const int g = lane >> 4, l16 = lane & 15;
    const int64_t per_batch = row_tiles * col_tiles;
    const int64_t bt = (int64_t)blockIdx.x / per_batch;
    const int64_t rest = (int64_t)blockIdx.x - bt * per_batch;
    const int64_t rt = rest / col_tiles, ct = rest - rt * col_tiles;
    A += bt * sA;
    B += bt * sB;
    C += bt * sC;
    const int64_t i0 = rt * kFR, p0 = ct * kBC;

    // staging: A 128 x 16 (two threads per row, 8 consecutive each); B 16 x kBC
    // (column t % kBC of rows t / kBC + kBT u: unit-stride loads and LDS stores)
    const int ar = t >> 1, ac = (t & 1) * 8;
    const int br = t / kBC, bc = t % kBC;
    double pa[8], pb[kFK / kBT];
    auto load = [&](int64_t j0) {
        const int64_t row = i0 + ar;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int64_t col = j0 + ac + u;
            pa[u] = (row < M && col < K) ? A[row * lda + col] : 0.0;
        }
        const int64_t p = p0 + bc;
#pragma unroll
        for (int u = 0; u < kFK / kBT; ++u) {
            const int64_t j = j0 + br + kBT * u;
            pb[u] = (j < K && p < N) ? B[j * ldb + p] : 0.0;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int u = 0; u < 8; ++u) As[ar * kFApad + ac + u] = pa[u];
#pragma unroll
        for (int u = 0; u < kFK / kBT; ++u) Bs[(br + kBT * u) * kBpad + bc] = pb[u];
    };

    f64x4 acc[kRF][4];
#pragma unroll
    for (int r = 0; r < kRF; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = f64x4{0.0, 0.0, 0.0, 0.0};

    load(0);
    for (int64_t j0 = 0; j0 < K; j0 += kFK) {
        __syncthreads();
        store();
        __syncthreads();
        if (j0 + kFK < K) load(j0 + kFK);
#pragma unroll
        for (int kk = 0; kk < kFK; kk += 4) {
            double a[kRF], b[4];
#pragma unroll
            for (int r = 0; r < kRF; ++r) a[r] = As[(wr * kRF * 16 + r * 16 + l16) * kFApad + kk + g];
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] = Bs[(kk + g) * kBpad + wc * 64 + c * 16 + l16];
#pragma unroll
            for (int r = 0; r < kRF; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int64_t p = p0 + wc * 64 + c * 16 + l16;
        if (p >= N) continue;
#pragma unroll
        for (int r = 0; r < kRF; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t row = i0 + wr * kRF * 16 + r * 16 + g + 4 * q;
                if (row < M) C[row * ldc + p] = acc[r][c][q];
            }
    }
}

__global__ void k_fourier_zero(double* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__device__ __forceinline__ void phase_sincos(double th, double* s, double* c) {
    if (fabs(th) <= kFastPhaseMax) sincos_fast(th, s, c); else sincos(th, s, c);
}

// c = f (z1 - i z2) as [cr; ci]: 2 x N doubles (N = all modes, C order)
__global__ void k_fourier_coeff(int64_t N, const double* __restrict__ fac, const double* __restrict__ z1,
                                const double* __restrict__ z2, double* __restrict__ c) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    c[j] = fac[j] * z1[j];
    c[N + j] = -(fac[j] * z2[j]);
}

// Rotation-form table of one axis, rows (point, part) x columns (part', m):
// row (i, 0) = [cos(k_m x_i), -sin(k_m x_i)], row (i, 1) = [sin, cos] (2 n x 2 M)
__global__ void k_fourier_rot_table(int64_t n, int64_t M, const double* __restrict__ x,
                                    const double* __restrict__ k, double* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * M) return;
    const int64_t i = e / M, m = e - i * M;
    double s, c;
    phase_sincos(k[m] * x[i], &s, &c);
    double* r0 = out + (2 * i) * (2 * M);
    double* r1 = r0 + 2 * M;
    r0[m] = c;
    r0[M + m] = -s;
    r1[m] = s;
    r1[M + m] = c;
}

// Last-stage table [cos(k_m x_i); -sin(k_m x_i)] as rows (part, m) x columns i (2 M x n)
__global__ void k_fourier_last_table(int64_t n, int64_t M, const double* __restrict__ x,
                                     const double* __restrict__ k, double* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * M) return;
    const int64_t m = e / n, i = e - m * n;
    double s, c;
    phase_sincos(k[m] * x[i], &s, &c);
    out[m * n + i] = c;
    out[(M + m) * n + i] = -s;
}

struct FourierTensorPlan {
    int64_t d, N, P;
    int64_t M[3], n[3];
    int64_t off_c, off_e1, off_e2, off_t1, off_t2, off_last, bytes;  // workspace offsets (doubles)
};

inline bool fourier_tensor_plan(int64_t d, const int64_t* mode_no, const int64_t* shape, FourierTensorPlan& p) {
    if (d < 1 || d > 3 || !mode_no || !shape) return false;
    p.d = d;
    p.N = 1;
    p.P = 1;
    for (int i = 0; i < 3; ++i) {
        p.M[i] = i < d ? mode_no[i] : 1;
        p.n[i] = i < d ? shape[i] : 1;
        if (p.M[i] < 0 || p.n[i] < 0) return false;
        p.N *= p.M[i];
        p.P *= p.n[i];
    }
    const int64_t Md = p.M[d - 1], nd = p.n[d - 1];
    int64_t o = 0;
    auto take = [&](int64_t doubles) {
        const int64_t at = o;
        o += (doubles + 31) / 32 * 32;  // 256-byte aligned
        return at;
    };
    p.off_c = take(2 * p.N);
    p.off_e1 = d >= 2 ? take(2 * p.n[0] * 2 * p.M[0]) : 0;
    p.off_e2 = d == 3 ? take(2 * p.n[1] * 2 * p.M[1]) : 0;
    p.off_t1 = d >= 2 ? take(2 * p.n[0] * p.M[1] * p.M[2]) : 0;  // 2 n1 x (M2 M3)
    p.off_t2 = d == 3 ? take(p.n[0] * 2 * p.n[1] * p.M[2]) : 0;
    p.off_last = take(2 * Md * nd);
    p.bytes = o * 8;
    return true;
}

inline int dgemm(int64_t M, int64_t N, int64_t K, int64_t batch, const double* A, int64_t lda, int64_t sA,
                 const double* B, int64_t ldb, int64_t sB, double* C, int64_t ldc, int64_t sC, hipStream_t s) {
    if (M == 0 || N == 0 || batch == 0) return GSX_OK;
    const int64_t rt = ceil_div(M, (int64_t)kFR);
    // 128-column tiles unless that leaves the chip (256 CUs) under-filled or
    // half of each tile empty: then 64-column tiles
    const int64_t ct2 = ceil_div(N, (int64_t)128), ct1 = ceil_div(N, (int64_t)64);
    const bool narrow = N <= 64 || batch * rt * ct2 < 256;
    const int64_t ct = narrow ? ct1 : ct2;
    if (batch * rt * ct > 0x7fffffffLL) return GSX_E_INVALID;
    if (narrow)
        hipLaunchKernelGGL(k_fourier_dgemm<1>, dim3((unsigned)(batch * rt * ct)), dim3(256), 0, s, M, N, K, A, lda,
                           sA, B, ldb, sB, C, ldc, sC, rt, ct);
    else
        hipLaunchKernelGGL(k_fourier_dgemm<2>, dim3((unsigned)(batch * rt * ct)), dim3(256), 0, s, M, N, K, A, lda,
                           sA, B, ldb, sB, C, ldc, sC, rt, ct);
    return launch_status();
}

}  // namespace gsx

extern "C" {

int64_t gsx_summate_fourier_tensor_workspace(int64_t d, const int64_t* mode_no, const int64_t* shape) {
    gsx::FourierTensorPlan p;
    if (!gsx::fourier_tensor_plan(d, mode_no, shape, p)) return GSX_E_INVALID;
    return p.bytes;
}

int gsx_summate_fourier_tensor_f64(int64_t d, const int64_t* mode_no, const double* kaxes, const int64_t* shape,
                                   const double* axes, const double* spectrum_factor, const double* z1,
                                   const double* z2, double* out, void* workspace, int64_t workspace_bytes,
                                   void* stream) {
    gsx::FourierTensorPlan p;
    if (!gsx::fourier_tensor_plan(d, mode_no, shape, p)) return GSX_E_INVALID;
    if (p.P == 0) return GSX_OK;
    if (!axes || !out) return GSX_E_INVALID;
    hipStream_t s = gsx::as_stream(stream);
    if (p.N == 0) {  // empty sum
        hipLaunchKernelGGL(gsx::k_fourier_zero, dim3((unsigned)gsx::ceil_div(p.P, (int64_t)256)), dim3(256), 0, s,
                           out, p.P);
        return gsx::launch_status();
    }
    if (!kaxes || !spectrum_factor || !z1 || !z2) return GSX_E_INVALID;
    if (!workspace || workspace_bytes < p.bytes) return GSX_E_WORKSPACE;
    double* ws = static_cast<double*>(workspace);
    const double* ax[3] = {axes, axes + p.n[0], axes + p.n[0] + p.n[1]};
    const double* ka[3] = {kaxes, kaxes + p.M[0], kaxes + p.M[0] + p.M[1]};
    double* c = ws + p.off_c;
    auto blocks = [](int64_t n) { return dim3((unsigned)gsx::ceil_div(n, (int64_t)256)); };
    if (gsx::ceil_div(p.N, (int64_t)256) > 0x7fffffffLL) return GSX_E_INVALID;
    for (int i = 0; i < d; ++i)
        if (gsx::ceil_div(p.n[i] * p.M[i], (int64_t)256) > 0x7fffffffLL) return GSX_E_INVALID;
    hipLaunchKernelGGL(gsx::k_fourier_coeff, blocks(p.N), dim3(256), 0, s, p.N, spectrum_factor, z1, z2, c);
    int st = gsx::launch_status();
    if (st != GSX_OK) return st;
    const int64_t Md = p.M[d - 1], nd = p.n[d - 1];
    double* last = ws + p.off_last;
    hipLaunchKernelGGL(gsx::k_fourier_last_table, blocks(nd * Md), dim3(256), 0, s, nd, Md, ax[d - 1], ka[d - 1],
                       last);
    if ((st = gsx::launch_status()) != GSX_OK) return st;
    if (d == 1)  // u (1 x n1) = [cr, ci] (1 x 2 M1) . [cos; -sin] (2 M1 x n1)
        return gsx::dgemm(1, nd, 2 * Md, 1, c, 2 * Md, 0, last, nd, 0, out, nd, 0, s);
    // first stage: T1 (2 n1 x R) = E1rot (2 n1 x 2 M1) . [cr; ci] (2 M1 x R), R = M2 ... Md
    const int64_t R = p.M[1] * p.M[2];
    double* e1 = ws + p.off_e1;
    double* t1 = ws + p.off_t1;
    hipLaunchKernelGGL(gsx::k_fourier_rot_table, blocks(p.n[0] * p.M[0]), dim3(256), 0, s, p.n[0], p.M[0], ax[0],
                       ka[0], e1);
    if ((st = gsx::launch_status()) != GSX_OK) return st;
    // [cr; ci] (2 M1 x R): c holds all cr then all ci, mode index (m1, rest) in C order
    st = gsx::dgemm(2 * p.n[0], R, 2 * p.M[0], 1, e1, 2 * p.M[0], 0, c, R, 0, t1, R, 0, s);
    if (st != GSX_OK) return st;
    if (d == 2)  // u (n1 x n2) = T1 as n1 rows of [T1r, T1i] (2 M2) . last (2 M2 x n2)
        return gsx::dgemm(p.n[0], nd, 2 * Md, 1, t1, 2 * Md, 0, last, nd, 0, out, nd, 0, s);
    // middle stage, batched over x: T2(x) (2 n2 x M3) = E2rot (2 n2 x 2 M2) . T1(x) (2 M2 x M3)
    double* e2 = ws + p.off_e2;
    double* t2 = ws + p.off_t2;
    hipLaunchKernelGGL(gsx::k_fourier_rot_table, blocks(p.n[1] * p.M[1]), dim3(256), 0, s, p.n[1], p.M[1], ax[1],
                       ka[1], e2);
    if ((st = gsx::launch_status()) != GSX_OK) return st;
    st = gsx::dgemm(2 * p.n[1], p.M[2], 2 * p.M[1], p.n[0], e2, 2 * p.M[1], 0, t1, p.M[2], 2 * R, t2, p.M[2],
                    2 * p.n[1] * p.M[2], s);
    if (st != GSX_OK) return st;
    // last stage: u (n1 n2 x n3) = T2 as n1 n2 rows of [T2r, T2i] (2 M3) . last (2 M3 x n3)
    return gsx::dgemm(p.n[0] * p.n[1], nd, 2 * Md, 1, t2, 2 * Md, 0, last, nd, 0, out, nd, 0, s);
}

}  // extern "C"

// ------------------------------------------------------------------------
// Randomisation-method sum on a structured grid as an FP64 GEMM.
//
// u[x][(y z)] = Re sum_j c_j e^{i k0_j x} e^{i k1_j y} e^{i k2_j z},
// c_j = f_j (z1_j - i z2_j): rows are axis 0, columns the flattened (y z),
// A[x][(part, j)] = [cos(k0_j x), -sin(k0_j x)] (n0 x 2N) and
// B[(part, j)][col] = [Re; Im] of c_j e^{i k1_j y} e^{i k2_j z} (2N x n1 n2),
// built in column chunks of bounded size and multiplied on the FP64 matrix
// cores (k_fourier_dgemm).  The phase of each axis is the rounded product
// k_i x_i, as in the per-point sum; the sum of those phases is not rounded,
// so the result differs from the per-point sum by about |k . x| 2^-52
// (relative) per term.
namespace gsx {

constexpr int64_t kGemmChunkBytes = int64_t(1) << 29;  // B chunk: at most 512 MB

// A (n0 x 2N): row x = [cos(k0_j x) (N), -sin(k0_j x) (N)]
__global__ void k_grid_gemm_a(int64_t n0, int64_t N, const double* __restrict__ x, const double* __restrict__ k0,
                              double* __restrict__ a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n0 * N) return;
    const int64_t i = e / N, j = e - i * N;
    double s, c;
    phase_sincos(k0[j] * x[i], &s, &c);
    a[i * 2 * N + j] = c;
    a[i * 2 * N + N + j] = -s;
}

// E (N x n) complex, mode-major: [j][i] = (cos, sin)(k_j x_i) (k null: 1); with
// coeff, times f_j (z1_j - i z2_j)
__global__ void k_grid_gemm_e(int64_t n, int64_t N, const double* __restrict__ x, const double* __restrict__ k,
                              const double* __restrict__ fac, const double* __restrict__ z1,
                              const double* __restrict__ z2, bool coeff, double2* __restrict__ e) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * N) return;
    const int64_t j = t / n, i = t - j * n;
    double s = 0.0, c = 1.0;
    if (k) phase_sincos(k[j] * x[i], &s, &c);
    if (coeff) {
        const double f = fac ? fac[j] : 1.0;
        const double cr = f * z1[j], ci = -(f * z2[j]);
        e[t] = double2{cr * c - ci * s, cr * s + ci * c};
    } else {
        e[t] = double2{c, s};
    }
}

// B chunk (2N x C): columns col0 .. col0 + C - 1 of the flattened (i1, i2);
// e1: c_j e^{i k1_j y} (N x n1), e2: e^{i k2_j z} (N x n2) or null (2D), mode-major
// so that neighbouring columns read neighbouring entries
__global__ void k_grid_gemm_b(int64_t N, int64_t n2, int64_t col0, int64_t C, int64_t cols,
                              const double2* __restrict__ e1, const double2* __restrict__ e2,
                              double* __restrict__ b) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * C) return;
    const int64_t j = t / C, cc = t - j * C;
    const int64_t col = col0 + cc;
    double re = 0.0, im = 0.0;
    if (col < cols) {
        const int64_t i1 = e2 ? col / n2 : col, i2 = e2 ? col - i1 * n2 : 0;
        const double2 u = e1[j * (cols / n2) + i1];
        if (e2) {
            const double2 v = e2[j * n2 + i2];
            re = u.x * v.x - u.y * v.y;
            im = u.x * v.y + u.y * v.x;
        } else {
            re = u.x;
            im = u.y;
        }
    }
    b[j * C + cc] = re;
    b[(N + j) * C + cc] = im;
}

struct GridGemmPlan {
    int64_t d, N, n0, n1, n2, cols, C;
    int64_t off_a, off_e1, off_e2, off_b, bytes;  // doubles
};

inline bool grid_gemm_plan(int64_t d, int64_t N, const int64_t* shape, GridGemmPlan& p) {
    if (d < 1 || d > 3 || N < 0 || !shape) return false;
    for (int i = 0; i < d; ++i)
        if (shape[i] < 0) return false;
    p.d = d;
    p.N = N;
    p.n0 = shape[0];
    p.n1 = d >= 2 ? shape[1] : 1;
    p.n2 = d == 3 ? shape[2] : 1;
    p.cols = p.n1 * p.n2;
    const int64_t per_col = 2 * (N > 0 ? N : 1) * 8;
    int64_t C = kGemmChunkBytes / per_col;
    C = C >= 128 ? C / 128 * 128 : (C > 0 ? C : 1);
    p.C = C < p.cols ? C : (p.cols > 0 ? p.cols : 1);
    int64_t o = 0;
    auto take = [&](int64_t doubles) {
        const int64_t at = o;
        o += (doubles + 31) / 32 * 32;
        return at;
    };
    p.off_a = take(p.n0 * 2 * N);
    p.off_e1 = take(2 * p.n1 * N);
    p.off_e2 = d == 3 ? take(2 * p.n2 * N) : 0;
    p.off_b = take(2 * N * p.C);
    p.bytes = o * 8;
    return true;
}

}  // namespace gsx

extern "C" {

int64_t gsx_summate_grid_gemm_workspace(int64_t d, int64_t N, const int64_t* shape) {
    gsx::GridGemmPlan p;
    if (!gsx::grid_gemm_plan(d, N, shape, p)) return GSX_E_INVALID;
    return p.bytes;
}

int gsx_summate_grid_gemm_f64(int64_t d, int64_t N, const int64_t* shape, const double* axes, const double* cov,
                              const double* z1, const double* z2, const double* spectrum_factor, double* out,
                              void* workspace, int64_t workspace_bytes, void* stream) {
    gsx::GridGemmPlan p;
    if (!gsx::grid_gemm_plan(d, N, shape, p)) return GSX_E_INVALID;
    const int64_t P = p.n0 * p.cols;
    if (P == 0) return GSX_OK;
    if (!axes || !out) return GSX_E_INVALID;
    hipStream_t s = gsx::as_stream(stream);
    auto blocks = [](int64_t n) { return dim3((unsigned)gsx::ceil_div(n, (int64_t)256)); };
    if (N == 0) {
        if (gsx::ceil_div(P, (int64_t)256) > 0x7fffffffLL) return GSX_E_INVALID;
        hipLaunchKernelGGL(gsx::k_fourier_zero, blocks(P), dim3(256), 0, s, out, P);
        return gsx::launch_status();
    }
    if (!cov || !z1 || !z2) return GSX_E_INVALID;
    if (!workspace || workspace_bytes < p.bytes) return GSX_E_WORKSPACE;
    const int64_t big = p.n0 * N > p.n1 * N ? p.n0 * N : p.n1 * N;
    if (gsx::ceil_div(big, (int64_t)256) > 0x7fffffffLL || gsx::ceil_div(p.n2 * N, (int64_t)256) > 0x7fffffffLL ||
        gsx::ceil_div(N * p.C, (int64_t)256) > 0x7fffffffLL)
        return GSX_E_INVALID;
    double* ws = static_cast<double*>(workspace);
    double* a = ws + p.off_a;
    double2* e1 = reinterpret_cast<double2*>(ws + p.off_e1);
    double2* e2 = d == 3 ? reinterpret_cast<double2*>(ws + p.off_e2) : nullptr;
    double* b = ws + p.off_b;
    const double* ax0 = axes;
    hipLaunchKernelGGL(gsx::k_grid_gemm_a, blocks(p.n0 * N), dim3(256), 0, s, p.n0, N, ax0, cov, a);
    int st = gsx::launch_status();
    if (st != GSX_OK) return st;
    // e1: c_j e^{i k1_j y} (d >= 2), or c_j alone for the one column of a 1D grid
    if (d >= 2)
        hipLaunchKernelGGL(gsx::k_grid_gemm_e, blocks(p.n1 * N), dim3(256), 0, s, p.n1, N, axes + p.n0, cov + N,
                           spectrum_factor, z1, z2, true, e1);
    else
        hipLaunchKernelGGL(gsx::k_grid_gemm_e, blocks(N), dim3(256), 0, s, (int64_t)1, N, nullptr, nullptr,
                           spectrum_factor, z1, z2, true, e1);
    if ((st = gsx::launch_status()) != GSX_OK) return st;
    if (d == 3) {
        hipLaunchKernelGGL(gsx::k_grid_gemm_e, blocks(p.n2 * N), dim3(256), 0, s, p.n2, N, axes + p.n0 + p.n1,
                           cov + 2 * N, nullptr, nullptr, nullptr, false, e2);
        if ((st = gsx::launch_status()) != GSX_OK) return st;
    }
    for (int64_t c0 = 0; c0 < p.cols; c0 += p.C) {
        const int64_t cn = p.cols - c0 < p.C ? p.cols - c0 : p.C;
        hipLaunchKernelGGL(gsx::k_grid_gemm_b, blocks(N * p.C), dim3(256), 0, s, N, p.n2, c0, p.C, p.cols, e1, e2,
                           b);
        if ((st = gsx::launch_status()) != GSX_OK) return st;
        st = gsx::dgemm(p.n0, cn, 2 * N, 1, a, 2 * N, 0, b, p.C, 0, out + c0, p.cols, 0, s);
        if (st != GSX_OK) return st;
    }
    return GSX_OK;
}

}  // extern "C"
