// Structured-grid randomisation-method sum on the gfx950 matrix cores.
//
// On a tensor grid (axes x_i0, y_i1, z_i2) the reference sum
// (src/gstools/field/generator.py:193-199, positions from generate_grid,
// tools/geometric.py:340-356) factorises by the angle-addition identity:
//
//   u[i0, i1, i2] = Re sum_j c_j e^{i k0_j x_i0} e^{i k1_j y_i1} e^{i k2_j z_i2},
//   c_j = z1_j - i z2_j,
//
// i.e. for every column (i1, i2) a real GEMM with K = 2N:
//
//   U (n0 x n1 n2) = A (n0 x 2N) . B (2N x n1 n2)
//   A[i0, 2j] = cos(k0_j x_i0),           A[i0, 2j+1] = -sin(k0_j x_i0)
//   B[2j, c]  = Re(E1_j(i1) E2_j(i2)),    B[2j+1, c]  = Im(E1_j(i1) E2_j(i2))
//   E1_j(i1)  = e^{i k1_j y_i1},          E2_j(i2)    = c_j e^{i k2_j z_i2}
//
// (2D grids use E1 = 1 on a one-point middle axis.)  Every operand is split
// into a high and a low 16-bit part (bf16 or fp16) and the product is
// A_hi B_hi + A_hi B_lo + A_lo B_hi on v_mfma_f32_16x16x32_{bf16,f16},
// accumulated in FP32.
//
// Kernels:
//  * k_mfma_tables: phase factors in FP64 (sincos of the FP64 phase), A split
//    into hi/lo 16-bit rows [i0][2Np], E1/E2 as complex FP32 rows [i][Np].
//  * k_mfma_grid: block = 4 waves, tile = 128 rows x 256 columns; the A tile
//    of each K step (16 modes = 32 real K) is staged through double-buffered
//    LDS and shared by the 4 waves; each wave owns 64 columns (4 fragments of
//    16) and builds its B fragments in registers from E1/E2 (complex product
//    + hi/lo split), so B never touches memory.  Block -> tile mapping is
//    XCD-aware: the blocks resident on one XCD work on one row tile of A,
//    so that XCD's L2 holds the A rows its blocks stream.
//
// MFMA operand layouts (v_mfma_f32_16x16x32, 64 lanes, g = lane / 16):
//   A: lane holds row (lane % 16), K elements 8g .. 8g+7
//   B: lane holds column (lane % 16), K elements 8g .. 8g+7
//   C: lane holds column (lane % 16), rows 4g .. 4g+3
// K element 2t (+1) of a lane's 8 is Re (Im) of mode 4g + t of the K step.
#include "gsx_internal.h"

namespace gsx {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
typedef __attribute__((ext_vector_type(2))) _Float16 f16x2;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;

constexpr int kMBM = 128;            // rows per block
constexpr int kMWaves = 4;
constexpr int kMCF = 4;              // 16-column fragments per wave
constexpr int kMRF = kMBM / 16;      // 16-row fragments per wave (all block rows)
constexpr int kMKM = 16;             // modes per K step

struct MfmaGeom {
    int64_t n0, n1, n2;   // grid (n1 = 1 for 2D grids)
    int64_t n0p, Np;      // padded rows / modes
    int64_t ncols;        // n1 * n2
    int64_t out_cols;     // valid flat columns i1 * n2 + i2: n1 * n2, or the split axis' length
    int64_t rs, c1s;      // output offset of (row, i1, i2) = row * rs + i1 * c1s + i2
    int g2;               // 1: k_mfma_grid3 (16 i1 x 16 i2 column tiles); 16: k_mfma_grid
    int64_t t1, t2;       // column tiles along i1 / i2 (k_mfma_grid: 1 x tiles of 256 flat columns)
    int64_t row_tiles, col_tiles;
    int64_t a_elems, e1_elems, e2_elems;
};

inline MfmaGeom mfma_geom(int64_t d, int64_t N, const int64_t* shape) {
    MfmaGeom g;
    g.n0 = shape[0];
    g.n1 = d == 3 ? shape[1] : 1;
    g.n2 = d == 3 ? shape[2] : shape[1];
    g.n0p = ceil_div(g.n0, kMBM) * kMBM;
    g.Np = ceil_div(N, kMKM) * kMKM;
    g.ncols = g.n1 * g.n2;
    g.out_cols = g.ncols;
    g.rs = g.ncols;
    g.c1s = g.n2;
    g.g2 = (g.n1 >= 16 && g.n2 >= 16) ? 1 : 16;
    g.t1 = g.g2 == 1 ? ceil_div(g.n1, (int64_t)16) : 1;
    g.t2 = g.g2 == 1 ? ceil_div(g.n2, (int64_t)16) : ceil_div(g.ncols, (int64_t)256);
    g.row_tiles = g.n0p / kMBM;
    g.col_tiles = g.t1 * g.t2;
    g.a_elems = g.n0p * 2 * g.Np;  // 16-bit elements, hi and lo tables each
    g.e1_elems = g.n1 * g.Np;      // complex float
    g.e2_elems = g.n2 * g.Np;
    return g;
}

// Blocks resident at once (256 CUs x 2 blocks at 2 waves per SIMD): by default
// grids with fewer tiles split the K steps over floor(512 / blocks) blocks (at
// least 8 steps = 128 modes each), so the launch stays one round with no tail,
// and add the partial sums afterwards.  The fill target is a tuning choice
// (GSX_MFMA_FILL_LOG2: 256 .. 2048 blocks).
constexpr int64_t kMFillBlocks = 512;
constexpr int kMFillLog2Min = 8, kMFillLog2Max = 11;

// (>= 1 also with no K steps or no tiles: those launch nothing)
inline int64_t mfma_steps_per_split(const MfmaGeom& g, int64_t comps, int64_t fill = kMFillBlocks) {
    const int64_t blocks = comps * g.row_tiles * g.col_tiles;
    const int64_t steps = g.Np / kMKM;
    if (steps < 1 || blocks < 1) return 1;
    int64_t k = blocks >= fill ? 1 : fill / blocks;
    if (k > steps / 8) k = steps / 8;
    if (k < 1) k = 1;
    return ceil_div(steps, k);
}

inline int64_t mfma_ksplit(const MfmaGeom& g, int64_t comps, int64_t fill = kMFillBlocks) {
    const int64_t steps = g.Np / kMKM;
    return steps < 1 ? 1 : ceil_div(steps, mfma_steps_per_split(g, comps, fill));
}

// A hi / lo (16-bit), E1 and one E2 table per output component (complex FP32)
inline int64_t mfma_table_bytes(const MfmaGeom& g, int64_t comps) {
    return 2 * g.a_elems * 2 + (g.e1_elems + comps * g.e2_elems) * 8;
}

// tables, then (K split) the FP32 partial sums, 256-byte aligned
inline int64_t mfma_ws_bytes(const MfmaGeom& g, int64_t comps = 1, int64_t fill = kMFillBlocks) {
    const int64_t ks = mfma_ksplit(g, comps, fill);
    const int64_t tables = ceil_div(mfma_table_bytes(g, comps), (int64_t)256) * 256;
    return ks > 1 ? tables + ks * comps * g.n0 * g.out_cols * 4 : mfma_table_bytes(g, comps);
}

// A uniformly spaced last axis of a 2D grid, y_j = y_0 + j h, factorises:
// e^{i k y_(16a+b)} = e^{i k y_(16a)} e^{i k (y_b - y_0)}.  The 2D sum is then
// the 3D sum over virtual axes (y_(16a): ceil(n/16) values, y_b - y_0: 16
// values) with the mode's k_1 on both, run by k_mfma_grid3: two small phase
// tables replace the n x N table of the axis, and the GEMM's B operand is
// built in registers.  Output column 16a + b of the virtual grid is column j
// of the real one (columns >= n are not written).
constexpr int64_t kSplitMin = 256;  // n / 16 >= 16: the 3D kernel's column tile

inline MfmaGeom mfma_split_geom(int64_t N, const int64_t* shape) {
    const int64_t vshape[3] = {shape[0], ceil_div(shape[1], (int64_t)16), 16};
    MfmaGeom g = mfma_geom(3, N, vshape);
    g.out_cols = shape[1];
    g.rs = shape[1];
    return g;
}

// split workspace: tables and partial sums, then the two virtual axes
inline int64_t mfma_split_ws_bytes(const MfmaGeom& gv, int64_t fill = kMFillBlocks) {
    return ceil_div(mfma_ws_bytes(gv, 1, fill), (int64_t)256) * 256 + (gv.n1 + 16) * 8;
}

// The GEMM's row axis is axis 0, in 128-row tiles.  A 3D grid with a short
// axis 0 (n0 = 32 fills a quarter of each tile) runs with axes 0 and 1
// swapped when that needs fewer tiles: rows = axis 1, column axes (axis 0,
// axis 2), written through the output strides (rs = n2, c1s = n1 n2) in the
// caller's C order; axis 2 stays the fastest, so stores stay coalesced.
// Tile cost: a k_mfma_grid tile (per-lane gathers from the E tables) takes
// about 2.2 k_mfma_grid3 tiles (512x512x8 against 512^3 per mode-point).
struct MfmaPlan {
    MfmaGeom g;
    bool split;    // 2D grid, uniform last axis factorised (k_split_axis)
    bool swapped;  // 3D grid run with axes 0 and 1 swapped
    int64_t fill;  // K-split fill target (resident blocks)
};

inline double mfma_tile_cost(const MfmaGeom& g) {
    return (double)g.row_tiles * (double)g.col_tiles * (g.g2 == 1 ? 1.0 : 2.2);
}

inline MfmaPlan mfma_plan(int64_t d, int64_t N, const int64_t* shape, int64_t comps, bool uniform_last,
                          int64_t fill = kMFillBlocks, bool allow_swap = true) {
    MfmaPlan p;
    p.split = uniform_last && d == 2 && comps == 1 && shape[1] >= kSplitMin;
    p.swapped = false;
    p.fill = fill;
    p.g = p.split ? mfma_split_geom(N, shape) : mfma_geom(d, N, shape);
    if (allow_swap && d == 3 && shape[0] < shape[1]) {
        const int64_t sw[3] = {shape[1], shape[0], shape[2]};
        MfmaGeom gs = mfma_geom(3, N, sw);
        if (mfma_tile_cost(gs) < 0.8 * mfma_tile_cost(p.g)) {
            gs.rs = shape[2];
            gs.c1s = shape[1] * shape[2];
            p.g = gs;
            p.swapped = true;
        }
    }
    return p;
}

inline int64_t mfma_plan_ws_bytes(const MfmaPlan& p, int64_t comps) {
    return p.split ? mfma_split_ws_bytes(p.g, p.fill) : mfma_ws_bytes(p.g, comps, p.fill);
}

// Workspace enough for every tuning choice (fill target, swap) of a grid.
inline int64_t mfma_ws_any(int64_t d, int64_t N, const int64_t* shape, int64_t comps, bool uniform_last) {
    int64_t ws = 0;
    for (int f = kMFillLog2Min; f <= kMFillLog2Max; ++f)
        for (int sw = 0; sw < 2; ++sw) {
            const int64_t w = mfma_plan_ws_bytes(mfma_plan(d, N, shape, comps, uniform_last, (int64_t)1 << f,
                                                           sw == 1),
                                                 comps);
            if (w > ws) ws = w;
        }
    return ws;
}

// ---------------------------------------------------------------- splits ---
template <bool FP16>
__device__ __forceinline__ void split2(float a, float b, uint32_t& hi, uint32_t& lo) {
    // hi = nearest 16-bit values of (a, b), lo = nearest 16-bit values of the
    // exact remainders.  The float value of hi is taken from the packed word
    // (v_perm_b32) so the compiler does not convert a and b a second time.
    if constexpr (FP16) {
        hi = __builtin_bit_cast(uint32_t, f16x2{(_Float16)a, (_Float16)b});
        const f16x2 h = __builtin_bit_cast(f16x2, __builtin_amdgcn_perm(0u, hi, 0x03020100u));
        lo = __builtin_bit_cast(uint32_t, f16x2{(_Float16)(a - (float)h[0]), (_Float16)(b - (float)h[1])});
    } else {
        hi = __builtin_bit_cast(uint32_t, bf16x2{(__bf16)a, (__bf16)b});
        const float ha = __builtin_bit_cast(float, __builtin_amdgcn_perm(0u, hi, 0x01000c0cu));
        const float hb = __builtin_bit_cast(float, hi & 0xffff0000u);
        lo = __builtin_bit_cast(uint32_t, bf16x2{(__bf16)(a - ha), (__bf16)(b - hb)});
    }
}

// --------------------------------------------------------------- tables ----
// One thread per table entry: A (n0p x Np), E1 (n1 x Np), E2 (comps x n2 x Np).
// comps > 1 (incompressible fields): E2 of component c carries the projector
// weight p_c(k_j) = delta_c0 - k_jc k_j0 / |k_j|^2 (generator.py:479-494) in
// place of the spectrum factor.
// ka_row, k1_row, kz_row: the rows of cov holding the wave-vector components of A, E1
// and E2 (0, 1, d - 1; 1, 1 for a split axis's two virtual axes; 1, 0, 2 with axes 0
// and 1 swapped).  The projector weight always uses cov's own rows.
template <bool FP16>
__global__ __launch_bounds__(256) void k_mfma_tables(MfmaGeom g, int64_t d, int64_t N, int64_t comps,
                                                     int64_t ka_row, int64_t k1_row, int64_t kz_row,
                                                     const double* __restrict__ ax0,
                                                     const double* __restrict__ ax1,
                                                     const double* __restrict__ ax2,
                                                     const double* __restrict__ cov,
                                                     const double* __restrict__ z1,
                                                     const double* __restrict__ z2,
                                                     const double* __restrict__ fac,
                                                     uint32_t* __restrict__ a_hi,
                                                     uint32_t* __restrict__ a_lo,
                                                     float2* __restrict__ e1,
                                                     float2* __restrict__ e2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t na = g.n0p * g.Np, n1 = g.e1_elems, n2 = comps * g.e2_elems;
    if (t < na) {
        const int64_t row = t / g.Np, j = t - row * g.Np;
        float re = 0.0f, im = 0.0f;
        if (row < g.n0 && j < N) {
            double s, c;
            const double th = cov[ka_row * N + j] * ax0[row];
            if (fabs(th) <= kFastPhaseMax) sincos_fast(th, &s, &c); else sincos(th, &s, &c);
            re = (float)c;
            im = (float)(-s);
        }
        uint32_t hi, lo;
        split2<FP16>(re, im, hi, lo);
        a_hi[t] = hi;  // element pair (2j, 2j+1) of row `row`
        a_lo[t] = lo;
    } else if (t < na + n1) {
        const int64_t u = t - na;
        const int64_t i1 = u / g.Np, j = u - i1 * g.Np;
        float2 v = {0.0f, 0.0f};
        if (j < N) {
            if (d == 3) {
                double s, c;
                const double th = cov[k1_row * N + j] * ax1[i1];
                if (fabs(th) <= kFastPhaseMax) sincos_fast(th, &s, &c); else sincos(th, &s, &c);
                v = {(float)c, (float)s};
            } else {
                v = {1.0f, 0.0f};
            }
        }
        e1[u] = v;
    } else if (t < na + n1 + n2) {
        const int64_t u = t - na - n1;
        const int64_t comp = u / g.e2_elems, uc = u - comp * g.e2_elems;
        const int64_t i2 = uc / g.Np, j = uc - i2 * g.Np;
        float2 v = {0.0f, 0.0f};
        if (j < N) {
            double s, c;
            const double kz = cov[kz_row * N + j];
            const double th = kz * (d == 3 ? ax2[i2] : ax1[i2]);
            if (fabs(th) <= kFastPhaseMax) sincos_fast(th, &s, &c); else sincos(th, &s, &c);
            double f = fac ? fac[j] : 1.0;
            if (comps > 1) {
                double k2 = 0.0;
                for (int64_t e = 0; e < d; ++e) k2 += cov[e * N + j] * cov[e * N + j];
                f = (comp == 0 ? 1.0 : 0.0) - cov[comp * N + j] * cov[j] / k2;
            }
            const double a = f * z1[j], b = f * z2[j];
            // (a - i b)(c + i s)
            v = {(float)(a * c + b * s), (float)(a * s - b * c)};
        }
        e2[u] = v;
    }
}

// ---------------------------------------------------------------- GEMM -----
template <bool FP16>
__device__ __forceinline__ f32x4 mfma(u32x4 a, u32x4 b, f32x4 c) {
    if constexpr (FP16) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                      __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    }
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return {a.x * b.x - a.y * b.y, __builtin_fmaf(a.x, b.y, a.y * b.x)};
}

// E1ONE: the grid has a one-point middle axis (2D grids), E1 == 1, so B is E2.
// SPLIT: the K steps are split over blocks (small grids), partial sums written.
template <bool FP16, bool E1ONE, bool SPLIT>
__global__ __launch_bounds__(kMWaves * 64, 2) void k_mfma_grid(MfmaGeom g, double alpha,
                                                               const uint32_t* __restrict__ a_hi,
                                                               const uint32_t* __restrict__ a_lo,
                                                               const float2* __restrict__ e1,
                                                               const float2* __restrict__ e2,
                                                               float* __restrict__ out, int64_t out_stride,
                                                               int64_t comps, int64_t steps_per,
                                                               float* __restrict__ partial) {
    // two stages [hi/lo][row * 16 + chunk * 4]: 64-byte A rows, 16-byte chunk c of row r
    // stored at c ^ ((r >> 1) & 3) (the 8 lanes of a 16-byte LDS access phase hit 8 bank
    // groups).  The stages are separate LDS objects and the step loop is unrolled by two,
    // so the compiler sees that a step's LDS reads and the next step's global -> LDS loads
    // touch different objects and does not wait for those loads before the reads.
    __shared__ __align__(16) uint32_t lds0[2][kMBM * 16];
    __shared__ __align__(16) uint32_t lds1[2][kMBM * 16];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int grp = lane >> 4;  // K group g
    const int l16 = lane & 15;

    // Block index = (K split, component, tile), tile fastest.  Component:
    // incompressible fields have one E2 table and output row per component.
    // K split (small grids): the split runs steps [split * steps_per, +steps_per)
    // and writes its unscaled partial sum, added in a fixed order afterwards.
    // Then the XCD-aware tile mapping: dispatch places block b on XCD b % 8.
    const int64_t ntile = g.row_tiles * g.col_tiles;
    const int64_t rest = (int64_t)blockIdx.x / ntile;
    const int64_t b = (int64_t)blockIdx.x - rest * ntile;
    const int64_t comp = rest % comps, split = SPLIT ? rest / comps : 0;
    e2 += comp * g.e2_elems;
    out = SPLIT ? partial + rest * out_stride : out + comp * out_stride;
    int64_t rt, ct;
    {
        const int64_t xcd = b & 7, w = b >> 3;
        const int64_t R = g.row_tiles;
        if (R <= 8 && (8 % R) == 0 && (g.col_tiles % (8 / R)) == 0) {
            const int64_t per = 8 / R;  // XCDs per row tile
            rt = xcd / per;
            ct = w * per + (xcd % per);
        } else {
            rt = b % R;
            ct = b / R;
        }
    }
    // Columns are flat (c = i1 * n2 + i2, 256 per tile: 16 per fragment), so a
    // short last axis (thin 3D slabs) wastes no lanes; the
    // clamped (i1, i2) of each fragment column index the tables.
    const float2* e1p[kMCF];
    const float2* e2p[kMCF];
#pragma unroll
    for (int cf = 0; cf < kMCF; ++cf) {
        const int64_t c = ct * 256 + 16 * (wave * kMCF + cf) + l16;
        const int64_t cc = c < g.ncols ? c : g.ncols - 1;
        const int64_t i1 = cc / g.n2, i2 = cc - i1 * g.n2;
        e1p[cf] = e1 + i1 * g.Np + 4 * grp;
        e2p[cf] = e2 + i2 * g.Np + 4 * grp;
    }

    f32x4 acc[kMRF][kMCF];
#pragma unroll
    for (int rf = 0; rf < kMRF; ++rf)
#pragma unroll
        for (int cf = 0; cf < kMCF; ++cf) acc[rf][cf] = f32x4{0.f, 0.f, 0.f, 0.f};

    // A staging: 128 rows x 64 B per K step for hi and for lo = 512 chunks of
    // 16 B each, straight from global memory into LDS (global_load_lds_dwordx4):
    // lane l of an instruction writes the wave's LDS base + 16 l, so the lane
    // filling slot q (row q >> 2) fetches the chunk the swizzle puts there.
    const int64_t row0 = rt * kMBM;
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    int64_t aoff[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int q = h * 256 + wv * 64 + lane;
        const int r = q >> 2;
        aoff[h] = (row0 + r) * g.Np + (((q & 3) ^ ((r >> 1) & 3)) * 4);
    }
    typedef uint32_t AStage[2][kMBM * 16];
    auto load_a = [&](int64_t step, AStage& S) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int slot = (h * 256 + wv * 64) * 4;
            __builtin_amdgcn_global_load_lds((const void*)(a_hi + aoff[h] + step * kMKM),
                                             (__attribute__((address_space(3))) void*)&S[0][slot], 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(a_lo + aoff[h] + step * kMKM),
                                             (__attribute__((address_space(3))) void*)&S[1][slot], 16, 0, 0);
        }
    };

    const int64_t steps = g.Np / kMKM;
    const int64_t s0 = SPLIT ? split * steps_per : 0;
    const int64_t s1 = SPLIT ? (s0 + steps_per < steps ? s0 + steps_per : steps) : steps;
    load_a(s0, lds0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();

    const int asw = (l16 >> 1) & 3;
    // one K step on stage C while stage Nx receives the next step's A tile
    auto step = [&](const AStage& C, AStage& Nx, int64_t s) {
        if (s + 1 < s1) load_a(s + 1, Nx);

        // B fragments for this K step, built in registers.
        u32x4 bh[kMCF], bl[kMCF];
        const int64_t moff = s * kMKM;
        float2 zf[4];
#pragma unroll
        for (int cf = 0; cf < kMCF; ++cf) {
            float2 yf[4];
            if constexpr (!E1ONE) {
                const float4* p = reinterpret_cast<const float4*>(e1p[cf] + moff);
                const float4 q0 = p[0], q1 = p[1];
                yf[0] = {q0.x, q0.y};
                yf[1] = {q0.z, q0.w};
                yf[2] = {q1.x, q1.y};
                yf[3] = {q1.z, q1.w};
            }
            {
                const float4* p = reinterpret_cast<const float4*>(e2p[cf] + moff);
                const float4 q0 = p[0], q1 = p[1];
                zf[0] = {q0.x, q0.y};
                zf[1] = {q0.z, q0.w};
                zf[2] = {q1.x, q1.y};
                zf[3] = {q1.z, q1.w};
            }
            uint32_t h4[4], l4[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float2 v = E1ONE ? zf[t] : cmul(yf[t], zf[t]);
                split2<FP16>(v.x, v.y, h4[t], l4[t]);
            }
            bh[cf] = u32x4{h4[0], h4[1], h4[2], h4[3]};
            bl[cf] = u32x4{l4[0], l4[1], l4[2], l4[3]};
        }

#pragma unroll
        for (int rf = 0; rf < kMRF; ++rf) {
            const int pos = (rf * 16 + l16) * 16 + ((grp ^ asw) * 4);
            const u32x4 ah = *reinterpret_cast<const u32x4*>(&C[0][pos]);
            const u32x4 al = *reinterpret_cast<const u32x4*>(&C[1][pos]);
#pragma unroll
            for (int cf = 0; cf < kMCF; ++cf) {
                acc[rf][cf] = mfma<FP16>(ah, bh[cf], acc[rf][cf]);
                acc[rf][cf] = mfma<FP16>(ah, bl[cf], acc[rf][cf]);
                acc[rf][cf] = mfma<FP16>(al, bh[cf], acc[rf][cf]);
            }
        }
        // the next step's LDS loads have landed (the barrier's own wait covers LDS, not them)
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    };
    for (int64_t s = s0; s < s1; s += 2) {
        step(lds0, lds1, s);
        if (s + 1 < s1) step(lds1, lds0, s + 1);
    }

    const float fa = SPLIT ? 1.0f : (float)alpha;  // partial sums are scaled when added
    // output offsets i1 * c1s + i2, computed after the loop (not held in registers through it)
    int64_t col[kMCF];
    bool cok[kMCF];
#pragma unroll
    for (int cf = 0; cf < kMCF; ++cf) {
        const int64_t c = ct * 256 + 16 * (wave * kMCF + cf) + l16;
        cok[cf] = c < g.ncols;
        const int64_t i1 = c / g.n2;
        col[cf] = i1 * g.c1s + (c - i1 * g.n2);
    }
#pragma unroll
    for (int rf = 0; rf < kMRF; ++rf) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t row = row0 + rf * 16 + grp * 4 + i;
            if (row < g.n0) {
#pragma unroll
                for (int cf = 0; cf < kMCF; ++cf)
                    if (cok[cf]) out[row * g.rs + col[cf]] = fa * acc[rf][cf][i];
            }
        }
    }
}

// 3D grids with n1, n2 >= 16: block tile = 128 rows x (16 i1 x 16 i2) columns.
// Each K step stages, through double-buffered LDS: the A tile (hi and lo,
// 128 rows x 16 element pairs) and the E1 / E2 slices of the block's 16 i1
// and 16 i2 values (16 modes each), loaded one step ahead by the whole block
// straight from global memory into LDS (global_load_lds_dwordx4: no staging
// registers, no LDS store instructions).  Layouts are XOR-swizzled so that the 8
// lanes of one 16-byte LDS access phase (32 banks x 4 B = 8 bank groups of
// 16 B) hit 8 distinct groups, for the reads (8 rows) and the stores:
//   A row r (64 B = 4 chunks, two rows per 128 B):  chunk c stored at c ^ ((r >> 1) & 3)
//   E row r (128 B = 8 chunks):                      chunk c stored at c ^ (r & 7)
struct alignas(16) Stage3 {
    uint32_t ah[kMBM * 16];
    uint32_t al[kMBM * 16];
    float e1[16 * 32];
    float e2[16 * 32];
};

template <bool FP16, bool SPLIT>
__global__ __launch_bounds__(kMWaves * 64, 2) void k_mfma_grid3(MfmaGeom g, double alpha,
                                                                const uint32_t* __restrict__ a_hi,
                                                                const uint32_t* __restrict__ a_lo,
                                                                const float2* __restrict__ e1,
                                                                const float2* __restrict__ e2,
                                                                float* __restrict__ out, int64_t out_stride,
                                                                int64_t comps, int64_t steps_per,
                                                                float* __restrict__ partial) {
    // two stages as separate LDS objects, the step loop unrolled by two: the compiler
    // then sees that a step's LDS reads and the next step's global -> LDS loads touch
    // different objects and does not wait for those loads before the reads
    __shared__ Stage3 st0;
    __shared__ Stage3 st1;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int grp = lane >> 4;
    const int l16 = lane & 15;

    // block index = (K split, component, tile), as in k_mfma_grid
    const int64_t ntile = g.row_tiles * g.col_tiles;
    const int64_t rest = (int64_t)blockIdx.x / ntile;
    const int64_t b = (int64_t)blockIdx.x - rest * ntile;
    const int64_t comp = rest % comps, split = SPLIT ? rest / comps : 0;
    e2 += comp * g.e2_elems;
    out = SPLIT ? partial + rest * out_stride : out + comp * out_stride;
    int64_t rt, ct;
    {
        const int64_t xcd = b & 7, w = b >> 3;
        const int64_t R = g.row_tiles;
        if (R <= 8 && (8 % R) == 0 && (g.col_tiles % (8 / R)) == 0) {
            const int64_t per = 8 / R;
            rt = xcd / per;
            ct = w * per + (xcd % per);
        } else {
            rt = b % R;
            ct = b / R;
        }
    }
    const int64_t tile1 = ct / g.t2, tile2 = ct - tile1 * g.t2;
    const int64_t i1b = tile1 * 16, i2b = tile2 * 16, row0 = rt * kMBM;

    // each K step, each thread fills two 16-byte LDS slots of the A tile (hi and lo: slots
    // tid and tid + 256, row q >> 2) and one of E: threads 0..127 -> E1, 128..255 -> E2
    // (row (tid & 127) >> 3)
    const int erow = (tid & 127) >> 3;
    const bool is_e1 = tid < 128;
    const int64_t eidx = is_e1 ? (i1b + erow < g.n1 ? i1b + erow : g.n1 - 1)
                               : (i2b + erow < g.n2 ? i2b + erow : g.n2 - 1);
    // global -> LDS loads: lane l of an instruction writes 16 B at the wave's LDS base + 16 l,
    // so each lane fetches the global chunk that the swizzled layout puts at its LDS slot
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    int64_t aoff[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int q = h * 256 + wv * 64 + lane;
        const int r = q >> 2;
        aoff[h] = (row0 + r) * g.Np + (((q & 3) ^ ((r >> 1) & 3)) * 4);
    }
    const float2* ebase = (is_e1 ? e1 : e2) + eidx * g.Np + (((tid & 7) ^ (erow & 7)) * 2);
    auto load = [&](int64_t step, Stage3& S) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int slot = (h * 256 + wv * 64) * 4;
            __builtin_amdgcn_global_load_lds((const void*)(a_hi + aoff[h] + step * kMKM),
                                             (__attribute__((address_space(3))) void*)&S.ah[slot], 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(a_lo + aoff[h] + step * kMKM),
                                             (__attribute__((address_space(3))) void*)&S.al[slot], 16, 0, 0);
        }
        float* es = (wv < 2 ? S.e1 : S.e2) + (wv & 1) * 256;
        __builtin_amdgcn_global_load_lds((const void*)(ebase + step * kMKM),
                                         (__attribute__((address_space(3))) void*)es, 16, 0, 0);
    };

    f32x4 acc[kMRF][kMCF];
#pragma unroll
    for (int rf = 0; rf < kMRF; ++rf)
#pragma unroll
        for (int cf = 0; cf < kMCF; ++cf) acc[rf][cf] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int64_t steps = g.Np / kMKM;
    const int64_t s0 = SPLIT ? split * steps_per : 0;
    const int64_t s1 = SPLIT ? (s0 + steps_per < steps ? s0 + steps_per : steps) : steps;
    load(s0, st0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    const int zsw = l16 & 7;
    // one K step on stage C while stage Nx receives the next step's tiles
    auto step = [&](const Stage3& C, Stage3& Nx, int64_t s) {
        if (s + 1 < s1) load(s + 1, Nx);
        // B fragments: E2 of this lane's i2 (row l16) times E1 of each fragment's i1
        float2 zf[4];
        {
            const float* zr = C.e2 + l16 * 32;
            const float4 q0 = *reinterpret_cast<const float4*>(zr + (((2 * grp) ^ zsw) * 4));
            const float4 q1 = *reinterpret_cast<const float4*>(zr + (((2 * grp + 1) ^ zsw) * 4));
            zf[0] = {q0.x, q0.y};
            zf[1] = {q0.z, q0.w};
            zf[2] = {q1.x, q1.y};
            zf[3] = {q1.z, q1.w};
        }
        u32x4 bh[kMCF], bl[kMCF];
#pragma unroll
        for (int cf = 0; cf < kMCF; ++cf) {
            const int er = wave * kMCF + cf;
            const int ysw = er & 7;
            const float* yr = C.e1 + er * 32;
            const float4 q0 = *reinterpret_cast<const float4*>(yr + (((2 * grp) ^ ysw) * 4));
            const float4 q1 = *reinterpret_cast<const float4*>(yr + (((2 * grp + 1) ^ ysw) * 4));
            const float2 yf[4] = {{q0.x, q0.y}, {q0.z, q0.w}, {q1.x, q1.y}, {q1.z, q1.w}};
            uint32_t h4[4], l4[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float2 v = cmul(yf[t], zf[t]);
                split2<FP16>(v.x, v.y, h4[t], l4[t]);
            }
            bh[cf] = u32x4{h4[0], h4[1], h4[2], h4[3]};
            bl[cf] = u32x4{l4[0], l4[1], l4[2], l4[3]};
        }

        const int asw = (l16 >> 1) & 3;
#pragma unroll
        for (int rf = 0; rf < kMRF; ++rf) {
            const int pos = (rf * 16 + l16) * 16 + ((grp ^ asw) * 4);
            const u32x4 ah = *reinterpret_cast<const u32x4*>(&C.ah[pos]);
            const u32x4 al = *reinterpret_cast<const u32x4*>(&C.al[pos]);
#pragma unroll
            for (int cf = 0; cf < kMCF; ++cf) {
                acc[rf][cf] = mfma<FP16>(ah, bh[cf], acc[rf][cf]);
                acc[rf][cf] = mfma<FP16>(ah, bl[cf], acc[rf][cf]);
                acc[rf][cf] = mfma<FP16>(al, bh[cf], acc[rf][cf]);
            }
        }
        // the next step's LDS loads have landed (the barrier's own wait covers LDS, not them)
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    };
    for (int64_t s = s0; s < s1; s += 2) {
        step(st0, st1, s);
        if (s + 1 < s1) step(st1, st0, s + 1);
    }

    const float fa = SPLIT ? 1.0f : (float)alpha;  // partial sums are scaled when added
    const int64_t i2 = i2b + l16;
#pragma unroll
    for (int cf = 0; cf < kMCF; ++cf) {
        const int64_t i1 = i1b + wave * kMCF + cf;
        if (i1 >= g.n1 || i2 >= g.n2) continue;
        if (i1 * g.n2 + i2 >= g.out_cols) continue;  // past the end of a split axis
        const int64_t col = i1 * g.c1s + i2;
#pragma unroll
        for (int rf = 0; rf < kMRF; ++rf) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t row = row0 + rf * 16 + grp * 4 + i;
                if (row < g.n0) out[row * g.rs + col] = fa * acc[rf][cf][i];
            }
        }
    }
}

// out[i] = alpha * sum_k partial[k * n + i], k in order (K-split partial sums)
__global__ void k_mfma_ksum(const float* __restrict__ partial, int64_t ksplit, int64_t n, float alpha,
                            float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f;
    for (int64_t k = 0; k < ksplit; ++k) acc += partial[k * n + i];
    out[i] = alpha * acc;
}

__global__ void k_fill_zero_f32(float* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// Virtual axes of a split last axis (one block): vy[a] = y[16a], vz[b] = y[b] - y[0].
// Checks y[j] == y[16 floor(j/16)] + (y[j mod 16] - y[0]) to 4 ulp of max|y| for every
// j (the same FP64 operations as the host's check); any miss (or a NaN) makes every vy
// NaN, so the whole output is NaN rather than a wrong field.
__global__ __launch_bounds__(256) void k_split_axis(const double* __restrict__ y, int64_t n,
                                                    double* __restrict__ vy, double* __restrict__ vz) {
    __shared__ double s_max[256];
    __shared__ int s_bad;
    const int t = threadIdx.x;
    double m = 0.0;
    for (int64_t j = t; j < n; j += 256) m = fmax(m, fabs(y[j]));
    s_max[t] = m;
    if (t == 0) s_bad = 0;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) s_max[t] = fmax(s_max[t], s_max[t + w]);
        __syncthreads();
    }
    const double tol = 0x1p-50 * s_max[0];
    const double y0 = y[0];
    bool bad = false;
    for (int64_t j = t; j < n; j += 256) {
        const double r = y[j & ~(int64_t)15] + (y[j & 15] - y0);
        if (!(fabs(y[j] - r) <= tol)) bad = true;
    }
    if (bad) s_bad = 1;
    __syncthreads();
    const double poison = s_bad ? __builtin_nan("") : 0.0;
    for (int64_t a = t; a < (n + 15) / 16; a += 256) vy[a] = y[16 * a] + poison;
    if (t < 16) vz[t] = y[t] - y0;
}

template <bool FP16>
int mfma_launch(const MfmaGeom& g, int64_t fill, int64_t d, int64_t N, int64_t comps, int64_t ka_row, int64_t k1_row,
                int64_t kz_row, const double* ax0,
                const double* ax1, const double* ax2, const double* cov, const double* z1, const double* z2,
                const double* fac, double alpha, float* out, void* ws, hipStream_t s) {
    uint32_t* a_hi = static_cast<uint32_t*>(ws);
    uint32_t* a_lo = a_hi + g.a_elems / 2;
    float2* e1 = reinterpret_cast<float2*>(a_lo + g.a_elems / 2);
    float2* e2 = e1 + g.e1_elems;
    const int64_t entries = g.n0p * g.Np + g.e1_elems + comps * g.e2_elems;
    hipLaunchKernelGGL((k_mfma_tables<FP16>), dim3((unsigned)ceil_div(entries, 256)), dim3(256), 0, s, g, d,
                       N, comps, ka_row, k1_row, kz_row, ax0, ax1, ax2, cov, z1, z2, fac, a_hi, a_lo, e1, e2);
    int st = launch_status();
    if (st != GSX_OK) return st;
    const int64_t steps_per = mfma_steps_per_split(g, comps, fill);
    const int64_t ksplit = mfma_ksplit(g, comps, fill);
    const int64_t blocks = ksplit * comps * g.row_tiles * g.col_tiles;
    const int64_t stride = g.n0 * g.out_cols;  // output values per component
    float* partial = nullptr;
    if (ksplit > 1)
        partial = reinterpret_cast<float*>(static_cast<char*>(ws) +
                                           ceil_div(mfma_table_bytes(g, comps), (int64_t)256) * 256);
    const dim3 grid((unsigned)blocks), block(kMWaves * 64);
    if (g.g2 == 1) {
        if (ksplit > 1)
            hipLaunchKernelGGL((k_mfma_grid3<FP16, true>), grid, block, 0, s, g, alpha, a_hi, a_lo, e1, e2, out,
                               stride, comps, steps_per, partial);
        else
            hipLaunchKernelGGL((k_mfma_grid3<FP16, false>), grid, block, 0, s, g, alpha, a_hi, a_lo, e1, e2, out,
                               stride, comps, steps_per, partial);
    } else if (d == 2) {
        if (ksplit > 1)
            hipLaunchKernelGGL((k_mfma_grid<FP16, true, true>), grid, block, 0, s, g, alpha, a_hi, a_lo, e1, e2,
                               out, stride, comps, steps_per, partial);
        else
            hipLaunchKernelGGL((k_mfma_grid<FP16, true, false>), grid, block, 0, s, g, alpha, a_hi, a_lo, e1, e2,
                               out, stride, comps, steps_per, partial);
    } else {
        if (ksplit > 1)
            hipLaunchKernelGGL((k_mfma_grid<FP16, false, true>), grid, block, 0, s, g, alpha, a_hi, a_lo, e1, e2,
                               out, stride, comps, steps_per, partial);
        else
            hipLaunchKernelGGL((k_mfma_grid<FP16, false, false>), grid, block, 0, s, g, alpha, a_hi, a_lo, e1,
                               e2, out, stride, comps, steps_per, partial);
    }
    st = launch_status();
    if (st != GSX_OK || ksplit == 1) return st;
    const int64_t n = comps * stride;
    hipLaunchKernelGGL(k_mfma_ksum, dim3((unsigned)ceil_div(n, (int64_t)256)), dim3(256), 0, s, partial, ksplit, n,
                       (float)alpha, out);
    return launch_status();
}

// comps: 1 (scalar sum) or d (incompressible: out holds d components of prod(shape) values)
int mfma_entry(int64_t d, int64_t N, const int64_t* shape, const double* axes, const double* cov,
               const double* z1, const double* z2, const double* fac, int variant, double alpha, float* out,
               void* workspace, int64_t workspace_bytes, void* stream, int64_t comps = 1) {
    if (d == 1) return GSX_E_UNSUPPORTED;
    if ((d != 2 && d != 3) || N < 0 || !shape) return GSX_E_INVALID;
    const bool uniform_last = (variant & GSX_MFMA_UNIFORM_LAST) != 0;
    const bool no_swap = (variant & GSX_MFMA_NO_SWAP) != 0;
    const int fill_log2 = (variant >> GSX_MFMA_FILL_SHIFT) & 0xf;
    variant &= ~(GSX_MFMA_UNIFORM_LAST | GSX_MFMA_NO_SWAP | (0xf << GSX_MFMA_FILL_SHIFT));
    if (variant != GSX_MFMA_BF16X3 && variant != GSX_MFMA_FP16X3) return GSX_E_INVALID;
    if (fill_log2 != 0 && (fill_log2 < kMFillLog2Min || fill_log2 > kMFillLog2Max)) return GSX_E_INVALID;
    const int64_t fill = fill_log2 ? ((int64_t)1 << fill_log2) : kMFillBlocks;
    int64_t P = 1;
    for (int i = 0; i < d; ++i) {
        if (shape[i] < 0) return GSX_E_INVALID;
        P *= shape[i];
    }
    if (P == 0) return GSX_OK;
    if (!axes || !out || (N > 0 && (!cov || !z1 || !z2))) return GSX_E_INVALID;
    hipStream_t s = as_stream(stream);
    if (N == 0) {  // empty sum
        hipLaunchKernelGGL(k_fill_zero_f32, dim3((unsigned)ceil_div(comps * P, 256)), dim3(256), 0, s, out,
                           comps * P);
        return launch_status();
    }
    const MfmaPlan plan = mfma_plan(d, N, shape, comps, uniform_last, fill, !no_swap);
    const MfmaGeom& g = plan.g;
    const bool split = plan.split;
    if (mfma_ksplit(g, comps, fill) * comps * g.row_tiles * g.col_tiles > 0x7fffffffLL ||
        ceil_div(comps * P, (int64_t)256) > 0x7fffffffLL)
        return GSX_E_INVALID;
    if (!workspace || workspace_bytes < mfma_plan_ws_bytes(plan, comps)) return GSX_E_WORKSPACE;
    const double* ax0 = axes;
    const double* ax1 = axes + shape[0];
    const double* ax2 = d == 3 ? ax1 + shape[1] : nullptr;
    int64_t dk = d, ka_row = 0, k1_row = 1, kz_row = d - 1;
    if (plan.swapped) {
        ax0 = axes + shape[0];
        ax1 = axes;
        ka_row = 1;
        k1_row = 0;
    }
    if (split) {
        double* vy = reinterpret_cast<double*>(static_cast<char*>(workspace) +
                                               ceil_div(mfma_ws_bytes(g, 1, fill), (int64_t)256) * 256);
        double* vz = vy + g.n1;
        hipLaunchKernelGGL(k_split_axis, dim3(1), dim3(256), 0, s, ax1, shape[1], vy, vz);
        const int st = launch_status();
        if (st != GSX_OK) return st;
        ax1 = vy;
        ax2 = vz;
        dk = 3;      // tables of the 3D form
        kz_row = 1;  // both virtual axes carry the mode's k_1
    }
    if (variant == GSX_MFMA_FP16X3)
        return mfma_launch<true>(g, fill, dk, N, comps, ka_row, k1_row, kz_row, ax0, ax1, ax2, cov, z1, z2, fac,
                                 alpha, out, workspace, s);
    return mfma_launch<false>(g, fill, dk, N, comps, ka_row, k1_row, kz_row, ax0, ax1, ax2, cov, z1, z2, fac,
                              alpha, out, workspace, s);
}

}  // namespace gsx

extern "C" {

int64_t gsx_summate_grid_mfma_workspace(int64_t d, int64_t N, const int64_t* shape) {
    if ((d != 2 && d != 3) || N < 0 || !shape) return GSX_E_INVALID;
    for (int i = 0; i < d; ++i)
        if (shape[i] < 0) return GSX_E_INVALID;
    // enough for GSX_MFMA_UNIFORM_LAST and every tuning choice as well
    const int64_t ws = gsx::mfma_ws_any(d, N, shape, 1, false);
    const int64_t ws_split = gsx::mfma_ws_any(d, N, shape, 1, true);
    return ws > ws_split ? ws : ws_split;
}

int gsx_summate_grid_mfma_plan(int64_t d, int64_t N, const int64_t* shape, int64_t comps, int variant,
                               int64_t* info) {
    if ((d != 2 && d != 3) || N < 0 || !shape || !info || (comps != 1 && comps != d)) return GSX_E_INVALID;
    for (int i = 0; i < d; ++i)
        if (shape[i] < 0) return GSX_E_INVALID;
    const int fill_log2 = (variant >> GSX_MFMA_FILL_SHIFT) & 0xf;
    if (fill_log2 != 0 && (fill_log2 < gsx::kMFillLog2Min || fill_log2 > gsx::kMFillLog2Max)) return GSX_E_INVALID;
    const int64_t fill = fill_log2 ? ((int64_t)1 << fill_log2) : gsx::kMFillBlocks;
    const bool uniform_last = (variant & GSX_MFMA_UNIFORM_LAST) != 0;
    const gsx::MfmaPlan p = gsx::mfma_plan(d, N, shape, comps, uniform_last, fill, (variant & GSX_MFMA_NO_SWAP) == 0);
    const int64_t ks = gsx::mfma_ksplit(p.g, comps, fill);
    info[0] = p.swapped;
    info[1] = p.split;
    info[2] = p.g.row_tiles;
    info[3] = p.g.col_tiles;
    info[4] = ks;
    info[5] = ks * comps * p.g.row_tiles * p.g.col_tiles;
    info[6] = gsx::mfma_plan_ws_bytes(p, comps);
    info[7] = p.g.g2 == 1 ? 3 : 2;  // 3: the 16 x 16 column-tile kernel, 2: flat columns
    return GSX_OK;
}

int gsx_summate_grid_mfma(int64_t d, int64_t N, const int64_t* shape, const double* axes, const double* cov,
                          const double* z1, const double* z2, int variant, double alpha, float* out,
                          void* workspace, int64_t workspace_bytes, void* stream) {
    return gsx::mfma_entry(d, N, shape, axes, cov, z1, z2, nullptr, variant, alpha, out, workspace,
                           workspace_bytes, stream);
}

int64_t gsx_summate_incompr_grid_mfma_workspace(int64_t d, int64_t N, const int64_t* shape) {
    if ((d != 2 && d != 3) || N < 0 || !shape) return GSX_E_INVALID;
    for (int i = 0; i < d; ++i)
        if (shape[i] < 0) return GSX_E_INVALID;
    return gsx::mfma_ws_any(d, N, shape, d, false);
}

int gsx_summate_incompr_grid_mfma(int64_t d, int64_t N, const int64_t* shape, const double* axes, const double* cov,
                                  const double* z1, const double* z2, int variant, double alpha, float* out,
                                  void* workspace, int64_t workspace_bytes, void* stream) {
    if (d == 1) return GSX_E_INVALID;  // incompressible fields are 2D or 3D
    return gsx::mfma_entry(d, N, shape, axes, cov, z1, z2, nullptr, variant, alpha, out, workspace,
                           workspace_bytes, stream, d);
}

int gsx_summate_fourier_grid_mfma(int64_t d, int64_t N, const int64_t* shape, const double* axes,
                                  const double* spectrum_factor, const double* modes, const double* z1,
                                  const double* z2, int variant, double alpha, float* out, void* workspace,
                                  int64_t workspace_bytes, void* stream) {
    if (d == 1) return GSX_E_UNSUPPORTED;
    if (N > 0 && !spectrum_factor) return GSX_E_INVALID;
    return gsx::mfma_entry(d, N, shape, axes, modes, z1, z2, spectrum_factor, variant, alpha, out, workspace,
                           workspace_bytes, stream);
}

}  // extern "C"
