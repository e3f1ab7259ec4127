// Empirical variograms on gfx950: pair binning (unstructured / directional)
// and axis lags (structured / masked structured).
//
// Reference contract: the external kernels called at
// src/gstools/variogram/variogram.py:44-113 (gstools-cython >= 1):
//   pairs j < k with edges[b] <= |x_j - x_k| < edges[b+1]; for directions a
//   pair also needs acos(|s| / dist) < angles_tol (s = (x_k - x_j) . dir,
//   distance-0 pairs pass) and, for bandwidth > 0, a band distance
//   |(x_k - x_j) - s dir| < bandwidth; with separate_dirs the first matching
//   direction takes the pair; every field m with neither value NaN adds
//   est(f[m,k] - f[m,j]) (x^2 or sqrt|x|) and one count.  The comparisons
//   are written exactly as there, so a NaN distance falls in every bin.
//
// Pair kernel: a persistent grid walks the upper triangle of 256 x 256 point
// tiles (tile (tj, tk), tk >= tj, round robin over blocks).  Thread t owns
// point j = 256 tj + t (coordinates and field values in registers); the k
// tile's coordinates and field values are staged in LDS and read as
// broadcasts.  Distances are FP64 with contraction off: the bits of the
// reference's distance, hence the same bin.  Bins: binary search when the
// edges are non-decreasing, a full scan otherwise.  Every thread accumulates
// into its own LDS histogram column ([slot][thread], conflict-free, no
// atomics); at the end each block sums its columns into per-block partials
// and a second kernel sums the blocks in a fixed order: counts exact, sums
// bitwise reproducible.  Slots are (direction, bin) pairs; a launch covers a
// range of slots so that the histograms fit the LDS.
//
// Structured kernel: lag k along the first axis of an (n0, n1) array; block =
// (lag, 256-column tile), each thread sums over i for its column, blocks
// write per-tile partials reduced in a fixed order.
#include <type_traits>

#include "gsx_internal.h"

namespace gsx {

constexpr int kVT = 256;        // points per tile (threads per block)
constexpr int kVPad = 8;        // zero doubles after the staged field rows
constexpr int kVFields = 8;     // fields per launch
constexpr int kVMaxDim = 3;
constexpr int kVBlocks = 2048;  // persistent blocks

struct VarioArgs {
    int64_t F, P, B, D;
    int cressie;
    int separate;
    int directional;
    double angles_tol, bandwidth;
    int64_t h0, h1;  // slot range of this launch: slot = d * B + b
    int64_t nt;      // tiles per side
};

constexpr double kDeg2Rad = 3.14159265358979323846 / 180.0;

// Great-circle angle (radians) of two (lat, lon) points in degrees, the haversine
// form of the backend's distance_type 'h' (selected at
// src/gstools/variogram/variogram.py:446 for latlon=True), in the same operation
// order as the oracle; coslat_* = cos(lat_* * pi / 180), taken once per point
// (the same bits as per pair).  Point j precedes point k.
__device__ __forceinline__ double haversine(double lat_j, double lon_j, double coslat_j, double lat_k, double lon_k,
                                            double coslat_k) {
#pragma clang fp contract(off)
    const double diff_lat = (lat_k - lat_j) * kDeg2Rad;
    const double diff_lon = (lon_k - lon_j) * kDeg2Rad;
    const double sl = sin(diff_lat / 2.0), so = sin(diff_lon / 2.0);
    const double arg = sl * sl + coslat_j * coslat_k * (so * so);
    return 2.0 * atan2(sqrt(arg), sqrt(1.0 - arg));
}

__device__ __forceinline__ double vario_est(int cressie, double x) { return cressie ? sqrt(fabs(x)) : x * x; }

// Largest b with e[b] <= x for non-decreasing e[0..B]; -1 if x < e[0].
__device__ __forceinline__ int64_t upper_edge(const double* e, int64_t B, double x) {
    int64_t lo = 0, hi = B + 1;  // answer in [lo - 1, hi - 1]
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (e[mid] <= x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - 1;
}

// T(e) = the smallest double x >= 0 with sqrt(x) >= e (sqrt correctly rounded,
// hence monotone): sqrt(d2) >= e exactly when d2 >= T(e), so bins can be found on
// squared distances with the reference's sqrt-then-compare decisions.  e <= 0
// gives -1 (every d2 >= 0 passes), e = +inf gives +inf, NaN stays NaN.
__device__ double sq_threshold(double e) {
    if (e != e) return e;
    if (e <= 0.0) return -1.0;
    double x = e * e;
    while (sqrt(x) < e) x = nextafter(x, __builtin_inf());
    for (;;) {
        const double y = nextafter(x, 0.0);
        if (y >= 0.0 && sqrt(y) >= e && y < x)
            x = y;
        else
            break;
    }
    return x;
}

// Q(tol) = the smallest double q in [0, 1) with acos(q) < tol (acos non-increasing):
// for q in [0, 1), acos(q) < tol exactly when q >= Q(tol); 1.0 when no q in [0, 1)
// qualifies (tol <= 0 or NaN), 0.0 when every one does.  Bisection over the bit
// patterns of [0, 1] with the same acos the per-pair test would call.
__device__ double acos_threshold(double tol) {
    if (!(acos(1.0) < tol)) return 1.0;  // tol <= 0 or NaN: no q qualifies
    if (acos(0.0) < tol) return 0.0;
    long long lo = __double_as_longlong(0.0), hi = __double_as_longlong(1.0);  // acos(lo) >= tol > acos(hi)
    while (hi - lo > 1) {
        const long long mid = lo + (hi - lo) / 2;
        if (acos(__longlong_as_double(mid)) < tol)
            hi = mid;
        else
            lo = mid;
    }
    return __longlong_as_double(hi);
}

// Smallest-index form of the bin search for non-decreasing edges: g in [-1, B]
// with e[g] <= x < e[g+1] (e[-1] = -inf, e[B+1] = +inf), starting from a guess
// (uniform edges: the guess is exact but for rounding, so the loops run 0-1 times).
__device__ __forceinline__ int bin_from_guess(const double* e, int B, double x, int g) {
    while (g >= 0 && e[g] > x) --g;
    while (g < B && e[g + 1] <= x) ++g;
    return g;
}

// One pair into slot idx of this thread's histogram row: LDS adds without a
// return value.  The row is the thread's own and adds to one location are
// ordered as the program issues them, so every slot adds its pairs in pair
// order (the same roundings as reading, adding and writing back each pair).
__device__ __forceinline__ void hist_add(double* h_sum, unsigned* h_cnt, int idx, double v) {
    __hip_atomic_fetch_add(&h_sum[idx], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&h_cnt[idx], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// FAST: one field, no directions; a block whose edges are non-decreasing runs a
// straight-line pair loop (the general loop covers everything else).
// CRESSIE selects the estimator of the FAST loop at compile time (a runtime flag
// would be if-converted into computing both estimators for every pair).
// HAV: (lat, lon) points in degrees, great-circle distances (general loop, DIM 2,
// a third staged row holds cos(lat)).
// DIRF: directional estimate of one field with non-decreasing edges: waves whose
// points and staged tile are all finite run a straight-line loop over four pairs
// per step and the directions in turn (others take the general loop).
template <int DIM, bool FAST, bool CRESSIE, bool HAV = false, bool DIRF = false>
__global__ __launch_bounds__(kVT) void k_vario_pairs(VarioArgs a, const double* __restrict__ f,
                                                     const double* __restrict__ edges,
                                                     const double* __restrict__ pos,
                                                     const double* __restrict__ dirs,
                                                     double* __restrict__ part_sum,
                                                     unsigned long long* __restrict__ part_cnt) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char vlds[];
    const int64_t NH = a.h1 - a.h0;
    const int t = threadIdx.x;
    double* s_edges = reinterpret_cast<double*>(vlds);
    // squared-distance thresholds of the edges, padded: s_thp[0] = -inf,
    // s_thp[i + 1] = T(e_i), s_thp[B + 2] = +inf, so that bins -1..B read
    // their two bounding thresholds at one address with no clamping
    double* s_thp = s_edges + (a.B + 1);
    double* s_thr = s_thp + 1;
    // directional tests on thresholds: acos(q) < angles_tol <=> q >= s_tol[0] (q in
    // [0, 1)), sqrt(bd) < bandwidth <=> bd < s_tol[1] (bd >= 0 or NaN)
    double* s_tol = s_thp + (a.B + 3);
    double* s_dirs = s_tol + 2;
    static_assert(!HAV || (DIM == 2 && !FAST), "haversine distances: 2D, general loop");
    static_assert(!DIRF || (!FAST && !HAV), "directional straight-line loop: its own variant");
    constexpr int NPOS = HAV ? DIM + 1 : DIM;  // staged coordinate rows
    // staged tile: coordinates then field values, rows of kVT, then kVPad zeros;
    // a read up to 7 entries past a row lands in the next row or in the pad
    double* s_pos = s_dirs + a.D * DIM;                     // [NPOS][kVT]
    double* s_f = s_pos + NPOS * kVT;                       // [F][kVT]
    double* h_sum = s_f + a.F * kVT + kVPad;                // [NH + 1][kVT], row NH scratch
    unsigned* h_cnt = reinterpret_cast<unsigned*>(h_sum + (NH + 1) * kVT);  // [NH + 1][kVT]

    for (int64_t i = t; i <= a.B; i += kVT) {
        const double e = edges[i];
        s_edges[i] = e;
        s_thr[i] = sq_threshold(e);
    }
    if (a.directional && t == kVT - 1) {
        s_tol[0] = acos_threshold(a.angles_tol);
        s_tol[1] = a.bandwidth > 0.0 ? sq_threshold(a.bandwidth) : 0.0;
    }
    if (t == 0) {
        s_thp[0] = -__builtin_inf();
        s_thp[a.B + 2] = __builtin_inf();
    }
    if (a.directional)
        for (int64_t i = t; i < a.D * DIM; i += kVT) s_dirs[i] = dirs[i];
    if (t < kVPad) s_f[a.F * kVT + t] = 0.0;
    for (int64_t h = 0; h <= NH; ++h) {
        h_sum[h * kVT + t] = 0.0;
        h_cnt[h * kVT + t] = 0u;
    }
    __syncthreads();
    bool mono = true;
    for (int64_t i = 0; i < a.B; ++i) mono = mono && (s_edges[i] <= s_edges[i + 1]);
    // near-uniform edges: bin guess from the spacing (exactness comes from the edge checks)
    const int Bi = (int)a.B, NHi = (int)NH, h0i = (int)a.h0;
    const double e0 = s_edges[0];
    const double w = Bi > 0 ? (s_edges[Bi] - e0) / Bi : 0.0;
    // uniform: every edge within w/4 of e0 + i w, and w not below 4e-6 of the edge
    // magnitudes (an FP32 distance then guesses the bin to within one); the FP32
    // guess also needs the last edge below 1e18 (larger squared distances overflow)
    const double emax = fmax(fabs(e0), fabs(s_edges[Bi > 0 ? Bi : 0]));
    // (w >= 1e-15: squared distances that underflow FP32 lie within 1e-19 of 0,
    // far below one bin)
    bool uniform = mono && Bi > 0 && w >= 1.0e-15 && emax < 1.0e18 && w >= 4.0e-6 * emax;
    for (int i = 0; uniform && i <= Bi; ++i) uniform = fabs(s_edges[i] - (e0 + i * w)) <= 0.25 * w;
    const double inv_w = uniform ? 1.0 / w : 0.0;
    // guess q = d / w - e0 / w in one FP32 fma
    const float inv_wf = (float)inv_w, qoff = (float)(-e0 * inv_w);
    // directions to test: all of those whose slots this launch records, and
    // with separate_dirs also the earlier ones (they may take the pair)
    const int64_t d_last = a.directional ? (a.h1 - 1) / a.B : 0;
    const int64_t d_first = a.directional ? (a.separate ? 0 : a.h0 / a.B) : 0;

    double nan_sum = 0.0;  // NaN-distance pairs of this thread (every bin takes them)
    unsigned nan_cnt = 0u;
    const int64_t ntp = a.nt * a.nt;
    for (int64_t tt = blockIdx.x; tt < ntp; tt += gridDim.x) {
        const int64_t tj = tt / a.nt, tk = tt - tj * a.nt;
        if (tk < tj) continue;
        const int64_t k0 = tk * kVT;
        const int64_t kn = (a.P - k0 < kVT) ? a.P - k0 : kVT;
        __syncthreads();  // previous tile's readers are done
        // a non-finite coordinate (inf - inf is a NaN distance) or first-field value
        // among the staged points
        bool stage_nan = false;
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            const double v = (t < kn) ? pos[c * a.P + k0 + t] : 0.0;
            stage_nan = stage_nan || !__builtin_isfinite(v);
            s_pos[c * kVT + t] = v;
        }
        if constexpr (HAV) s_pos[DIM * kVT + t] = cos(s_pos[t] * kDeg2Rad);
        for (int64_t m = 0; m < a.F; ++m) {
            const double v = (t < kn) ? f[m * a.P + k0 + t] : 0.0;
            stage_nan = stage_nan || (m == 0 && !__builtin_isfinite(v));
            s_f[m * kVT + t] = v;
        }
        const bool tile_nan = __syncthreads_or(stage_nan ? 1 : 0) != 0;
        const int64_t j = tj * kVT + t;
        if (j >= a.P) continue;
        double xj[DIM];
        bool xj_nan = false;
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            xj[c] = pos[c * a.P + j];
            xj_nan = xj_nan || !__builtin_isfinite(xj[c]);
        }
        double fj[kVFields];
#pragma unroll
        for (int m = 0; m < kVFields; ++m) fj[m] = (m < a.F) ? f[m * a.P + j] : 0.0;
        const int kstart = (tj == tk) ? t + 1 : 0;
        if constexpr (FAST) {
            if (mono) {
                const double fj0 = fj[0];
                const bool fj_ok = (a.F > 0) && (fj0 == fj0);
                const int kend = (int)kn;
                // Four pairs per step.  Every pair updates one histogram slot; pairs
                // that are not counted (out of range, outside this launch's slots,
                // NaN) go to the scratch row NH, so values and counts need no
                // selects.  The four slots are read together; a later pair in the
                // same slot as an earlier one adds onto that pair's new value
                // (the latest earlier match), and the writes go out in pair order,
                // so the last write to a slot holds every pair's contribution.
                // CLEAN: every coordinate and field value of this wave's points and
                // of the staged tile is finite, so no pair has a NaN distance or a
                // NaN field value.
                auto run = [&](auto uni, auto clean) {
                    constexpr bool UNI = decltype(uni)::value;
                    constexpr bool CLEAN = decltype(clean)::value;
                    // eight pairs per iteration: their bins are found together (the
                    // threshold reads overlap), then two updates of four slots each,
                    // in order (the second group's reads see the first group's writes)
                    for (int k8 = kstart; k8 < kend; k8 += 8) {
                        int slot[8];
                        double val[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            // past kend (at most 7 entries) the reads see staged values,
                            // zeros or the pad, all finite when CLEAN; such pairs go to
                            // the scratch row
                            const int kk = k8 + u;
                            const bool live = kk < kend;
                            double d2 = 0.0;
#pragma unroll
                            for (int c = 0; c < DIM; ++c) {
                                const double dd = s_pos[c * kVT + kk] - xj[c];
                                d2 += dd * dd;
                            }
                            const double fk = s_f[kk];
                            const double x = fk - fj0;
                            const double v = CRESSIE ? sqrt(fabs(x)) : x * x;
                            bool ok = live;
                            if constexpr (!CLEAN) {
                                ok = ok && fj_ok && (fk == fk);
                                const bool dnan = d2 != d2;  // in every bin by the reference's comparisons
                                nan_sum += (dnan && ok) ? v : 0.0;
                                nan_cnt += (dnan && ok) ? 1u : 0u;
                                ok = ok && !dnan;
                            }
                            int g;
                            if constexpr (UNI) {
                                // guess from an FP32 distance, within 0.1 bin of (d - e0) / w
                                // for the edges `uniform` admits (every edge within w/4 of
                                // uniform), so the truncated guess is within one bin of the
                                // true one and one threshold step is exact; thresholds are
                                // non-decreasing, so at most one of dec / inc holds
                                float q = __builtin_fmaf(__builtin_amdgcn_sqrtf((float)d2), inv_wf, qoff);
                                if constexpr (!CLEAN) q = (d2 == d2) ? q : -1.0f;
                                q = __builtin_amdgcn_fmed3f(q, -1.0f, (float)Bi);
                                const int g0 = (int)q;  // -1 .. B
                                const double tdown = s_thp[g0 + 1], tup = s_thp[g0 + 2];  // T(e_g0), T(e_g0+1)
                                g = g0 - (tdown > d2 ? 1 : 0) + (tup <= d2 ? 1 : 0);
                            } else {
                                g = (int)upper_edge(s_thr, a.B, d2);
                            }
                            // g is made opaque so the search stays unconditional: otherwise
                            // the compiler moves it, loads included, under a branch on `ok`
                            // and the four pairs' threshold reads no longer overlap
                            asm volatile("" : "+v"(g));
                            // launch slots lie in [0, B): one unsigned test covers the bin range
                            const unsigned sl = (unsigned)(g - h0i);
                            slot[u] = (ok && sl < (unsigned)NHi) ? (int)sl : NHi;
                            val[u] = v;
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u) hist_add(h_sum, h_cnt, slot[u] * kVT + t, val[u]);
                    }
                };
                const bool clean = !tile_nan && !__any((xj_nan || !__builtin_isfinite(fj0) || a.F < 1) ? 1 : 0);
                if (uniform) {
                    if (clean)
                        run(std::integral_constant<bool, true>{}, std::integral_constant<bool, true>{});
                    else
                        run(std::integral_constant<bool, true>{}, std::integral_constant<bool, false>{});
                } else {
                    if (clean)
                        run(std::integral_constant<bool, false>{}, std::integral_constant<bool, true>{});
                    else
                        run(std::integral_constant<bool, false>{}, std::integral_constant<bool, false>{});
                }
                continue;
            }
        }
        if constexpr (DIRF) {
            const double fj0 = fj[0];
            const bool clean = !tile_nan && !__any((xj_nan || !__builtin_isfinite(fj0) || a.F != 1) ? 1 : 0);
            if (mono && clean) {
                // Finite points: no NaN distance or value.  Per pair: the bin from
                // the squared distance and the edges' thresholds (exact, as in the
                // omnidirectional loop), dist = sqrt(d2) for the angle test; per
                // direction the band / angle tests on their thresholds, then one
                // histogram slot per pair updated as in the omnidirectional loop
                // (scratch row NH, carried values, writes in pair order).  With
                // separate directions the first matching direction takes the pair.
                const int kend = (int)kn;
                const bool band = a.bandwidth > 0.0;
                const double qtol = s_tol[0], btol = s_tol[1];
                for (int k4 = kstart; k4 < kend; k4 += 4) {
                    double dv[4][DIM], dist[4], val[4];
                    int g[4];
                    bool taken[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int kk = k4 + u;  // past kend: staged values or the pad
                        double d2 = 0.0;
#pragma unroll
                        for (int c = 0; c < DIM; ++c) {
                            dv[u][c] = s_pos[c * kVT + kk] - xj[c];
                            d2 += dv[u][c] * dv[u][c];
                        }
                        const double x = s_f[kk] - fj0;
                        val[u] = CRESSIE ? sqrt(fabs(x)) : x * x;
                        int gu;
                        if (uniform) {
                            float q = __builtin_fmaf(__builtin_amdgcn_sqrtf((float)d2), inv_wf, qoff);
                            q = __builtin_amdgcn_fmed3f(q, -1.0f, (float)Bi);
                            const int g0 = (int)q;
                            gu = g0 - (s_thp[g0 + 1] > d2 ? 1 : 0) + (s_thp[g0 + 2] <= d2 ? 1 : 0);
                        } else {
                            gu = (int)upper_edge(s_thr, a.B, d2);
                        }
                        g[u] = (kk < kend && (unsigned)gu < (unsigned)Bi) ? gu : -1;
                        dist[u] = sqrt(d2);
                        taken[u] = false;
                    }
                    for (int64_t d = d_first; d <= d_last; ++d) {
                        const double* ud = s_dirs + d * DIM;
                        int slot[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            double sp = 0.0;
#pragma unroll
                            for (int c = 0; c < DIM; ++c) sp += dv[u][c] * ud[c];
                            bool in = g[u] >= 0 && !taken[u];
                            if (band) {
                                double bd = 0.0;
#pragma unroll
                                for (int c = 0; c < DIM; ++c) {
                                    const double r = dv[u][c] - sp * ud[c];
                                    bd += r * r;
                                }
                                in = in && bd < btol;  // sqrt(bd) < bandwidth
                            }
                            if (dist[u] > 0.0) {
                                const double q = fabs(sp) / dist[u];
                                in = in && (!(q < 1.0) || q >= qtol);  // acos(q) < angles_tol
                            }
                            if (a.separate) taken[u] = taken[u] || in;
                            const unsigned sl = (unsigned)(d * a.B + g[u] - a.h0);
                            slot[u] = (in && sl < (unsigned)NHi) ? (int)sl : NHi;
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) hist_add(h_sum, h_cnt, slot[u] * kVT + t, val[u]);
                    }
                }
                continue;
            }
        }
        const double coslat_j = HAV ? cos(xj[0] * kDeg2Rad) : 0.0;
        for (int kk = kstart; kk < kn; ++kk) {
            double dv[DIM];
            double dist;
            if constexpr (HAV) {
                dist = haversine(xj[0], xj[1], coslat_j, s_pos[kk], s_pos[kVT + kk], s_pos[2 * kVT + kk]);
#pragma unroll
                for (int c = 0; c < DIM; ++c) dv[c] = 0.0;  // (directions are not allowed for lat-lon)
            } else {
                double d2 = 0.0;
#pragma unroll
                for (int c = 0; c < DIM; ++c) {
                    dv[c] = s_pos[c * kVT + kk] - xj[c];  // x_k - x_j
                    d2 += dv[c] * dv[c];
                }
                dist = sqrt(d2);
            }
            int64_t b_lo, b_hi;  // candidate bins [b_lo, b_hi)
            if (mono && dist == dist) {
                const int64_t b = upper_edge(s_edges, a.B, dist);
                if (b < 0 || b >= a.B) continue;
                b_lo = b;
                b_hi = b + 1;
            } else {
                b_lo = 0;
                b_hi = a.B;
            }
            for (int64_t b = b_lo; b < b_hi; ++b) {
                if (dist < s_edges[b] || dist >= s_edges[b + 1]) continue;
                for (int64_t d = d_first; d <= d_last; ++d) {
                    if (a.directional) {
                        const double* u = s_dirs + d * DIM;
                        double sp = 0.0;
#pragma unroll
                        for (int c = 0; c < DIM; ++c) sp += dv[c] * u[c];
                        bool in_band = true, in_angle = true;
                        if (a.bandwidth > 0.0) {
                            double bd = 0.0;
#pragma unroll
                            for (int c = 0; c < DIM; ++c) {
                                const double r = dv[c] - sp * u[c];
                                bd += r * r;
                            }
                            in_band = bd < s_tol[1];  // sqrt(bd) < bandwidth
                        }
                        if (dist > 0.0) {
                            const double q = fabs(sp) / dist;
                            if (q < 1.0) in_angle = q >= s_tol[0];  // acos(q) < angles_tol
                        }
                        if (!(in_band && in_angle)) continue;
                    }
                    const int64_t h = d * a.B + b - a.h0;
                    if (h >= 0 && h < NH) {
                        double s = 0.0;
                        unsigned c = 0;
#pragma unroll
                        for (int m = 0; m < kVFields; ++m) {
                            if (m >= a.F) break;
                            const double fk = s_f[m * kVT + kk];
                            if (fk != fk || fj[m] != fj[m]) continue;
                            ++c;
                            s += vario_est(a.cressie, fk - fj[m]);
                        }
                        h_sum[h * kVT + t] += s;
                        h_cnt[h * kVT + t] += c;
                    }
                    if (a.separate) break;
                }
            }
        }
    }
    if (nan_cnt) {
        for (int h = 0; h < NHi; ++h) {
            const int b = h + h0i;
            if (b < Bi) {
                h_sum[h * kVT + t] += nan_sum;
                h_cnt[h * kVT + t] += nan_cnt;
            }
        }
    }
    __syncthreads();
    // block partials: slot h summed over the threads in order
    for (int64_t h = t; h < NH; h += kVT) {
        double s = 0.0;
        unsigned long long c = 0;
        for (int i = 0; i < kVT; ++i) {
            s += h_sum[h * kVT + i];
            c += h_cnt[h * kVT + i];
        }
        part_sum[(int64_t)blockIdx.x * NH + h] = s;
        part_cnt[(int64_t)blockIdx.x * NH + h] = c;
    }
}

// Sum the block partials of slot h in block order; accumulate into the totals.
__global__ void k_vario_collect(int64_t blocks, int64_t NH, int64_t h0, const double* __restrict__ part_sum,
                                const unsigned long long* __restrict__ part_cnt, double* __restrict__ tot_sum,
                                long long* __restrict__ tot_cnt, int accumulate) {
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= NH) return;
    double s = 0.0;
    unsigned long long c = 0;
    for (int64_t b = 0; b < blocks; ++b) {
        s += part_sum[b * NH + h];
        c += part_cnt[b * NH + h];
    }
    if (accumulate) {
        tot_sum[h0 + h] += s;
        tot_cnt[h0 + h] += (long long)c;
    } else {
        tot_sum[h0 + h] = s;
        tot_cnt[h0 + h] = (long long)c;
    }
}

// gamma from sums and counts (Matheron or Cressie), N = max(count, 1)
__global__ void k_vario_normalize(int64_t n, int cressie, double* __restrict__ gamma,
                                  const long long* __restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double cnt = (double)(counts[i] > 0 ? counts[i] : 1);
    const double s = gamma[i];
    if (cressie) {
        const double q = s / cnt;
        gamma[i] = 0.5 * (q * q * (q * q)) / (0.457 + 0.494 / cnt + 0.045 / (cnt * cnt));
    } else {
        gamma[i] = s / (2.0 * cnt);
    }
}

// structured lags: block b takes lag k = b / tiles and column tile b % tiles.
// A tile holds C = 2^cols_log2 columns (256, or the next power of two >= n1
// for narrow fields); the block's 256 / C row phases split the rows i, so a
// 1D series (n1 = 1) spreads each lag over all 256 lanes.  Lanes of one row
// phase read consecutive columns (coalesced), and for C = 1 consecutive rows.
// The block's partial sum is reduced in a fixed order.
__global__ __launch_bounds__(256) void k_vario_lags(int64_t n0, int64_t n1, int64_t tiles, int cols_log2,
                                                    const double* __restrict__ f,
                                                    const unsigned char* __restrict__ mask, int cressie,
                                                    double* __restrict__ part_sum,
                                                    unsigned long long* __restrict__ part_cnt) {
    __shared__ double rs[256];
    __shared__ unsigned long long rc[256];
    const int64_t b = blockIdx.x;
    const int64_t k = b / tiles;
    const int64_t tile = b - k * tiles;
    const int c = (int)threadIdx.x & ((1 << cols_log2) - 1);
    const int r = (int)threadIdx.x >> cols_log2;
    const int phases = 256 >> cols_log2;
    const int64_t j = (tile << cols_log2) + c;
    double s = 0.0;
    unsigned long long cnt = 0;
    if (k > 0 && j < n1) {
        for (int64_t i = r; i + k < n0; i += phases) {
            if (mask && (mask[(i + k) * n1 + j] || mask[i * n1 + j])) continue;
            ++cnt;
            s += vario_est(cressie, f[(i + k) * n1 + j] - f[i * n1 + j]);
        }
    }
    rs[threadIdx.x] = s;
    rc[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            rs[threadIdx.x] += rs[threadIdx.x + w];
            rc[threadIdx.x] += rc[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_sum[k * tiles + tile] = rs[0];
        part_cnt[k * tiles + tile] = rc[0];
    }
}

// Wide arrays (n1 >= 64): block = 64 columns x 4 KW lags over one chunk of rows;
// wave w takes lags kb = k0 + KW w .. kb + KW - 1 of the block's group, lane c
// column 64 t + c.  A thread walks its column's rows KW at a time with the KW
// values f[i .. i + KW - 1] and the 2 KW - 1 values f[i + kb .. i + kb + 2 KW - 2]
// in registers (3 KW - 1 loads serve KW^2 pairs, each pair the same difference
// as in k_vario_lags; the rows are addressed from wave-uniform row pointers with
// 32-bit lane offsets).  KW = 16 without a mask, 8 with one (its flags take
// registers too).  Sums run over rows in order per thread, then over the 64
// columns (fixed shuffle order), then over the blocks
// (k_vario_lag_collect_wave).  Unmasked counts are closed-form: (rows with a
// pair) x (columns) per block.
constexpr int kLagWMax = 16;

template <int kLagW, bool MASK, bool CRESSIE>
__global__ __launch_bounds__(256) void k_vario_lags_win(int64_t n0, int64_t n1, int64_t ctiles, int64_t rchunks,
                                                        int64_t rows_per, const double* __restrict__ f,
                                                        const unsigned char* __restrict__ mask,
                                                        double* __restrict__ part_sum,
                                                        unsigned long long* __restrict__ part_cnt) {
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    const int64_t b = blockIdx.x;
    const int64_t rc = b % rchunks, rest = b / rchunks;
    const int64_t ct = rest % ctiles, lg = rest / ctiles;
    const int64_t kb = lg * 4 * kLagW + __builtin_amdgcn_readfirstlane(wave) * kLagW;  // this wave's first lag
    const int64_t j = ct * 64 + lane;
    const bool col_ok = j < n1;
    const int jc = (int)(col_ok ? j : n1 - 1);  // loads of lanes past the last column stay in bounds
    const int n1i = (int)n1;
    const int64_t i_lo = rc * rows_per;
    const int64_t i_hi = i_lo + rows_per < n0 ? i_lo + rows_per : n0;
    const int64_t i_end = i_hi < n0 - kb ? i_hi : n0 - kb;  // rows with a pair at lag kb
    double sum[kLagW];
    unsigned cnt[kLagW];
#pragma unroll
    for (int t = 0; t < kLagW; ++t) {
        sum[t] = 0.0;
        cnt[t] = 0u;
    }
    for (int64_t i0 = i_lo; i0 < i_end; i0 += kLagW) {
        double a[kLagW], w[2 * kLagW - 1];
        bool ma[kLagW], mw[2 * kLagW - 1];
        // rows i0 .. and i0 + kb .. (clamped at the last row: those pairs are skipped)
        const int64_t ra = i0, rw = i0 + kb;
        const int na = (int)(n0 - ra < kLagW ? n0 - ra : kLagW);                // rows of a in the array
        const int nw = (int)(n0 - rw < 2 * kLagW - 1 ? n0 - rw : 2 * kLagW - 1);  // rows of w in the array
        const double* fa = f + ra * n1;
        const double* fw = f + rw * n1;
#pragma unroll
        for (int u = 0; u < kLagW; ++u) {
            const int o = (u < na ? u : na - 1) * n1i + jc;
            a[u] = fa[o];
            if constexpr (MASK) ma[u] = mask[ra * n1 + o] != 0;
        }
#pragma unroll
        for (int v = 0; v < 2 * kLagW - 1; ++v) {
            const int o = (v < nw ? v : nw - 1) * n1i + jc;
            w[v] = fw[o];
            if constexpr (MASK) mw[v] = mask[rw * n1 + o] != 0;
        }
        // every pair (i0 + u, kb + t) exists: rows in the chunk, i + k < n0, k > 0
        const bool full = kb > 0 && i0 + kLagW <= i_hi && nw == 2 * kLagW - 1;
#pragma unroll
        for (int u = 0; u < kLagW; ++u)
#pragma unroll
            for (int t = 0; t < kLagW; ++t) {
                const double d = w[u + t] - a[u];
                const double e = CRESSIE ? sqrt(fabs(d)) : d * d;
                bool ok = full || (i0 + u < i_hi && u + t < nw && kb + t > 0);
                if constexpr (MASK) ok = ok && !ma[u] && !mw[u + t];
                if (MASK || !full) {
                    sum[t] += ok ? e : 0.0;
                    cnt[t] += ok ? 1u : 0u;
                } else {
                    sum[t] += e;
                }
            }
    }
    const int64_t cols = n1 - ct * 64 < 64 ? n1 - ct * 64 : 64;
    const int64_t T = ctiles * rchunks;
#pragma unroll
    for (int t = 0; t < kLagW; ++t) {
        double v = col_ok ? sum[t] : 0.0;
        unsigned c = col_ok ? cnt[t] : 0u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            v += __shfl_xor(v, o);
            if constexpr (MASK) c += __shfl_xor(c, o);
        }
        const int64_t k = kb + t;
        if (lane == 0 && k < n0) {
            part_sum[k * T + ct * rchunks + rc] = v;
            if constexpr (MASK) {
                part_cnt[k * T + ct * rchunks + rc] = c;
            } else {
                const int64_t top = i_hi < n0 - k ? i_hi : n0 - k;
                const int64_t rows = (k > 0 && top > i_lo) ? top - i_lo : 0;
                part_cnt[k * T + ct * rchunks + rc] = (unsigned long long)(rows * cols);
            }
        }
    }
}

// k_vario_lag_collect for many partials per lag: one wave per lag, lane l adding
// partials l, l + 64, ... in order, then the 64 lanes in a fixed shuffle order
__global__ __launch_bounds__(256) void k_vario_lag_collect_wave(int64_t n0, int64_t tiles,
                                                                const double* __restrict__ part_sum,
                                                                const unsigned long long* __restrict__ part_cnt,
                                                                double* __restrict__ gamma,
                                                                long long* __restrict__ counts) {
    const int lane = (int)threadIdx.x & 63;
    const int64_t k = (int64_t)blockIdx.x * 4 + ((int)threadIdx.x >> 6);
    if (k >= n0) return;  // whole wave
    double s = 0.0;
    unsigned long long c = 0;
    for (int64_t b = lane; b < tiles; b += 64) {
        s += part_sum[k * tiles + b];
        c += part_cnt[k * tiles + b];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o);
        c += __shfl_xor(c, o);
    }
    if (lane == 0) {
        gamma[k] = s;
        counts[k] = (long long)c;
    }
}

// row chunks of the wide-array lag kernel: enough blocks to fill the chip
// (from 64-lag groups for either lag width, so that both share one workspace size)
inline int64_t lag_win_rchunks(int64_t n0, int64_t n1) {
    const int64_t groups = ceil_div(n0, (int64_t)64) * ceil_div(n1, (int64_t)64);
    int64_t rc = ceil_div((int64_t)4096, groups);
    const int64_t max_rc = ceil_div(n0, (int64_t)256);  // at least 256 rows per chunk
    if (rc > max_rc) rc = max_rc;
    if (rc > 64) rc = 64;
    return rc < 1 ? 1 : rc;
}

inline int64_t lag_win_rows(int64_t n0, int64_t rchunks) {
    return ceil_div(ceil_div(n0, rchunks), (int64_t)kLagWMax) * kLagWMax;
}

__global__ void k_vario_lag_collect(int64_t n0, int64_t tiles, const double* __restrict__ part_sum,
                                    const unsigned long long* __restrict__ part_cnt, double* __restrict__ gamma,
                                    long long* __restrict__ counts) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n0) return;
    double s = 0.0;
    unsigned long long c = 0;
    for (int64_t b = 0; b < tiles; ++b) {
        s += part_sum[k * tiles + b];
        c += part_cnt[k * tiles + b];
    }
    gamma[k] = s;
    counts[k] = (long long)c;
}

// LDS bytes of a pair launch with NH histogram slots (+ one scratch row)
inline size_t vario_lds_bytes(int64_t B, int64_t D, int dim, int64_t F, int64_t NH) {
    return sizeof(double) * (2 * (size_t)(B + 1) + 4 + (size_t)(D * dim) + (size_t)dim * kVT + (size_t)F * kVT + kVPad +
                             (size_t)(NH + 1) * kVT) +
           sizeof(unsigned) * (size_t)(NH + 1) * kVT;
}

int vario_lds_limit(size_t* out) {
    static size_t cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return GSX_E_NO_DEVICE;
    if (!cached[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            return GSX_E_NO_DEVICE;
        cached[dev] = (size_t)v;
    }
    *out = cached[dev];
    return GSX_OK;
}

// histogram slots one launch can hold, for fc fields per launch
inline int64_t vario_slots_per_launch(size_t limit, int64_t B, int64_t D, int dim, int64_t fc) {
    const size_t fixed = vario_lds_bytes(B, D, dim, fc, 0);
    if (limit <= fixed) return 0;
    return (int64_t)((limit - fixed) / ((sizeof(double) + sizeof(unsigned)) * kVT));
}

template <int DIM, bool FAST, bool CRESSIE, bool HAV = false, bool DIRF = false>
int launch_pairs_t(const VarioArgs& a, size_t lds, int64_t blocks, hipStream_t s, const double* f,
                   const double* edges, const double* pos, const double* dirs, double* part_sum,
                   unsigned long long* part_cnt) {
    if (lds > 65536 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_vario_pairs<DIM, FAST, CRESSIE, HAV, DIRF>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return GSX_E_LAUNCH;
    hipLaunchKernelGGL((k_vario_pairs<DIM, FAST, CRESSIE, HAV, DIRF>), dim3((unsigned)blocks), dim3(kVT), lds, s, a,
                       f, edges, pos, dirs, part_sum, part_cnt);
    return launch_status();
}

template <int DIM>
int launch_pairs(const VarioArgs& a, size_t lds, int64_t blocks, hipStream_t s, const double* f,
                 const double* edges, const double* pos, const double* dirs, double* part_sum,
                 unsigned long long* part_cnt) {
    if (!a.directional && a.F <= 1) {
        if (a.cressie)
            return launch_pairs_t<DIM, true, true>(a, lds, blocks, s, f, edges, pos, dirs, part_sum, part_cnt);
        return launch_pairs_t<DIM, true, false>(a, lds, blocks, s, f, edges, pos, dirs, part_sum, part_cnt);
    }
    if (a.directional && a.F == 1) {
        if (a.cressie)
            return launch_pairs_t<DIM, false, true, false, true>(a, lds, blocks, s, f, edges, pos, dirs, part_sum,
                                                                 part_cnt);
        return launch_pairs_t<DIM, false, false, false, true>(a, lds, blocks, s, f, edges, pos, dirs, part_sum,
                                                              part_cnt);
    }
    return launch_pairs_t<DIM, false, false>(a, lds, blocks, s, f, edges, pos, dirs, part_sum, part_cnt);
}

inline int64_t vario_pair_blocks(int64_t P) {
    const int64_t nt = ceil_div(P, (int64_t)kVT);
    const int64_t ntri = nt * (nt + 1) / 2;
    return ntri < kVBlocks ? (ntri > 0 ? ntri : 1) : kVBlocks;
}

int vario_pairs(int64_t F, int64_t P, int64_t dim, int64_t B, int64_t D, const double* f, const double* edges,
                const double* pos, const double* dirs, double angles_tol, double bandwidth, int separate,
                int estimator, int directional, double* gamma, int64_t* counts, void* ws, int64_t ws_bytes,
                void* stream, bool haversine = false) {
    if (F < 0 || P < 0 || dim < 1 || dim > kVMaxDim || B < 0 || D < 1 || (estimator != 0 && estimator != 1))
        return GSX_E_INVALID;
    if (haversine && (dim != 2 || directional)) return GSX_E_INVALID;
    const int dim_stage = haversine ? 3 : (int)dim;  // staged coordinate rows
    const int64_t NHall = D * B;
    if (NHall == 0) return GSX_OK;
    if (B > 0x3fffffff || NHall > 0x3fffffff) return GSX_E_INVALID;
    if (!gamma || !counts || !edges || (directional && !dirs) || (P > 0 && (!pos || (F > 0 && !f))))
        return GSX_E_INVALID;
    size_t limit = 0;
    int st = vario_lds_limit(&limit);
    if (st != GSX_OK) return st;
    const int64_t fc_max = F < kVFields ? (F > 0 ? F : 1) : kVFields;  // fields staged per launch
    const int64_t per = vario_slots_per_launch(limit, B, directional ? D : 1, dim_stage, fc_max);
    if (per < 1) return GSX_E_UNSUPPORTED;
    const int64_t blocks = vario_pair_blocks(P);
    const int64_t chunk_slots = per < NHall ? per : NHall;
    const int64_t need = blocks * chunk_slots * (int64_t)(sizeof(double) + sizeof(unsigned long long));
    if (!ws || ws_bytes < need) return GSX_E_WORKSPACE;
    // per-thread counts are 32-bit: bound the pairs one thread can see
    const int64_t nt = ceil_div(P, (int64_t)kVT);
    const int64_t tiles_per_block = ceil_div(nt * nt, blocks);
    if ((double)tiles_per_block * kVT * (double)(F > 0 ? F : 1) > 4.0e9) return GSX_E_UNSUPPORTED;
    hipStream_t s = as_stream(stream);
    double* part_sum = static_cast<double*>(ws);
    unsigned long long* part_cnt = reinterpret_cast<unsigned long long*>(part_sum + blocks * chunk_slots);
    long long* tot_cnt = reinterpret_cast<long long*>(counts);
    for (int64_t h0 = 0; h0 < NHall; h0 += chunk_slots) {
        const int64_t h1 = (h0 + chunk_slots < NHall) ? h0 + chunk_slots : NHall;
        for (int64_t m0 = 0; m0 < (F > 0 ? F : 1); m0 += kVFields) {
            const int64_t fc = F - m0 < kVFields ? F - m0 : kVFields;
            VarioArgs a;
            a.F = F > 0 ? fc : 0;
            a.P = P;
            a.B = B;
            a.D = directional ? D : 1;
            a.cressie = estimator;
            a.separate = separate;
            a.directional = directional;
            a.angles_tol = angles_tol;
            a.bandwidth = bandwidth;
            a.h0 = h0;
            a.h1 = h1;
            a.nt = nt;
            const size_t lds = vario_lds_bytes(B, a.D, dim_stage, fc_max, h1 - h0);
            if (P > 1) {
                const double* fm = F > 0 ? f + m0 * P : nullptr;
                if (haversine)
                    st = launch_pairs_t<2, false, false, true>(a, lds, blocks, s, fm, edges, pos, dirs, part_sum,
                                                               part_cnt);
                else if (dim == 1)
                    st = launch_pairs<1>(a, lds, blocks, s, fm, edges, pos, dirs, part_sum, part_cnt);
                else if (dim == 2)
                    st = launch_pairs<2>(a, lds, blocks, s, fm, edges, pos, dirs, part_sum, part_cnt);
                else
                    st = launch_pairs<3>(a, lds, blocks, s, fm, edges, pos, dirs, part_sum, part_cnt);
                if (st != GSX_OK) return st;
                hipLaunchKernelGGL(k_vario_collect, dim3((unsigned)ceil_div(h1 - h0, (int64_t)256)), dim3(256), 0,
                                   s, blocks, h1 - h0, h0, part_sum, part_cnt, gamma, tot_cnt, m0 > 0 ? 1 : 0);
            } else {
                hipLaunchKernelGGL(k_vario_collect, dim3((unsigned)ceil_div(h1 - h0, (int64_t)256)), dim3(256), 0,
                                   s, (int64_t)0, h1 - h0, h0, part_sum, part_cnt, gamma, tot_cnt, 0);
            }
            st = launch_status();
            if (st != GSX_OK) return st;
        }
    }
    hipLaunchKernelGGL(k_vario_normalize, dim3((unsigned)ceil_div(NHall, (int64_t)256)), dim3(256), 0, s, NHall,
                       estimator, gamma, tot_cnt);
    return launch_status();
}

// structured lags of an (n0, n1 >= 64) array with the register-window kernel
inline int lags_wide(int64_t n0, int64_t n1, const double* f, const uint8_t* mask, int estimator, double* gamma,
                     int64_t* counts, void* workspace, int64_t workspace_bytes, void* stream) {
    const int64_t ctiles = ceil_div(n1, (int64_t)64), rchunks = lag_win_rchunks(n0, n1);
    const int64_t rows_per = lag_win_rows(n0, rchunks);
    const int64_t T = ctiles * rchunks;
    const int kw = mask ? 8 : 16;  // lags per wave
    const int64_t blocks = ceil_div(n0, (int64_t)(4 * kw)) * T;
    // 32-bit lane offsets within 2 kw rows
    if (blocks > 0x7fffffffLL || n1 > 0x7fffffffLL / (2 * kLagWMax)) return GSX_E_UNSUPPORTED;
    if (!workspace || workspace_bytes < n0 * T * 16) return GSX_E_WORKSPACE;
    hipStream_t s = as_stream(stream);
    double* part_sum = static_cast<double*>(workspace);
    unsigned long long* part_cnt = reinterpret_cast<unsigned long long*>(part_sum + n0 * T);
    const unsigned char* m = mask;
    const dim3 grid((unsigned)blocks), block(256);
    if (mask) {
        if (estimator)
            hipLaunchKernelGGL((k_vario_lags_win<8, true, true>), grid, block, 0, s, n0, n1, ctiles, rchunks, rows_per,
                               f, m, part_sum, part_cnt);
        else
            hipLaunchKernelGGL((k_vario_lags_win<8, true, false>), grid, block, 0, s, n0, n1, ctiles, rchunks, rows_per,
                               f, m, part_sum, part_cnt);
    } else {
        if (estimator)
            hipLaunchKernelGGL((k_vario_lags_win<16, false, true>), grid, block, 0, s, n0, n1, ctiles, rchunks,
                               rows_per, f, m, part_sum, part_cnt);
        else
            hipLaunchKernelGGL((k_vario_lags_win<16, false, false>), grid, block, 0, s, n0, n1, ctiles, rchunks,
                               rows_per, f, m, part_sum, part_cnt);
    }
    int st = launch_status();
    if (st != GSX_OK) return st;
    long long* cnt = reinterpret_cast<long long*>(counts);
    hipLaunchKernelGGL(k_vario_lag_collect_wave, dim3((unsigned)ceil_div(n0, (int64_t)4)), dim3(256), 0, s, n0, T,
                       part_sum, part_cnt, gamma, cnt);
    st = launch_status();
    if (st != GSX_OK) return st;
    hipLaunchKernelGGL(k_vario_normalize, dim3((unsigned)ceil_div(n0, (int64_t)256)), dim3(256), 0, s, n0, estimator,
                       gamma, cnt);
    return launch_status();
}

}  // namespace gsx

extern "C" {

int64_t gsx_vario_workspace(int64_t kind, int64_t P_or_n0, int64_t dim_or_n1, int64_t B, int64_t D) {
    if (kind == 2) {  // structured: n0 lags x column tiles (x row chunks for wide arrays)
        if (P_or_n0 < 0 || dim_or_n1 < 0) return GSX_E_INVALID;
        const int64_t tiles = gsx::ceil_div(dim_or_n1, (int64_t)256);
        if (dim_or_n1 >= 64)
            return P_or_n0 * gsx::ceil_div(dim_or_n1, (int64_t)64) * gsx::lag_win_rchunks(P_or_n0, dim_or_n1) * 16;
        return P_or_n0 * (tiles > 0 ? tiles : 1) * 16;
    }
    if (P_or_n0 < 0 || dim_or_n1 < 1 || dim_or_n1 > gsx::kVMaxDim || B < 0 || D < 1) return GSX_E_INVALID;
    // per-block partials of every (direction, bin) slot: enough for any split into launches
    const int64_t slots = (kind == 1 ? D : 1) * B;
    return gsx::vario_pair_blocks(P_or_n0) * (slots > 0 ? slots : 1) * 16;
}

int gsx_vario_unstructured_f64(int64_t F, int64_t P, int64_t dim, int64_t B, const double* f, const double* edges,
                               const double* pos, int estimator, double* gamma, int64_t* counts, void* workspace,
                               int64_t workspace_bytes, void* stream) {
    return gsx::vario_pairs(F, P, dim, B, 1, f, edges, pos, nullptr, 0.0, -1.0, 0, estimator, 0, gamma, counts,
                            workspace, workspace_bytes, stream);
}

int gsx_vario_unstructured_haversine_f64(int64_t F, int64_t P, int64_t B, const double* f, const double* edges,
                                         const double* latlon, int estimator, double* gamma, int64_t* counts,
                                         void* workspace, int64_t workspace_bytes, void* stream) {
    return gsx::vario_pairs(F, P, 2, B, 1, f, edges, latlon, nullptr, 0.0, -1.0, 0, estimator, 0, gamma, counts,
                            workspace, workspace_bytes, stream, true);
}

int gsx_vario_directional_f64(int64_t F, int64_t P, int64_t dim, int64_t B, int64_t D, const double* f,
                              const double* edges, const double* pos, const double* directions, double angles_tol,
                              double bandwidth, int separate_dirs, int estimator, double* gamma, int64_t* counts,
                              void* workspace, int64_t workspace_bytes, void* stream) {
    return gsx::vario_pairs(F, P, dim, B, D, f, edges, pos, directions, angles_tol, bandwidth, separate_dirs ? 1 : 0,
                            estimator, 1, gamma, counts, workspace, workspace_bytes, stream);
}

int gsx_vario_structured_f64(int64_t n0, int64_t n1, const double* f, const uint8_t* mask, int estimator,
                             double* gamma, int64_t* counts, void* workspace, int64_t workspace_bytes,
                             void* stream) {
    if (n0 < 0 || n1 < 0 || (estimator != 0 && estimator != 1)) return GSX_E_INVALID;
    if (n0 == 0) return GSX_OK;
    if (!gamma || !counts || (n1 > 0 && !f)) return GSX_E_INVALID;
    if (n1 >= 64) return gsx::lags_wide(n0, n1, f, mask, estimator, gamma, counts, workspace, workspace_bytes, stream);
    const int64_t tiles = gsx::ceil_div(n1, (int64_t)256) > 0 ? gsx::ceil_div(n1, (int64_t)256) : 1;
    // columns per tile: 256, or the next power of two >= n1 (then one tile)
    int cols_log2 = 0;
    while (cols_log2 < 8 && ((int64_t)1 << cols_log2) < n1) ++cols_log2;
    if (n0 > 0x7fffffffLL / tiles) return GSX_E_UNSUPPORTED;  // one block per (lag, tile)
    if (!workspace || workspace_bytes < n0 * tiles * 16) return GSX_E_WORKSPACE;
    hipStream_t s = gsx::as_stream(stream);
    double* part_sum = static_cast<double*>(workspace);
    unsigned long long* part_cnt = reinterpret_cast<unsigned long long*>(part_sum + n0 * tiles);
    hipLaunchKernelGGL(gsx::k_vario_lags, dim3((unsigned)(n0 * tiles)), dim3(256), 0, s, n0, n1, tiles, cols_log2,
                       f, mask, estimator, part_sum, part_cnt);
    int st = gsx::launch_status();
    if (st != GSX_OK) return st;
    long long* cnt = reinterpret_cast<long long*>(counts);
    hipLaunchKernelGGL(gsx::k_vario_lag_collect, dim3((unsigned)gsx::ceil_div(n0, (int64_t)256)), dim3(256), 0, s, n0,
                       tiles, part_sum, part_cnt, gamma, cnt);
    st = gsx::launch_status();
    if (st != GSX_OK) return st;
    hipLaunchKernelGGL(gsx::k_vario_normalize, dim3((unsigned)gsx::ceil_div(n0, (int64_t)256)), dim3(256), 0, s, n0,
                       estimator, gamma, cnt);
    return gsx::launch_status();
}

}  // extern "C"
