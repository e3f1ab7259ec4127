/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement of the empirical-variogram kernels that GSTools calls
 * through `_unstructured`, `_directional`, `_structured` and `_ma_structured`
 * (src/gstools/variogram/variogram.py:44-113, called at :438-466 and
 * :560-565).  The kernels live in the third-party package gstools-cython
 * >=1,<2 (pyproject.toml:45, imported at variogram.py:12-15), which is not
 * vendored; this file restates their published algorithm:
 *
 *   unstructured: for every bin i, every pair j < k of points with
 *       edges[i] <= |x_j - x_k| < edges[i+1] (Euclidean, dimensions summed
 *       in order), and every field m with neither value NaN:
 *       counts[i] += 1, sums[i] += est(f[m,k] - f[m,j]).
 *   directional: as unstructured, but a pair also counts for direction d
 *       (unit vector) when its projection s = (x_k - x_j).dir_d satisfies
 *       acos(|s| / dist) < angles_tol (pairs at distance 0 always pass) and,
 *       for bandwidth > 0, |(x_k - x_j) - s dir_d| < bandwidth; with
 *       separate_dirs the first matching direction takes the pair.
 *   structured: lag k along the first axis of a (n0, n1) array:
 *       counts[k] = (n0 - k) n1, sums[k] = sum est(f[i+k, j] - f[i, j]);
 *       ma_structured skips pairs with a masked end.
 *   est: x^2 (Matheron) or sqrt(|x|) (Cressie); normalisation
 *       Matheron gamma = sum / (2 max(N, 1)),
 *       Cressie  gamma = 0.5 (sum / N)^4 / (0.457 + 0.494 / N + 0.045 / N^2)
 *       with N = max(count, 1)  (variogram.py:172-185 states both formulas).
 *
 * Built without fused multiply-add.
 */
#include <math.h>
#include <stdint.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static double est(int cressie, double d) { return cressie ? sqrt(fabs(d)) : d * d; }

static void normalize(int cressie, int64_t n, double *sums, const int64_t *counts) {
    for (int64_t i = 0; i < n; ++i) {
        const double cnt = (double)(counts[i] > 0 ? counts[i] : 1);
        if (cressie)
            sums[i] = 0.5 * pow(sums[i] / cnt, 4.0) / (0.457 + 0.494 / cnt + 0.045 / (cnt * cnt));
        else
            sums[i] = sums[i] / (2.0 * cnt);
    }
}

static double dist_euclid(int64_t dim, int64_t P, const double *pos, int64_t j, int64_t k) {
    double s = 0.0;
    for (int64_t d = 0; d < dim; ++d) {
        const double t = pos[d * P + j] - pos[d * P + k];
        s += t * t;
    }
    return sqrt(s);
}

/* Great-circle angle (radians) between points j and k of pos = (lat, lon) in
 * degrees: the haversine form of the published backend (gstools-cython >= 1,
 * < 2, variogram kernel with distance_type 'h', selected by
 * src/gstools/variogram/variogram.py:446 for latlon=True):
 *   a = sin^2(dlat / 2) + cos(lat_j) cos(lat_k) sin^2(dlon / 2),
 *   d = 2 atan2(sqrt(a), sqrt(1 - a)),  angles in radians = degrees * (pi / 180). */
static double dist_haversine(int64_t P, const double *pos, int64_t j, int64_t k) {
    const double deg_2_rad = M_PI / 180.0;
    const double diff_lat = (pos[k] - pos[j]) * deg_2_rad;
    const double diff_lon = (pos[P + k] - pos[P + j]) * deg_2_rad;
    const double sl = sin(diff_lat / 2.0), so = sin(diff_lon / 2.0);
    const double arg = sl * sl + cos(pos[j] * deg_2_rad) * cos(pos[k] * deg_2_rad) * (so * so);
    return 2.0 * atan2(sqrt(arg), sqrt(1.0 - arg));
}

/* f (F, P), pos (dim, P), edges (B + 1) -> gamma (B), counts (B); haversine != 0:
 * pos is (lat, lon) in degrees and the distance the great-circle angle */
void oracle_vario_unstructured(int64_t F, int64_t P, int64_t dim, int64_t B, const double *f,
                               const double *edges, const double *pos, int cressie, int haversine,
                               double *gamma, int64_t *counts) {
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < B; ++i) {
        double s = 0.0;
        int64_t c = 0;
        for (int64_t j = 0; j < P - 1; ++j)
            for (int64_t k = j + 1; k < P; ++k) {
                const double dist = haversine ? dist_haversine(P, pos, j, k) : dist_euclid(dim, P, pos, j, k);
                if (dist < edges[i] || dist >= edges[i + 1]) continue;
                for (int64_t m = 0; m < F; ++m) {
                    const double a = f[m * P + k], b = f[m * P + j];
                    if (isnan(a) || isnan(b)) continue;
                    ++c;
                    s += est(cressie, a - b);
                }
            }
        gamma[i] = s;
        counts[i] = c;
    }
    normalize(cressie, B, gamma, counts);
}

static int dir_test(int64_t dim, int64_t P, const double *pos, double dist, const double *dir,
                    double angles_tol, double bandwidth, int64_t k, int64_t j) {
    double s = 0.0;
    for (int64_t d = 0; d < dim; ++d) s += (pos[d * P + k] - pos[d * P + j]) * dir[d];
    int in_band = 1, in_angle = 1;
    if (bandwidth > 0.0) {
        double b = 0.0;
        for (int64_t d = 0; d < dim; ++d) {
            const double t = (pos[d * P + k] - pos[d * P + j]) - s * dir[d];
            b += t * t;
        }
        in_band = sqrt(b) < bandwidth;
    }
    if (dist > 0.0) {
        const double t = fabs(s) / dist;
        if (t < 1.0) in_angle = acos(t) < angles_tol;
    }
    return in_band && in_angle;
}

/* directions (D, dim) unit vectors -> gamma (D, B), counts (D, B) */
void oracle_vario_directional(int64_t F, int64_t P, int64_t dim, int64_t B, int64_t D, const double *f,
                              const double *edges, const double *pos, const double *dirs, double angles_tol,
                              double bandwidth, int separate, int cressie, double *gamma, int64_t *counts) {
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < B; ++i) {
        for (int64_t d = 0; d < D; ++d) {
            gamma[d * B + i] = 0.0;
            counts[d * B + i] = 0;
        }
        for (int64_t j = 0; j < P - 1; ++j)
            for (int64_t k = j + 1; k < P; ++k) {
                const double dist = dist_euclid(dim, P, pos, j, k);
                if (dist < edges[i] || dist >= edges[i + 1]) continue;
                for (int64_t d = 0; d < D; ++d) {
                    if (!dir_test(dim, P, pos, dist, dirs + d * dim, angles_tol, bandwidth, k, j)) continue;
                    for (int64_t m = 0; m < F; ++m) {
                        const double a = f[m * P + k], b = f[m * P + j];
                        if (isnan(a) || isnan(b)) continue;
                        counts[d * B + i] += 1;
                        gamma[d * B + i] += est(cressie, a - b);
                    }
                    if (separate) break;
                }
            }
    }
    for (int64_t d = 0; d < D; ++d) normalize(cressie, B, gamma + d * B, counts + d * B);
}

/* f (n0, n1); mask (n0, n1) of 0/1 or NULL -> gamma (n0) */
void oracle_vario_structured(int64_t n0, int64_t n1, const double *f, const uint8_t *mask, int cressie,
                             double *gamma, int64_t *counts) {
#pragma omp parallel for schedule(dynamic)
    for (int64_t k = 0; k < n0; ++k) {
        double s = 0.0;
        int64_t c = 0;
        if (k > 0)
            for (int64_t i = 0; i + k < n0; ++i)
                for (int64_t j = 0; j < n1; ++j) {
                    if (mask && (mask[(i + k) * n1 + j] || mask[i * n1 + j])) continue;
                    ++c;
                    s += est(cressie, f[(i + k) * n1 + j] - f[i * n1 + j]);
                }
        gamma[k] = s;
        counts[k] = c;
    }
    normalize(cressie, n0, gamma, counts);
}
