/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement of the kriging summation that GSTools' Krige calls through
 * `_calc_field_krige` / `_calc_field_krige_and_variance`
 * (src/gstools/krige/base.py:42-61, called at :307-317).  The arithmetic
 * lives in the third-party package gstools-cython >=1,<2 (pyproject.toml:45,
 * imported at krige/base.py:16-19), which is not vendored; this file restates
 * its published loop:
 *
 *   for every point k (parallel):
 *       for i in 0..K-1:
 *           w = sum_j krig_mat[i, j] * krig_vecs[j, k]     (j in order)
 *           field[k] += cond[i] * w
 *           error[k] += krig_vecs[i, k] * w                (variance form only)
 *
 * krig_mat (K, K) and krig_vecs (K, P) row-major, cond (K,).  The caller
 * turns `error` into the kriging variance max(sill - error, 0)
 * (krige/base.py:292-296).  Built without fused multiply-add.
 */
#include <stdint.h>

void oracle_krige(int64_t K, int64_t P, const double *krig_mat,
                  const double *krig_vecs, const double *cond, double *field,
                  double *error) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < P; ++k) {
        double f = 0.0, e = 0.0;
        for (int64_t i = 0; i < K; ++i) {
            double w = 0.0;
            for (int64_t j = 0; j < K; ++j) w += krig_mat[i * K + j] * krig_vecs[j * P + k];
            f += cond[i] * w;
            e += krig_vecs[i * P + k] * w;
        }
        field[k] = f;
        if (error) error[k] = e;
    }
}
