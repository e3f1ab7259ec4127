/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement of the randomisation-method summation that GSTools'
 * RandMeth / IncomprRandMeth call through `_summate` / `_summate_incompr`
 * (src/gstools/field/generator.py:42-64).  The arithmetic itself lives in
 * the third-party package gstools-cython >=1,<2 (pyproject.toml:45), which is
 * not vendored in the reference; this file restates its published formula:
 *
 *   summate         generator.py:193-199
 *       out[i] = sum_j z1[j] cos(theta_ij) + z2[j] sin(theta_ij)
 *       theta_ij = sum_d cov[d, j] * pos[d, i]          (d in order)
 *   summate_fourier generator.py:619-624 (Fourier.__call__, :659-691)
 *       out[i] = sum_j f[j] (z1[j] cos(theta_ij) + z2[j] sin(theta_ij)),
 *       f = spectrum factor sqrt(S(|k_j|) prod(dk)) computed by the caller
 *   summate_incompr generator.py:479-494
 *       out[c, i] = sum_j p_c(k_j) (z1[j] cos + z2[j] sin),
 *       p_c(k) = delta_{c0} - k_c k_0 / |k|^2
 *
 * Loop order mirrors the reference's thread-parallel CPU loop: parallel over
 * points, serial over modes in index order, phase accumulated from 0 in
 * dimension order.  Built with -O3 -fopenmp -ffp-contract=off and no
 * fast-math so that no multiply-add is fused: the phase is the exact
 * rounded sequence ((k0*x0) + k1*x1) + k2*x2 that a plain C loop produces.
 *
 * Only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark load this library.  The shipped product never calls it.
 */
#include <math.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* cov: (d, N) row-major; pos: (d, P) row-major; out: (P,) */
void oracle_summate(int64_t d, int64_t N, int64_t P, const double *cov,
                    const double *z1, const double *z2, const double *pos,
                    double *out, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < P; ++i) {
        double acc = 0.0;
        for (int64_t j = 0; j < N; ++j) {
            double phase = 0.0;
            for (int64_t k = 0; k < d; ++k) phase += cov[k * N + j] * pos[k * P + i];
            acc += z1[j] * cos(phase) + z2[j] * sin(phase);
        }
        out[i] = acc;
    }
}

/* Fourier method: the same sum with a per-mode spectrum factor f (N,). */
void oracle_summate_fourier(int64_t d, int64_t N, int64_t P, const double *f,
                            const double *cov, const double *z1,
                            const double *z2, const double *pos, double *out,
                            int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < P; ++i) {
        double acc = 0.0;
        for (int64_t j = 0; j < N; ++j) {
            double phase = 0.0;
            for (int64_t k = 0; k < d; ++k) phase += cov[k * N + j] * pos[k * P + i];
            acc += f[j] * (z1[j] * cos(phase) + z2[j] * sin(phase));
        }
        out[i] = acc;
    }
}

/* out: (d, P) row-major.  A k = 0 mode gives 0/0 like the reference formula. */
void oracle_summate_incompr(int64_t d, int64_t N, int64_t P, const double *cov,
                            const double *z1, const double *z2,
                            const double *pos, double *out, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < P; ++i) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int64_t j = 0; j < N; ++j) {
            double phase = 0.0, k2 = 0.0;
            for (int64_t k = 0; k < d; ++k) {
                phase += cov[k * N + j] * pos[k * P + i];
                k2 += cov[k * N + j] * cov[k * N + j];
            }
            double w = z1[j] * cos(phase) + z2[j] * sin(phase);
            double k0 = cov[j];
            for (int64_t c = 0; c < d; ++c) {
                double proj = (c == 0 ? 1.0 : 0.0) - cov[c * N + j] * k0 / k2;
                acc[c] += proj * w;
            }
        }
        for (int64_t c = 0; c < d; ++c) out[c * P + i] = acc[c];
    }
}

/*
 * Adjoint of summate for an upstream gradient g (P,):
 *   gz1[j]   = sum_i g_i cos(theta_ij)
 *   gz2[j]   = sum_i g_i sin(theta_ij)
 *   gk[d, j] = sum_i g_i x_id (z2[j] cos(theta_ij) - z1[j] sin(theta_ij))
 * (unscaled; the caller multiplies by sqrt(var / N)).  Parallel over modes.
 */
void oracle_summate_adjoint(int64_t d, int64_t N, int64_t P, const double *cov,
                            const double *z1, const double *z2,
                            const double *pos, const double *g, double *gz1,
                            double *gz2, double *gk, int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < N; ++j) {
        double c_acc = 0.0, s_acc = 0.0, kx[3] = {0.0, 0.0, 0.0};
        for (int64_t i = 0; i < P; ++i) {
            double phase = 0.0;
            for (int64_t k = 0; k < d; ++k) phase += cov[k * N + j] * pos[k * P + i];
            double c = cos(phase), s = sin(phase);
            c_acc += g[i] * c;
            s_acc += g[i] * s;
            double w = g[i] * (z2[j] * c - z1[j] * s);
            for (int64_t k = 0; k < d; ++k) kx[k] += w * pos[k * P + i];
        }
        gz1[j] = c_acc;
        gz2[j] = s_acc;
        for (int64_t k = 0; k < d; ++k) gk[k * N + j] = kx[k];
    }
}
