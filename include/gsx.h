/*
 * gsx -- MI355X (gfx950) randomisation-method field summation, C ABI.
 *
 * Drop-in boundary for the native summation that GSTools' generators call:
 *
 *   reference                                   replaced by
 *   ---------------------------------------------------------------------
 *   summate(cov_samples, z_1, z_2, pos, nt)      gsx_summate_f64 / _f32
 *     src/gstools/field/generator.py:22, :42-48, call :266-268
 *   summate_incompr(cov_samples, z_1, z_2, pos)  gsx_summate_incompr_f64 / _f32
 *     src/gstools/field/generator.py:24, :51-64, call :554-560
 *   summate on a meshgrid from Field.pre_pos     gsx_summate_grid_f64 / _f32,
 *     src/gstools/field/base.py:254-297 ->         gsx_summate_grid_mfma, gsx_summate_grid_gemm_f64
 *     tools/geometric.py:340-356
 *   summate_incompr on a meshgrid                gsx_summate_incompr_grid_f64 / _f32,
 *     src/gstools/field/generator.py:529-567      gsx_summate_incompr_grid_mfma
 *   summate_fourier(spectrum_factor, modes,       gsx_summate_fourier_f64 / _f32,
 *     z_1, z_2, pos)                              gsx_summate_fourier_grid_f64 / _f32,
 *     src/gstools/field/generator.py:23, :67-75,  gsx_summate_fourier_grid_mfma,
 *     call :681-688                               gsx_summate_fourier_tensor_f64
 *   (no reference counterpart: gradient of        gsx_summate_adjoint_f64 / _f32
 *    summate for calibration; formula is the
 *    derivative of generator.py:193-199)
 *   calc_field_krige(_and_variance)(krig_mat,     gsx_krige_f64, gsx_krige_workspace
 *     krig_vecs, cond)
 *     src/gstools/krige/base.py:16-19, :42-61,
 *     call :307-317
 *   variogram kernels unstructured / directional  gsx_vario_unstructured_f64,
 *     / structured / ma_structured                gsx_vario_directional_f64,
 *     src/gstools/variogram/variogram.py:12-15,   gsx_vario_structured_f64,
 *     :44-113, call :438-466, :560-565            gsx_vario_workspace
 *   unstructured(..., distance_type='h')          gsx_vario_unstructured_haversine_f64
 *     (latlon=True, variogram.py:446)
 *   emcee stretch-move draws of                   gsx_stretch_draws (host only),
 *     RNG.sample_ln_pdf for many chains,          gsx_stretch_run_device,
 *     src/gstools/random/rng.py:38-104            gsx_stretch_pick_device
 *   (no reference counterpart: many realizations  gsx_batch_*
 *    in one launch; each item is one
 *    RandMeth.__call__, generator.py:243-270)
 *
 * Conventions (all entry points):
 *   - every array argument is a DEVICE pointer owned by the caller (except
 *     gsx_stretch_draws, which works on host memory only); inputs
 *     are read-only, outputs are written in full;
 *   - cov is (d, N) row-major, pos is (d, P) row-major, incompressible
 *     outputs are (d, P) row-major -- the reference's numpy layouts;
 *   - outputs are the UNSCALED sum (the reference's summate); callers apply
 *     sqrt(var / N) as RandMeth.__call__ does (generator.py:270);
 *   - work is enqueued on `stream` (a hipStream_t, 0 = null stream) and the
 *     call returns without synchronising; no call allocates device memory
 *     except gsx_batch_plan_create (the plan's descriptors) -- scratch space
 *     comes from a caller-provided workspace; calls on different streams are
 *     independent (calls sharing one workspace must share one stream);
 *   - `shape` arrays are HOST memory;
 *   - the return value is GSX_OK (0) or a negative GSX_E_* status;
 *     gsx_status_string() names it.
 */
#ifndef GSX_H
#define GSX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSX_OK 0
#define GSX_E_INVALID (-1)   /* bad dimension, size or null pointer       */
#define GSX_E_LAUNCH (-2)    /* kernel launch or HIP runtime failure       */
#define GSX_E_NO_DEVICE (-3) /* no usable HIP device                       */
#define GSX_E_WORKSPACE (-4) /* workspace pointer null or too small        */
#define GSX_E_UNSUPPORTED (-5) /* valid request this build does not cover  */
#define GSX_E_SCRATCH (-6)   /* request exceeds the static scratch array of
                                gsx_summate_adjoint_f32_noalloc           */

/* MFMA split-precision variants for gsx_summate_grid_mfma. */
#define GSX_MFMA_BF16X3 0 /* bf16 hi/lo split, hi*hi + hi*lo + lo*hi */
#define GSX_MFMA_FP16X3 1 /* fp16 hi/lo split, hi*hi + hi*lo + lo*hi */
/* Tuning choices or-ed into the variant (results differ in rounding only):
 * GSX_MFMA_FILL_LOG2(f), f in 8..11: grids with fewer than 2^f blocks split
 * the mode (K) steps over about 2^f / blocks blocks (at least 128 modes each)
 * and sum the partial sums in a fixed order; 0 = the default 2^9.
 * GSX_MFMA_NO_SWAP: never run a 3D grid with axes 0 and 1 swapped (by
 * default a short axis 0 is swapped with axis 1 when that needs fewer
 * tiles).  The workspace sizes below cover every choice. */
#define GSX_MFMA_FILL_SHIFT 8
#define GSX_MFMA_FILL_LOG2(f) ((f) << GSX_MFMA_FILL_SHIFT)
#define GSX_MFMA_NO_SWAP 0x1000
/* Flag or-ed into the variant: the caller states that the last axis of a 2D
 * grid is uniformly spaced (y_j = y_0 + j h: y_j == y_(16a) + (y_b - y_0),
 * j = 16a + b, to 4 ulp of max|y|).  With at least 256 points on it the
 * library factorises that axis into two short ones (no n x N phase table;
 * faster).  Checked on the device: a non-uniform axis gives an all-NaN
 * output.  Ignored for 3D grids, incompressible sums and shorter axes. */
#define GSX_MFMA_UNIFORM_LAST 16

const char* gsx_status_string(int status);
int gsx_version(void);

/* Plain FP64 summation, out (P,) float64 = sum_j z1 cos(k_j.x) + z2 sin(k_j.x).
 * Phase computed as ((k0*x0) + k1*x1) + k2*x2 without fused multiply-add,
 * like the reference loop.  Replaces summate (generator.py:42-48). d in 1..3. */
int gsx_summate_f64(int64_t d, int64_t N, int64_t P, const double* cov,
                    const double* z1, const double* z2, const double* pos,
                    double* out, void* stream);

/* FP32 summation of the same sum, out (P,) float32.  Phase in FP64 (range
 * reduction is the precision-critical step), cosine and mode sum in FP32 with
 * blocked FP32 -> FP64 accumulation.  Replaces summate where an FP32 field is
 * acceptable (error <= 1e-5 relative to FP64).  The phase is reduced by magic
 * addition for the modes whose |k/2pi|_1 times the largest |x|_inf of the
 * point's aligned 2048-point segment stays within 2^18 revolutions, by
 * fract-and-convert for the others: a point's bits depend on its segment's
 * largest coordinate as well as its own (the batch plan applies the same
 * rule, so its items equal their single launches bit for bit). */
int gsx_summate_f32(int64_t d, int64_t N, int64_t P, const double* cov,
                    const double* z1, const double* z2, const double* pos,
                    float* out, void* stream);

/* Incompressible vector sum, out (d, P): projector e1 - k k0/|k|^2 per mode.
 * Replaces summate_incompr (generator.py:51-64, formula :479-494).  d in 2..3. */
int gsx_summate_incompr_f64(int64_t d, int64_t N, int64_t P, const double* cov,
                            const double* z1, const double* z2,
                            const double* pos, double* out, void* stream);
int gsx_summate_incompr_f32(int64_t d, int64_t N, int64_t P, const double* cov,
                            const double* z1, const double* z2,
                            const double* pos, float* out, void* stream);

/* Structured grid, direct sum: positions are the 'ij' meshgrid of the axes
 * (generate_grid, tools/geometric.py:340-356) and are never materialised.
 * shape[d] = axis lengths; axes = concatenated axis coordinates (sum(shape)
 * doubles); out (prod(shape),) in C order.  The FP64 form gives the same
 * phases, bit for bit, as gsx_summate_f64 on the materialised meshgrid. */
int gsx_summate_grid_f64(int64_t d, int64_t N, const int64_t* shape,
                         const double* axes, const double* cov,
                         const double* z1, const double* z2, double* out,
                         void* stream);
int gsx_summate_grid_f32(int64_t d, int64_t N, const int64_t* shape,
                         const double* axes, const double* cov,
                         const double* z1, const double* z2, float* out,
                         void* stream);

/* Structured grid on the matrix cores (d = 2 or 3): per-axis phase factors
 * e^{i k_j x} turn the sum into a real GEMM with K = 2N per z-plane, run with
 * split bf16/fp16 products accumulated in FP32.  `shape` and `axes` as above;
 * cov may already include a folded isometrize matrix (k' = M^T k).  out is
 * float32, scaled by `alpha` (pass 1.0 for the unscaled sum).  Workspace
 * bytes from gsx_summate_grid_mfma_workspace(). */
int64_t gsx_summate_grid_mfma_workspace(int64_t d, int64_t N, const int64_t* shape);
/* The matrix-core plan of a grid for one variant (comps 1, or d for the
 * incompressible form), host arithmetic only: info[0] axes 0 and 1 swapped,
 * [1] uniform last axis factorised, [2] row tiles, [3] column tiles, [4] K
 * split, [5] blocks, [6] workspace bytes of this variant, [7] 3 for the
 * 16 x 16 column-tile kernel, 2 for flat columns.  For tuning tools: two
 * variants with equal info launch the same kernels. */
int gsx_summate_grid_mfma_plan(int64_t d, int64_t N, const int64_t* shape,
                               int64_t comps, int variant, int64_t* info);
int gsx_summate_grid_mfma(int64_t d, int64_t N, const int64_t* shape,
                          const double* axes, const double* cov,
                          const double* z1, const double* z2, int variant,
                          double alpha, float* out, void* workspace,
                          int64_t workspace_bytes, void* stream);

/* Incompressible sum on a structured grid (summate_incompr at the meshgrid of
 * the axes, generator.py:479-494, :529-567 with generate_grid positions):
 * out (d, prod(shape)) in C order, d = 2 or 3.  The direct forms compute each
 * phase once for all d components, as gsx_summate_incompr_f64 / _f32 on the
 * materialised grid (same phases).  The matrix-core form runs one GEMM per
 * component in one launch: component c is the scalar sum with the projector
 * weight p_c(k_j) = delta_c0 - k_jc k_j0 / |k_j|^2 per mode, scaled by alpha;
 * workspace from gsx_summate_incompr_grid_mfma_workspace(). */
int gsx_summate_incompr_grid_f64(int64_t d, int64_t N, const int64_t* shape,
                                 const double* axes, const double* cov,
                                 const double* z1, const double* z2,
                                 double* out, void* stream);
int gsx_summate_incompr_grid_f32(int64_t d, int64_t N, const int64_t* shape,
                                 const double* axes, const double* cov,
                                 const double* z1, const double* z2,
                                 float* out, void* stream);
int64_t gsx_summate_incompr_grid_mfma_workspace(int64_t d, int64_t N,
                                                const int64_t* shape);
int gsx_summate_incompr_grid_mfma(int64_t d, int64_t N, const int64_t* shape,
                                  const double* axes, const double* cov,
                                  const double* z1, const double* z2,
                                  int variant, double alpha, float* out,
                                  void* workspace, int64_t workspace_bytes,
                                  void* stream);

/* Fourier method: out = sum_j f_j (z1_j cos(k_j.x) + z2_j sin(k_j.x)) with the
 * per-mode spectrum factor f = spectrum_factor (N,) (Fourier.reset_seed,
 * generator.py:812-818) and modes (d, N) the regular wavenumber grid.  Same
 * phase, precision and layout conventions as gsx_summate_f64 / _f32; the
 * grid forms take `shape` / `axes` like gsx_summate_grid_f64 and
 * gsx_summate_grid_mfma.  The output is unscaled, as Fourier.__call__ adds
 * no sqrt(var / N) (generator.py:681-691).  Replaces summate_fourier
 * (generator.py:67-75). */
int gsx_summate_fourier_f64(int64_t d, int64_t N, int64_t P,
                            const double* spectrum_factor, const double* modes,
                            const double* z1, const double* z2,
                            const double* pos, double* out, void* stream);
int gsx_summate_fourier_f32(int64_t d, int64_t N, int64_t P,
                            const double* spectrum_factor, const double* modes,
                            const double* z1, const double* z2,
                            const double* pos, float* out, void* stream);
int gsx_summate_fourier_grid_f64(int64_t d, int64_t N, const int64_t* shape,
                                 const double* axes,
                                 const double* spectrum_factor,
                                 const double* modes, const double* z1,
                                 const double* z2, double* out, void* stream);
int gsx_summate_fourier_grid_f32(int64_t d, int64_t N, const int64_t* shape,
                                 const double* axes,
                                 const double* spectrum_factor,
                                 const double* modes, const double* z1,
                                 const double* z2, float* out, void* stream);
int gsx_summate_fourier_grid_mfma(int64_t d, int64_t N, const int64_t* shape,
                                  const double* axes,
                                  const double* spectrum_factor,
                                  const double* modes, const double* z1,
                                  const double* z2, int variant, double alpha,
                                  float* out, void* workspace,
                                  int64_t workspace_bytes, void* stream);

/* Structured grid as an FP64 GEMM on the matrix cores (an opt-in FP64-class
 * form of gsx_summate_grid_f64 for the same meshgrid sum, generate_grid,
 * tools/geometric.py:340-356): rows are axis 0, columns the other axes, the
 * per-column complex phase products built in column chunks of at most 512
 * MB.  Each axis's phase is the rounded product k_i x_i; their sum is not
 * rounded, so the result differs from the per-point sum by about
 * |k . x| 2^-52 (relative) per term.  spectrum_factor may be null (the
 * randomisation method) or the Fourier method's factors.  out float64.
 * d in 1..3; workspace bytes from gsx_summate_grid_gemm_workspace(). */
int64_t gsx_summate_grid_gemm_workspace(int64_t d, int64_t N, const int64_t* shape);
int gsx_summate_grid_gemm_f64(int64_t d, int64_t N, const int64_t* shape,
                              const double* axes, const double* cov,
                              const double* z1, const double* z2,
                              const double* spectrum_factor, double* out,
                              void* workspace, int64_t workspace_bytes,
                              void* stream);

/* Fourier method on a structured grid when the modes are a tensor grid (the
 * Fourier generator's own: the 'ij' meshgrid of one wavenumber axis per
 * dimension, generator.py:838-862, summed at generator.py:659-691).  The sum
 * factorises into one contraction per axis, run as FP64 GEMMs on the matrix
 * cores: mode_no[d] wavenumbers per axis, kaxes = their concatenation
 * (sum(mode_no) doubles), spectrum_factor / z1 / z2 over all prod(mode_no)
 * modes in the meshgrid's C order; shape / axes as for the grid entries;
 * out (prod(shape),) float64, the unscaled sum.  FP64 throughout; it differs
 * from the per-point sum only by the rounding of the phase sum.  d in 1..3.
 * Workspace bytes from gsx_summate_fourier_tensor_workspace(). */
int64_t gsx_summate_fourier_tensor_workspace(int64_t d, const int64_t* mode_no,
                                             const int64_t* shape);
int gsx_summate_fourier_tensor_f64(int64_t d, const int64_t* mode_no,
                                   const double* kaxes, const int64_t* shape,
                                   const double* axes,
                                   const double* spectrum_factor,
                                   const double* z1, const double* z2,
                                   double* out, void* workspace,
                                   int64_t workspace_bytes, void* stream);

/* Adjoint of summate for an upstream gradient g (P,):
 *   gz1[j]   = sum_i g_i cos(k_j.x_i)
 *   gz2[j]   = sum_i g_i sin(k_j.x_i)
 *   gk[d][j] = sum_i g_i x_id (z2_j cos(k_j.x_i) - z1_j sin(k_j.x_i))
 * (unscaled, like the forward).  Deterministic two-stage reduction over
 * points.  Workspace bytes from gsx_summate_adjoint_workspace(). */
int64_t gsx_summate_adjoint_workspace(int64_t d, int64_t N, int64_t P);
int gsx_summate_adjoint_f64(int64_t d, int64_t N, int64_t P, const double* cov,
                            const double* z1, const double* z2,
                            const double* pos, const double* g, double* gz1,
                            double* gz2, double* gk, void* workspace,
                            int64_t workspace_bytes, void* stream);
int gsx_summate_adjoint_f32(int64_t d, int64_t N, int64_t P, const double* cov,
                            const double* z1, const double* z2,
                            const double* pos, const float* g, double* gz1,
                            double* gz2, double* gk, void* workspace,
                            int64_t workspace_bytes, void* stream);

/* Allocation-free form of gsx_summate_adjoint_f32 (no counterpart in the
 * reference).  For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same arguments and results as gsx_summate_adjoint_f32 minus the
 * workspace: the partial sums live in a static device array of
 * GSX_NOALLOC_SCRATCH_BYTES per GPU, held from the library's first use on
 * that GPU until the process ends.  It calls no hipMalloc / hipFree.
 *   - limit: (2 + 2d) * N * 8 bytes must fit the array, i.e.
 *     N <= GSX_NOALLOC_SCRATCH_BYTES / (16 (1 + d)) (1048576 modes for
 *     d = 1, 524288 for d = 3); larger N returns GSX_E_SCRATCH;
 *   - any P; whenever the workspace form's partial sums fit the array the
 *     results equal gsx_summate_adjoint_f32's bit for bit, otherwise they
 *     use fewer point chunks and differ in rounding only;
 *   - every call on one GPU shares that GPU's array: calls that overlap on
 *     one GPU (two streams, two threads) race and must be serialised by
 *     the caller. */
#define GSX_NOALLOC_SCRATCH_BYTES (32LL << 20)
int gsx_summate_adjoint_f32_noalloc(int64_t d, int64_t N, int64_t P,
                                    const double* cov, const double* z1,
                                    const double* z2, const double* pos,
                                    const float* g, double* gz1, double* gz2,
                                    double* gk, void* stream);

/* Kriging sums, FP64: for every target point k (column of krig_vecs),
 *   w = krig_mat . v_k,  field[k] = cond . w,  error[k] = v_k . w.
 * krig_mat (K, K) and krig_vecs (K, P) row-major, cond (K,).  error == NULL
 * computes the field only (calc_field_krige), otherwise both
 * (calc_field_krige_and_variance; the caller forms max(sill - error, 0)).
 * Replaces the external kernels called at src/gstools/krige/base.py:42-61.
 * Workspace bytes from gsx_krige_workspace(K, P, error != NULL), 16-byte
 * aligned (GSX_E_INVALID otherwise when error != NULL). */
int64_t gsx_krige_workspace(int64_t K, int64_t P, int variance);
/* Covariance rows of krig_vecs: out[i * ld_out + p] = C(|c_i - x_p|) for
 * conditioning points cond_pos (dim, K) and targets pos (dim, P), both
 * isometrized, row-major; distances rounded like scipy's cdist.  C =
 * var * cor(r / len_rescaled) for model_kind 0 Gaussian exp(-h^2),
 * 1 Exponential exp(-h), 2 Stable exp(-h^alpha), 3 Spherical; with
 * sill_at_zero >= 0 the value at r <= 1e-8 is sill_at_zero (cov_nugget).
 * K <= 65535.  Replaces the host evaluation at src/gstools/krige/base.py:374-378. */
int gsx_krige_cov_f64(int model_kind, double var, double len_rescaled,
                      double alpha, double sill_at_zero, int64_t dim,
                      int64_t K, int64_t P, const double* cond_pos,
                      const double* pos, double* out, int64_t ld_out,
                      void* stream);
int gsx_krige_f64(int64_t K, int64_t P, const double* krig_mat,
                  const double* krig_vecs, const double* cond, double* field,
                  double* error, void* workspace, int64_t workspace_bytes,
                  void* stream);

/* Empirical variograms, FP64.  Fields f (F, P) row-major (several fields at
 * the same points), pos (dim, P) row-major, dim 1..3, bin edges (B + 1).
 * gamma (B) and counts (B) receive the estimate and the pair counts: pairs
 * j < k with edges[b] <= |x_j - x_k| < edges[b+1] (Euclidean), every field
 * whose two values are not NaN adds est(f_k - f_j) and one count; estimator
 * 0 = Matheron (gamma = sum (f_k - f_j)^2 / (2 N)), 1 = Cressie
 * (gamma = 0.5 (sum |f_k - f_j|^0.5 / N)^4 / (0.457 + 0.494/N + 0.045/N^2)),
 * N = max(count, 1).  Directional: D unit directions (D, dim); a pair also
 * needs acos(|s| / dist) < angles_tol, s = (x_k - x_j) . dir (distance-0
 * pairs pass) and, if bandwidth > 0, a distance below bandwidth from the
 * direction line; separate_dirs != 0 gives each pair to the first matching
 * direction; gamma, counts are (D, B).  Structured: lags 0..n0-1 along the
 * first axis of an (n0, n1) array; mask (n0, n1) bytes, non-zero = masked,
 * or NULL; gamma, counts (n0); n1 < 64: any n0 with n0 * ceil(n1 / 256) <
 * 2^31, n1 >= 64: fewer than 2^31 blocks of 64 columns x 32 or 64 lags and
 * n1 < 2^26 (GSX_E_UNSUPPORTED beyond).  Counts are exact; sums are reduced in a
 * fixed order (bitwise reproducible).  Workspace bytes:
 * gsx_vario_workspace(kind, P, dim, B, D) for kind 0 (unstructured) or 1
 * (directional), gsx_vario_workspace(2, n0, n1, 0, 1) for structured.
 * Replaces the external kernels of src/gstools/variogram/variogram.py:44-113. */
int64_t gsx_vario_workspace(int64_t kind, int64_t P_or_n0, int64_t dim_or_n1,
                            int64_t B, int64_t D);
int gsx_vario_unstructured_f64(int64_t F, int64_t P, int64_t dim, int64_t B,
                               const double* f, const double* edges,
                               const double* pos, int estimator, double* gamma,
                               int64_t* counts, void* workspace,
                               int64_t workspace_bytes, void* stream);
/* Great-circle variant of gsx_vario_unstructured_f64 (the backend's
 * unstructured(..., distance_type='h'), src/gstools/variogram/variogram.py:446
 * with latlon=True): latlon (2, P) = latitude, longitude in degrees; a pair's
 * distance is its great-circle angle in radians (haversine), edges in
 * radians.  Workspace: gsx_vario_workspace(0, P, 2, B, 1). */
int gsx_vario_unstructured_haversine_f64(int64_t F, int64_t P, int64_t B,
                                         const double* f, const double* edges,
                                         const double* latlon, int estimator,
                                         double* gamma, int64_t* counts,
                                         void* workspace, int64_t workspace_bytes,
                                         void* stream);
int gsx_vario_directional_f64(int64_t F, int64_t P, int64_t dim, int64_t B,
                              int64_t D, const double* f, const double* edges,
                              const double* pos, const double* directions,
                              double angles_tol, double bandwidth,
                              int separate_dirs, int estimator, double* gamma,
                              int64_t* counts, void* workspace,
                              int64_t workspace_bytes, void* stream);
int gsx_vario_structured_f64(int64_t n0, int64_t n1, const double* f,
                             const uint8_t* mask, int estimator, double* gamma,
                             int64_t* counts, void* workspace,
                             int64_t workspace_bytes, void* stream);

/* Host-only (no device memory, no stream): every random draw of the stretch-
 * move ensemble sampler that RNG.sample_ln_pdf runs through emcee
 * (src/gstools/random/rng.py:38-104), for R independent chains.  Chain r's
 * generator state is states[r * 625 .. r * 625 + 624]: the 624 MT19937 key
 * words and the position, as numpy's RandomState.get_state() reports them;
 * it is advanced in place.  For steps[r] <= max_steps steps, each of the
 * (R, max_steps, nwalkers) outputs receives, per step: labels = the shuffled
 * arange(nwalkers) % 2; uz, partner, uacc = the stretch variates, partner
 * indices and acceptance uniforms of the walkers labelled 0 (first
 * ceil(nwalkers / 2) slots) and then labelled 1.  nthreads > 1 splits the
 * chains over host threads.  Entries past steps[r] are left untouched. */
int gsx_stretch_draws(int64_t R, uint32_t* states, const int64_t* steps,
                      int64_t max_steps, int32_t nwalkers, int8_t* labels,
                      double* uz, int64_t* partner, double* uacc,
                      int32_t nthreads);

/* Host-only: every legacy numpy draw RandMeth.reset_seed makes
 * (src/gstools/field/generator.py:346-387, random/rng.py:38-193) for R
 * realizations sharing one model: master RandomState(seeds[r]) (a seed in
 * [0, 2^32)) hands out stream seeds randint(1, 2^16); the streams give z_1,
 * z_2 = normal(size=n) (polar method, cached second normal), then for
 * kind >= 1 the sphere draws (dim 1: choice([-1, 1]) as -1.0 / 1.0 in ang1;
 * dim 2: uniform(0, 2 pi) in ang1; dim 3: that plus uniform(-1, 1) in ang2),
 * then for kind 1 (inverse cdf) uniform(size=n) in u_inv, for kind 2 (MCMC)
 * the initial walkers rand(nwalkers) * sample_around (R, nwalkers), the
 * fresh burn-in and production streams (R, 2, 625) as gsx_stretch_run_device
 * takes them, and the final choice(n_flat[r], n) indices in picks.  kind 0
 * (zero variance) stops after z_2.  Per-mode outputs are concatenated over r
 * (offsets = prefix sums of n_modes).  nthreads > 1 splits the realizations
 * over host threads. */
int gsx_mode_draws(int64_t R, const int64_t* seeds, const int64_t* n_modes,
                   int32_t dim, int32_t kind, int32_t nwalkers,
                   double sample_around, const int64_t* n_flat, double* z1,
                   double* z2, double* ang1, double* ang2, double* u_inv,
                   double* init, uint32_t* states, int64_t* picks,
                   int32_t nthreads);

/* Host-only: the two halves of one stretch-move half-move for every chain r
 * with steps[r] > step, one-dimensional walkers (emcee's arithmetic, every
 * operation rounded, no fused multiply-add).  Draws are row block_row of
 * (R, block, nwalkers) arrays made by gsx_stretch_draws.  Propose writes the
 * proposals of the walkers labelled `split` of the n_active running chains,
 * compactly, to q (n_active, ns):  q = c_p - (c_p - x) zz,
 * zz = ((a - 1) u + 1)^2 / a.  The caller evaluates the log-probability of q
 * (new_lp, same layout); accept replaces walker and log-probability where
 * (0 + new_lp) - log_prob > log_uacc (log of the acceptance uniforms, taken
 * by the caller) and, when `chain` is not NULL, copies each running chain's
 * walkers to chain[r][step] of an (R, chain_steps, nwalkers) array.
 * nwalkers <= 4096. */
int gsx_stretch_propose(int64_t R, int32_t nwalkers, const int64_t* steps,
                        int64_t step, int64_t block_row, int64_t block,
                        int32_t split, double a, const double* coords,
                        const int8_t* labels, const double* uz,
                        const int64_t* partner, double* q, int64_t* n_active);
int gsx_stretch_accept(int64_t R, int32_t nwalkers, const int64_t* steps,
                       int64_t step, int64_t block_row, int64_t block,
                       int32_t split, const int8_t* labels,
                       const double* log_uacc, const double* q,
                       const double* new_lp, double* coords, double* log_prob,
                       double* chain, int64_t chain_steps);

/* Closed-form radial spectral log-pdf for the device sampler:
 * log(rad_fac(dim, r) |S(r)|) of covmodel/models.py with the host-computed
 * constants c[] (the scalar factors exactly as numpy computes them):
 *   GAUSSIAN     c0 exp(-((r L / 2)^2)),             c0 = (L / 2 / sqrt(pi))^dim
 *   EXPONENTIAL  c0 / (pi (1 + (r L)^2))^((dim+1)/2), c0 = L^dim Gamma((dim+1)/2)
 *   MATERN       c0 exp(((c1 log(1 + (r L)^2 / c2) + c3) - c4) - c5),
 *                c0 = (L / sqrt(pi))^dim, c1 = -(nu + dim/2), c2 = nu,
 *                c3 = lgamma(nu + dim/2), c4 = lgamma(nu), c5 = dim log(sqrt(nu))
 * (L = len_rescaled; Matern with nu <= 20 only). */
#define GSX_STRETCH_GAUSSIAN 0
#define GSX_STRETCH_EXPONENTIAL 1
#define GSX_STRETCH_MATERN 2
typedef struct gsx_stretch_model {
    int32_t kind;
    int32_t dim;
    double len_rescaled;
    double c[6];
} gsx_stretch_model;

/* Device form of the stretch-move sampler of RNG.sample_ln_pdf
 * (src/gstools/random/rng.py:38-104) for R one-dimensional chains of
 * nwalkers <= 64 walkers: one block per chain runs burn_steps steps on
 * the chain's first MT19937 stream, then prod_steps[r] steps on its second,
 * recording the walkers after every production step in chain (R, max_steps,
 * nwalkers) (may be NULL).  models: DEVICE array of R gsx_stretch_model, the
 * log-pdf of each chain (an unknown kind reads as a NaN log-probability).
 * states (R, 2, 625) uint32: the legacy numpy key
 * (624 words) and position of the burn-in and production streams; init
 * (R, nwalkers) the initial walkers.  The draws, their order and the move's
 * arithmetic are those of gsx_stretch_draws / _propose / _accept; the
 * log-pdf is the model's closed form on the device.  final_coords (R,
 * nwalkers) may be NULL.  status (device int, 0 on entry) becomes 1 for an
 * infinite or NaN proposal, 2 for a NaN log-probability. */
int gsx_stretch_run_device(int64_t R, int32_t nwalkers, const uint32_t* states,
                           int64_t burn_steps, const int64_t* prod_steps,
                           int64_t max_steps, double a,
                           const gsx_stretch_model* models, const double* init,
                           double* chain, double* final_coords, int* status,
                           void* stream);
/* out[i] = chain[owner[i]][idx[i]] (flat index step * nwalkers + walker) for
 * i < total: the final choice of samples from each chain's recorded walkers. */
int gsx_stretch_pick_device(int64_t total, const int64_t* owner,
                            const int64_t* idx, int64_t max_steps,
                            int32_t nwalkers, const double* chain, double* out,
                            void* stream);

/* Batched FP32 summation of many realizations in one persistent launch.
 * Each item is one RandMeth / IncomprRandMeth call (generator.py:243-270,
 * :529-567) with its own modes and points; see gsx_batch_item. */
typedef struct gsx_batch_item {
    int32_t d;          /* 1..3                                            */
    int32_t kind;       /* 0 = scalar summate, 1 = incompressible (d >= 2) */
    int64_t N;          /* modes                                           */
    int64_t P;          /* points                                          */
    int64_t shape[3];   /* grid axis lengths when axes != NULL             */
    const double* cov;  /* (d, N)                                          */
    const double* z1;   /* (N,)                                            */
    const double* z2;   /* (N,)                                            */
    const double* pos;  /* (d, P), or NULL when axes is given              */
    const double* axes; /* concatenated grid axes, or NULL                 */
    float* out;         /* (P,) or (d, P), unscaled                        */
} gsx_batch_item;

/* Build a plan for `count` items (host array of descriptors whose pointers
 * are device pointers).  tile_points = points per work tile (512, 1024 or
 * 2048; 0 = 1024); mode_chunk = modes staged per LDS chunk (128 or 256;
 * 0 = 256).  The plan owns a device copy of the descriptors and of the tile
 * queue, ordered longest-processing-time first (descending N). */
typedef struct gsx_batch_plan gsx_batch_plan;
int gsx_batch_plan_create(const gsx_batch_item* items, int64_t count,
                          int64_t tile_points, int64_t mode_chunk,
                          gsx_batch_plan** plan);
/* One persistent launch over every tile of every item.  A plan has one
 * device tile counter, reset by every launch: use a plan on one stream at a
 * time (an eager launch must not overlap a replay of the plan's graph, nor a
 * launch on another stream); concurrent work needs one plan per stream. */
int gsx_batch_plan_launch(gsx_batch_plan* plan, void* stream);
/* Capture the launch into a HIP graph once, then replay it. */
int gsx_batch_plan_launch_graph(gsx_batch_plan* plan, void* stream);
int64_t gsx_batch_plan_tiles(const gsx_batch_plan* plan);
void gsx_batch_plan_destroy(gsx_batch_plan* plan);

/* Measurement aid, no counterpart in the reference: the matrix-core rate
 * the chip attains with nothing but independent back-to-back MFMAs (16
 * accumulators per wave, no memory traffic in the loop), for the roofline
 * figures of the GEMM-shaped kernels.  kind GSX_PROBE_BF16 runs
 * v_mfma_f32_16x16x32_bf16 (16384 flops each), GSX_PROBE_F64
 * v_mfma_f64_16x16x4_f64 (2048 flops each); every wave of `blocks` blocks
 * of 256 threads issues 16 * iters of them.  random_operands != 0 gives
 * each lane random operands (toggling inputs, as real data: more power, a
 * lower clock), 0 near-constant ones.  out: DEVICE array of blocks * 256
 * floats (bf16) or doubles (f64), one checksum per thread. */
#define GSX_PROBE_BF16 0
#define GSX_PROBE_F64 1
int gsx_probe_mfma(int kind, int random_operands, int64_t iters, int64_t blocks, void* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GSX_H */
