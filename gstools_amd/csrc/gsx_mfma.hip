// Structured-grid summation on the matrix cores (placeholder until the GEMM kernel lands).
#include "gsx_internal.h"

extern "C" {
int64_t gsx_summate_grid_mfma_workspace(int64_t d, int64_t N, const int64_t* shape) {
    (void)d; (void)N; (void)shape;
    return 0;
}
int gsx_summate_grid_mfma(int64_t, int64_t, const int64_t*, const double*, const double*, const double*,
                          const double*, int, double, float*, void*, int64_t, void*) {
    return GSX_E_UNSUPPORTED;
}
}
