// Batched persistent summation (placeholder until the persistent kernel lands).
#include "gsx_internal.h"

extern "C" {
int gsx_batch_plan_create(const gsx_batch_item*, int64_t, int64_t, gsx_batch_plan** plan) {
    if (plan) *plan = nullptr;
    return GSX_E_UNSUPPORTED;
}
int gsx_batch_plan_launch(gsx_batch_plan*, void*) { return GSX_E_UNSUPPORTED; }
int gsx_batch_plan_launch_graph(gsx_batch_plan*, void*) { return GSX_E_UNSUPPORTED; }
int64_t gsx_batch_plan_tiles(const gsx_batch_plan*) { return 0; }
void gsx_batch_plan_destroy(gsx_batch_plan*) {}
}
