// shp_solver.hip - solver scheduling for the contact step.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "shp_engine.h"

shp_status solver_alloc(shp_batch* b) {
  (void)b;
  return SHP_OK;
}

void launch_ordered_solver(shp_batch* b) { launch_serial_solver(b); }

void launch_colored_solver(shp_batch* b) { launch_serial_solver(b); }

extern "C" {

int32_t shp_mlp_param_count(void) { return 0; }

shp_status shp_batch_set_mlp(shp_batch* b, const float* params) {
  (void)b;
  (void)params;
  return SHP_E_UNSUPPORTED;
}

}  // extern "C"
