// Batched form of the ARD step (placeholder until the face-shell kernels land).
#include "ard_ctx_impl.h"

namespace ard {

int finalize_batched(Ctx& c) { return set_error(ARD_ERR_UNSUPPORTED, "batched form not built yet"); }
int step_batched(Ctx& c, int n) { return set_error(ARD_ERR_UNSUPPORTED, "batched form not built yet"); }
int batched_pressure_field(Ctx& c, int part, const float** out) {
  return set_error(ARD_ERR_UNSUPPORTED, "batched form not built yet");
}
int batched_set_force(Ctx& c, int part, int x, int y, int z, float f) {
  return set_error(ARD_ERR_UNSUPPORTED, "batched form not built yet");
}

}  // namespace ard
