// stage.hip — per-world record resolution ahead of the render kernels (no counterpart in the
// reference, whose kernels read the model arrays directly).
#include "abi_util.hpp"

namespace mjpbr {

// Every world's staged records into geo.world_stage, one block a world.
__global__ void __launch_bounds__(256) stage_prep_kernel(MjpbrScene sc) {
  const int w = blockIdx.x;
  stage_world_to(sc, w, sc.geo.world_stage + (int64_t)w * stage_floats(sc));
}

int launch_stage_prep(const MjpbrScene* sc, hipStream_t stream) {
  if (!sc->geo.world_stage) return MJPBR_OK;
  if (reinterpret_cast<uintptr_t>(sc->geo.world_stage) % 16) return set_error(MJPBR_EINVAL, "world_stage must be 16-byte aligned");
  if (sc->geo.nworld < 1) return set_error(MJPBR_EINVAL, "nworld must be at least 1");
  hipLaunchKernelGGL(stage_prep_kernel, dim3(sc->geo.nworld), dim3(256), 0, stream, *sc);
  return check_launch("stage_prep_kernel");
}

}  // namespace mjpbr

extern "C" int64_t mjpbr_world_stage_floats(const MjpbrScene* scene) {
  if (!scene) return mjpbr::set_error(MJPBR_EINVAL, "scene must not be NULL");
  if (scene->geo.tlas_depth < 1 || scene->geo.tlas_depth > 12)
    return mjpbr::set_error(MJPBR_EINVAL, "scene BVH depth must be in [1, 12], got %d", scene->geo.tlas_depth);
  if (scene->geo.ndirect < 0 || scene->geo.ndirect > scene->geo.ngeom)
    return mjpbr::set_error(MJPBR_EINVAL, "ndirect must be in [0, ngeom], got %d", scene->geo.ndirect);
  return mjpbr::stage_floats(*scene);
}
