// abi_util.hpp — host-side helpers of the C ABI: thread-local error messages, launch checks,
// tile maps of the render kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <stdarg.h>
#include <stdio.h>
#include "../../include/mjpbr.h"
#include "stage.hpp"

namespace mjpbr {

int set_error(int code, const char* fmt, ...);

inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(MJPBR_ELAUNCH, "%s: %s", what, hipGetErrorString(e));
  return MJPBR_OK;
}

inline size_t host_stage_bytes(const MjpbrScene* sc) { return (size_t)stage_floats(*sc) * 4; }

// LDS a render workgroup may use without lowering occupancy below 8 workgroups a compute unit.
constexpr size_t kRenderLdsPerGroup = 20 * 1024;

// Largest per-world record block the render kernels stage in LDS; larger worlds are read in
// place from geo.world_stage (the kFGlobal kernel instances).
constexpr size_t kMaxStagedLds = 64 * 1024;

// Whether the render kernels read world records from global memory; an error when the records
// need it and the scene carries no world_stage.
inline int global_records(const MjpbrScene* sc, bool* global) {
  *global = host_stage_bytes(sc) > kMaxStagedLds;
  if (*global && !sc->geo.world_stage)
    return set_error(MJPBR_EINVAL, "a world's records (%d geoms, %d lights) take %zu bytes, beyond the %zu staged in LDS: pass geo.world_stage",
                     sc->geo.ngeom, sc->lit.nlight, host_stage_bytes(sc), kMaxStagedLds);
  return MJPBR_OK;
}

// Sets tm's mesh-BVH cache (the top levels the render kernels stage in LDS after the world's
// records) for a scene whose records take `stage_bytes`; returns the bytes the cache adds.
inline size_t plan_mesh_cache(const MjpbrScene* sc, bool meshes, size_t stage_bytes, TileMap* tm) {
  tm->mesh_cache_nodes = tm->mesh_cache_n = 0;
  if (!meshes || stage_bytes >= kRenderLdsPerGroup) return 0;
  mesh_cache_shape(*sc, (int)(kRenderLdsPerGroup - stage_bytes), &tm->mesh_cache_nodes, &tm->mesh_cache_n);
  return (size_t)tm->mesh_cache_nodes * tm->mesh_cache_n * 48;
}

// Enqueues the per-world record resolution into sc->geo.world_stage (no-op when it is NULL);
// the render kernels launched after it on the stream copy the records into LDS.
int launch_stage_prep(const MjpbrScene* sc, hipStream_t stream);

inline int make_tile_map(const MjpbrScene* sc, TileMap* tm, int* nlaunch) {
  const MjpbrCameras& c = sc->cams;
  if (c.nrender < 1 || c.nrender > MJPBR_MAX_CAMS)
    return set_error(MJPBR_EINVAL, "nrender must be in [1, %d], got %d", MJPBR_MAX_CAMS, c.nrender);
  if (sc->geo.tlas_depth < 1 || sc->geo.tlas_depth > 12)
    return set_error(MJPBR_EINVAL, "scene BVH depth must be in [1, 12], got %d", sc->geo.tlas_depth);
  if (sc->geo.ndirect < 0 || sc->geo.ndirect > sc->geo.ngeom || (sc->geo.ndirect > 0 && !sc->geo.direct_geom))
    return set_error(MJPBR_EINVAL, "ndirect must be in [0, ngeom] with direct_geom set, got %d", sc->geo.ndirect);
  tm->nrender = c.nrender;
  tm->mesh_cache_nodes = tm->mesh_cache_n = 0;
  if (c.ray_base < 0) return set_error(MJPBR_EINVAL, "ray_base must not be negative, got %d", c.ray_base);
  int base = 0;
  int64_t rays = c.ray_base;
  for (int i = 0; i < c.nrender; ++i) {
    int w = c.cam_res[i][0], h = c.cam_res[i][1];
    if (w < 1 || h < 1) return set_error(MJPBR_EINVAL, "camera %d has resolution %dx%d", i, w, h);
    tm->tile_base[i] = base;
    tm->tiles_x[i] = (w + kTile - 1) / kTile;
    tm->ray_offset[i] = (int)std::min<int64_t>(rays, INT32_MAX);
    base += tm->tiles_x[i] * ((h + kTile - 1) / kTile);
    rays += (int64_t)w * h;
  }
  tm->tile_base[c.nrender] = base;
  // this launch's cameras end at or before the context's last pixel (at it when ray_base is 0 and
  // one launch renders every camera)
  if (rays > c.total_rays || (c.ray_base == 0 && rays != c.total_rays && c.nrender < MJPBR_MAX_CAMS))
    return set_error(MJPBR_EINVAL, "total_rays %d does not fit ray_base %d + the cameras' %lld pixels", c.total_rays, c.ray_base,
                     (long long)(rays - c.ray_base));
  tm->tiles_per_world = base;
  int64_t nb = (int64_t)base * sc->geo.nworld;
  if (nb > (1ll << 30)) return set_error(MJPBR_EINVAL, "too many tiles (%lld)", (long long)nb);
  tm->nblocks = (int)nb;
  int chunk = (tm->nblocks + 7) / 8;
  tm->xcd_chunk = chunk;
  *nlaunch = chunk * 8;
  return MJPBR_OK;
}

}  // namespace mjpbr
