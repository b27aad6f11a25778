// rt.hip — real-time tier megakernel (reference render.py:67-460).
//
// One workgroup = a 16 x 16 pixel tile of one world (a wave = an 8 x 8 tile); the world's scene
// BVH and geom frames are staged in LDS first. Each thread renders all N x N sub-pixel samples
// of its pixel: primary ray, split-sum IBL + direct lights with ray-traced (soft) shadows +
// emission, one mirror bounce, translucent layers, tone mapping and sRGB encoding.
#pragma clang fp contract(off)
#include <mutex>
#include "abi_util.hpp"
#include "rt_common.hpp"

namespace mjpbr {

// render.py:249-341. The camera-ray segments (translucent layers) and the mirror segment of
// each layer share one trace / surface / shading site: a segment loop with a mode flag instead of
// the reference's nested calls (shade_reflection), so the megakernel holds one inlined copy of
// the traversal and shading code. Arithmetic and random streams are those of the reference:
// the mirror segment traces from point + normal * RAY_OFFSET with max distance MJ_MAXVAL, is
// shaded with stream + 1 and blends as shaded * (1 - mirror) + reflected * mirror.
//
// The first sample of a pixel (`write_first`) writes depth, segmentation and the guides of its
// first hit as soon as they are known, so nothing of that hit stays live through shading.
template <unsigned F>
MJ_DEV V3 trace_pixel(const RtCtx& c, V3 origin, V3 dir, float max_dist, float footprint, float spread, V3 head_dir,
                      int stream, bool write_first, const MjpbrOutputs& out, const Pixel& pix, V3 dir_cam) {
  constexpr bool kMesh = F & kFMesh;
  const MjpbrSettings& s = c.s;
  const bool cull = s.cull_backfaces != 0;
  const bool layers = (F & kFLayer) && s.has_transparency;
  const bool reflect = (F & kFMirror) && s.has_reflection;
  V3 radiance = v3s(0.0f);
  float throughput = 1.0f, travelled = 0.0f;
  const int nlayer = layers ? s.max_layers : 1;
  int layer = 0;
  bool mirror_seg = false;
  // state of the layer whose mirror ray is being traced
  V3 layer_shaded = v3s(0.0f), layer_point = v3s(0.0f);
  float layer_mirror = 0.0f, layer_alpha = 1.0f, layer_dist = 0.0f, layer_cone = 0.0f;
  V3 ro = origin, rd = dir;
  float rmax = max_dist - travelled;
  for (;;) {
    Hit hit = trace_closest<kMesh>(c.sv, ro, rd, rmax, cull);
    const bool first = write_first && !mirror_seg && layer == 0;
    if (first) write_geometry(out, c.w, pix, hit, dir_cam);
    V3 shaded;
    V3 normal = v3s(0.0f);
    V3 point = v3s(0.0f);
    float cone = 0.0f;
    float alpha = 1.0f;
    if (hit.geom_id == -1) {
      if (first && s.render_guides) write_guides(out, c.w, pix, v3s(0.0f), v3s(0.0f));
      shaded = background(c, rd, spread);
      if (!mirror_seg) {
        radiance += shaded * throughput;
        break;
      }
    } else {
      point = ro + rd * hit.dist;
      cone = mirror_seg ? layer_cone + spread * hit.dist : footprint + spread * (travelled + hit.dist);
      Surface sf = eval_surface<kMesh, (F & kFTex) != 0>(c.sc, c.sv, s, c.w, hit, point, rd, cone, &normal);
      alpha = sf.alpha;
      if (first && s.render_guides) write_guides(out, c.w, pix, sf.base_color, normal);
      shaded = shade_surface<F>(c, sf, point, normal, -rd, head_dir, stream + (mirror_seg ? 1 : 0));
    }
    if (!mirror_seg) {
      float mirror = reflect && hit.geom_id >= 0 ? c.sv.surfs[hit.geom_id * kSurfRec + 12] : 0.0f;
      if (mirror > 0.0f) {
        layer_shaded = shaded;
        layer_point = point;
        layer_mirror = mirror;
        layer_alpha = alpha;
        layer_dist = hit.dist;
        layer_cone = cone;
        mirror_seg = true;
        ro = point + normal * kRayOffset;
        rd = dir - normal * (2.0f * dot(dir, normal));
        rmax = kMaxVal;
        continue;
      }
      layer_point = point;
      layer_alpha = alpha;
      layer_dist = hit.dist;
    } else {
      shaded = layer_shaded * (1.0f - layer_mirror) + shaded * layer_mirror;
      mirror_seg = false;
    }
    if (layers) {
      float opacity = clampf(layer_alpha, 0.0f, 1.0f);
      radiance += shaded * (throughput * opacity);
      throughput *= 1.0f - opacity;
      if (throughput < kMinThroughput) break;
      travelled += layer_dist;
      if (++layer >= nlayer) break;
      ro = layer_point + dir * kRayOffset;
      rd = dir;
      rmax = max_dist - travelled;
    } else {
      radiance += shaded;
      break;
    }
  }
  return radiance;
}

// Occupancy over registers: W waves per SIMD (W = 8: 64 VGPRs a lane) with spills to scratch
// beat fewer waves without, since waves mostly wait on dependent BVH, texture and spill loads;
// W is chosen per instance from measurements (rt_kernel instances table below).
//
// kFQuad instances (2x2 supersampling) run one lane a sample: four workgroups a 16x16 tile (8x8
// pixels each, on the tile's XCD), a wave a 4x4 pixel block with a pixel's four samples on
// adjacent lanes (sample = lane % 4). A wave then spans 16 pixels instead of 64, and the four
// lanes of a pixel trace nearly the same rays, so far fewer lanes idle in divergent branches
// (full-feature scene at 1024 x 128^2: 57.1 -> 43.4 ms at 8 waves a SIMD, 38.7 at 6; as
// 1024-thread workgroups a tile, 64.4 ms: a workgroup of 16 waves holds its compute unit's slots
// until its slowest wave ends). The
// pixel's sums are gathered on its first lane in sample order, as the per-pixel loop adds them,
// so the images are bit-identical.
template <unsigned F, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))
rt_kernel(MjpbrScene sc, MjpbrSettings s, MjpbrOutputs out, int seed, TileMap tm) {
  extern __shared__ float lds[];
  constexpr bool kQuad = (F & kFQuad) != 0;
  int b, quarter = 0;
  if constexpr (kQuad) {  // four workgroups a tile, on the tile's XCD: 8 x 8 pixels each
    const int xcd = blockIdx.x & 7, k4 = blockIdx.x >> 3;
    quarter = tm.xcd_chunk > 0 ? (k4 & 3) : (int)(blockIdx.x & 3);
    b = tm.xcd_chunk > 0 ? xcd * tm.xcd_chunk + (k4 >> 2) : (int)(blockIdx.x >> 2);
  } else {
    b = logical_block(tm);
  }
  if (b >= tm.nblocks) return;
  int w = b / tm.tiles_per_world;
  int t = b - w * tm.tiles_per_world;
  int cidx = 0;
  while (cidx + 1 < tm.nrender && t >= tm.tile_base[cidx + 1]) ++cidx;
  int tl = t - tm.tile_base[cidx];
  int tx = tl % tm.tiles_x[cidx], ty = tl / tm.tiles_x[cidx];
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int px, py;
  if constexpr (kQuad) {
    const int p = lane >> 2;  // pixel of the wave's 4x4 block; lane & 3 is the sample
    px = tx * kTile + (quarter & 1) * 8 + (wave & 1) * 4 + (p & 3);
    py = ty * kTile + (quarter >> 1) * 8 + (wave >> 1) * 4 + (p >> 2);
  } else {
    px = tx * kTile + (wave & 1) * 8 + (lane & 7);
    py = ty * kTile + (wave >> 1) * 8 + (lane >> 3);
  }

  SceneView sv = (F & kFGlobal) ? global_world(sc, w, lds, tm.mesh_cache_nodes, tm.mesh_cache_n)
                                 : stage_world(sc, w, lds, tm.mesh_cache_nodes, tm.mesh_cache_n);

  Pixel pix;
  if (!setup_pixel(sc.cams, out, s, w, cidx, px, py, tm.ray_offset[cidx], &pix)) return;
  float exposure = sc.env.exposure[w % sc.env.exposure_n];
  RtCtx c{sc, s, sv, w, 0, exposure};
  bool cull = s.cull_backfaces != 0;
  if (!pix.want_rgb) {
    if (kQuad && (lane & 3) != 0) return;
    V3 o, d, dc;
    float md;
    primary_ray(pix, s, 1, 0, 0, &o, &d, &md, &dc);
    Hit hit = trace_closest<(F & kFMesh) != 0, F == 0u>(sv, o, d, md, cull);
    write_geometry(out, w, pix, hit, dc);
    return;
  }
  const int n_aa = (F & kFSS) ? s.samples_per_axis : 1;
  float inv_aa = 1.0f / (float)n_aa;
  float pspread, pwidth;
  pixel_cone(pix, &pspread, &pwidth);
  float spread = pspread * inv_aa, width = pwidth * inv_aa;
  V3 head_dir = -V3{pix.xmat[2], pix.xmat[5], pix.xmat[8]};
  c.seed = seed + w * sc.cams.total_rays + pix.rayid;
  V3 display = v3s(0.0f), hdr = v3s(0.0f);
  if constexpr (kQuad) {  // the launch guarantees samples_per_axis == 2
    const int smp = lane & 3;
    V3 o, d, dc;
    float md;
    primary_ray(pix, s, 2, smp & 1, smp >> 1, &o, &d, &md, &dc);
    const V3 radiance = trace_pixel<F>(c, o, d, md, width, spread, head_dir, 2 * smp, smp == 0, out, pix, dc);
    const V3 shown = display_transform(radiance * exposure, s.tone_mapper);
    // a pixel's four lanes are all here (they return together): its sums in sample order
    const int first = lane & ~3;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      display += V3{pull_float(first + k, shown.x), pull_float(first + k, shown.y), pull_float(first + k, shown.z)};
      hdr += V3{pull_float(first + k, radiance.x), pull_float(first + k, radiance.y), pull_float(first + k, radiance.z)};
    }
    if (smp != 0) return;
  } else {
    for (int smp = 0; smp < n_aa * n_aa; ++smp) {
      V3 o, d, dc;
      float md;
      primary_ray(pix, s, n_aa, smp % n_aa, smp / n_aa, &o, &d, &md, &dc);
      V3 radiance = trace_pixel<F>(c, o, d, md, width, spread, head_dir, 2 * smp, smp == 0, out, pix, dc);
      display += display_transform(radiance * exposure, s.tone_mapper);
      hdr += radiance;
    }
  }
  float inv2 = inv_aa * inv_aa;
  int64_t adr = out.rgb_adr[pix.camid] + pix.pixelid;
  if (s.render_hdr) st3(out.hdr + 3 * ((int64_t)w * out.hdr_stride + adr), hdr * inv2);
  out.rgb[(int64_t)w * out.rgb_stride + adr] = encode_display(display * inv2);
}

}  // namespace mjpbr

using namespace mjpbr;

namespace {

using RtKernelFn = void (*)(MjpbrScene, MjpbrSettings, MjpbrOutputs, int, TileMap);

struct RtInstance {
  unsigned mask;
  RtKernelFn fn;
};

// Compiled instances, each launched for the settings whose required features it covers, at the
// occupancy measured best for it (waves per SIMD at 1024 worlds x 128^2: the reflective-plane
// textured scene runs 6 % faster at 6 than at 8 and 3 % slower at 5; the leanest instance 2 %
// faster at 7 than at 8 or 6; the full-feature one with its per-pixel sample loop 1 % slower at 7
// than at 8, with one lane a sample 38.7 ms at 6 against 38.8 at 7, 44.3 at 8, 39.3 at 5: its
// lanes diverge less, so fewer spills beat more waves). The textured mesh
// instance runs at 6: compiled for 8 (64 VGPRs, about 270 bytes a lane of spills) it shaded mesh
// hits wrongly after a source change in code it never executes, while the same source at 6 or 4
// waves matches the CPU oracle - its spill code at 8 is not trusted.
#define MJPBR_RT_INSTANCE(F, W) RtInstance{F, rt_kernel<F, W>}
const RtInstance kRtInstances[] = {
    MJPBR_RT_INSTANCE(0u, 7),
    MJPBR_RT_INSTANCE(kFTex, 8),
    MJPBR_RT_INSTANCE(kFMirror | kFTex, 6),
    MJPBR_RT_INSTANCE(kFMesh, 8),
    MJPBR_RT_INSTANCE(kFMesh | kFTex, 6),
    MJPBR_RT_INSTANCE(kFMesh | kFSS | kFSoft | kFMirror | kFTex, 8),
    MJPBR_RT_INSTANCE(kFAll, 8),
    MJPBR_RT_INSTANCE(kFAll | kFGlobal, 8),
    MJPBR_RT_INSTANCE(kFMesh | kFSS | kFSoft | kFMirror | kFTex | kFQuad, 6),
    MJPBR_RT_INSTANCE(kFAll | kFQuad, 6),
    MJPBR_RT_INSTANCE(kFAll | kFGlobal | kFQuad, 6),
};
#undef MJPBR_RT_INSTANCE

unsigned required_features(const MjpbrSettings& s) {
  unsigned f = 0;
  if (s.has_meshes) f |= kFMesh;
  if (s.samples_per_axis > 1) f |= kFSS;
  if (s.samples_per_axis == 2) f |= kFQuad;
  if (s.use_shadows && s.shadow_samples > 1) f |= kFSoft;
  if (s.has_reflection) f |= kFMirror;
  if (s.has_transparency) f |= kFLayer;
  if (s.textures) f |= kFTex;
  return f;
}

// The instance with the fewest features among those covering `need` (kFAll covers everything) and
// of the launch form it asks for: one lane a sample exactly when samples_per_axis is 2.
const RtInstance& pick_instance(unsigned need) {
  const RtInstance* best = nullptr;
  for (const RtInstance& k : kRtInstances)
    if ((k.mask & need) == need && (k.mask & kFQuad) == (need & kFQuad) &&
        (!best || __builtin_popcount(k.mask) < __builtin_popcount(best->mask)))
      best = &k;
  return *best;
}

}  // namespace

extern "C" int mjpbr_rt_kernel_features(const MjpbrSettings* s) {
  if (!s) return set_error(MJPBR_EINVAL, "settings must not be NULL");
  return (int)pick_instance(required_features(*s)).mask;
}

extern "C" int mjpbr_rt_render(const MjpbrScene* scene, const MjpbrSettings* s, MjpbrOutputs* out, int32_t seed,
                               void* stream) {
  TileMap tm;
  int nlaunch;
  int rc = make_tile_map(scene, &tm, &nlaunch);
  if (rc) return rc;
  if (s->samples_per_axis < 1) return set_error(MJPBR_EINVAL, "samples_per_axis must be at least 1");
  bool global;
  if ((rc = global_records(scene, &global))) return rc;
  size_t lds = global ? 0 : host_stage_bytes(scene);
  const RtInstance& k = pick_instance(required_features(*s) | (global ? kFGlobal : 0u));
  // mesh-BVH top levels in the LDS left to a workgroup at 8 per compute unit (the instances with
  // meshes also hold the waves' mesh job tables, static LDS)
  const size_t job_lds = (k.mask & kFMesh) ? (size_t)kJobTableBytes : 0;
  if ((k.mask & kFMesh) && scene->geo.ngeom >= (1 << kPendShift))
    return set_error(MJPBR_EINVAL, "ngeom must be below %d, got %d", 1 << kPendShift, scene->geo.ngeom);
  lds += plan_mesh_cache(scene, (k.mask & kFMesh) && s->has_meshes, lds + job_lds, &tm);
  const int64_t grid = (int64_t)nlaunch * ((k.mask & kFQuad) ? 4 : 1);
  if (grid > INT32_MAX) return set_error(MJPBR_EINVAL, "too many workgroups (%lld)", (long long)grid);
  if ((rc = launch_stage_prep(scene, (hipStream_t)stream))) return rc;
  hipLaunchKernelGGL(k.fn, dim3((unsigned)grid), dim3(256), lds, (hipStream_t)stream, *scene, *s, *out, seed, tm);
  return check_launch("rt_kernel");
}

// ---------------------------------------------------------------- allocation-free variant
// The per-world staging buffer of mjpbr_rt_render_noalloc: a file-scope device array, so the
// entry point needs no world_stage from the caller and allocates nothing. Each GPU the library
// runs on holds its own copy (MJPBR_NOALLOC_STAGE_BYTES) until the process ends. In host code
// the name `scratch` is not a device address: the address of this device's copy comes from
// hipGetSymbolAddress, once per device, and only that address is handed to kernels.
static __device__ __align__(256) float scratch[(size_t)MJPBR_NOALLOC_STAGE_BYTES / sizeof(float)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};

// Address of the current device's scratch (looked up on the first call on that device).
int scratch_address(float** ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(MJPBR_ELAUNCH, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return set_error(MJPBR_EINVAL, "device %d beyond the %d the library tracks", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return set_error(MJPBR_ELAUNCH, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return MJPBR_OK;
}

}  // namespace

extern "C" int mjpbr_rt_render_noalloc(const MjpbrScene* scene, const MjpbrSettings* s, MjpbrOutputs* out, int32_t seed,
                                       void* stream) {
  if (!scene || !s || !out) return set_error(MJPBR_EINVAL, "scene, settings and outputs must not be NULL");
  const int64_t floats = mjpbr_world_stage_floats(scene);
  if (floats < 0) return (int)floats;
  if (scene->geo.nworld < 1) return set_error(MJPBR_EINVAL, "nworld must be at least 1, got %d", scene->geo.nworld);
  const int64_t bytes = floats * (int64_t)sizeof(float) * scene->geo.nworld;
  if (bytes > (int64_t)MJPBR_NOALLOC_STAGE_BYTES)
    return set_error(MJPBR_EINVAL, "%d worlds need %lld bytes of staging; the allocation-free entry point holds %lld (use mjpbr_rt_render with a world_stage buffer)",
                     scene->geo.nworld, (long long)bytes, (long long)MJPBR_NOALLOC_STAGE_BYTES);
  float* ptr = nullptr;
  int rc = scratch_address(&ptr);
  if (rc) return rc;
  // Every world's block is rewritten by the staging launch before the render kernel reads it on
  // the same stream, so nothing needs clearing between calls.
  MjpbrScene sc = *scene;
  sc.geo.world_stage = ptr;
  return mjpbr_rt_render(&sc, s, out, seed, stream);
}
