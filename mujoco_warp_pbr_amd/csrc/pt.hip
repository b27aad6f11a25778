// pt.hip — path-traced tier megakernel (reference pathtrace.py:117-438).
//
// One thread per (world, pixel), `path_samples` paths each: next-event estimation of every light
// (bulb discs for soft shadows) and of the importance-sampled environment with power-heuristic
// MIS against BSDF sampling (GGX visible normals + cosine mixture), stochastic mirror and alpha,
// Russian roulette. Random numbers are counter-based per (sample, depth, slot), so the image is a
// deterministic function of the seed. With `denoise` the kernel writes the noisy HDR plus the
// first-hit features of the a-trous filter instead of RGB.
#pragma clang fp contract(off)
#include "abi_util.hpp"
#include "path.hpp"

namespace mjpbr {

// W waves per SIMD, from measurements at 1024 worlds: on scenes of more than 64 k mesh triangles
// 8 with spills (the mesh walks' dependent loads want the waves: the 200 k-triangle scene at 256^2,
// 1 sample, 182.5 ms against 189.9 at 6); with lighter meshes 6 (the full-feature scene at 128^2,
// 1 sample, 22.65 against 23.72 at 8; 6 and 8 both far ahead of 2); without meshes 4 (the
// reference suite's scenes at 16 samples, depth 4: material_grid 115.3 -> 94.5 ms, primitives
// 153.8 -> 118.4 ms against 8).
// kGlobal: the world's records are read in place from geo.world_stage (too large for the LDS).
// kCount: adds the rays traced to out.ray_counts (the same image).
constexpr int kHeavyMeshTriangles = 1 << 16;

template <bool kGlobal, bool kCount, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))
pt_kernel(MjpbrScene sc, MjpbrSettings s, MjpbrOutputs out, int seed, TileMap tm) {
  extern __shared__ float lds[];
  int b = logical_block(tm);
  if (b >= tm.nblocks) return;
  int w = b / tm.tiles_per_world;
  int t = b - w * tm.tiles_per_world;
  int cidx = 0;
  while (cidx + 1 < tm.nrender && t >= tm.tile_base[cidx + 1]) ++cidx;
  int tl = t - tm.tile_base[cidx];
  int tx = tl % tm.tiles_x[cidx], ty = tl / tm.tiles_x[cidx];
  int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int px = tx * kTile + (wave & 1) * 8 + (lane & 7);
  int py = ty * kTile + (wave >> 1) * 8 + (lane >> 3);
  SceneView sv = kGlobal ? global_world(sc, w, lds, tm.mesh_cache_nodes, tm.mesh_cache_n)
                        : stage_world(sc, w, lds, tm.mesh_cache_nodes, tm.mesh_cache_n);
  RayCount rays{0u, 0u, 0u};
  // the pixel's work; its early returns end here, where every lane of the wave is active again
  auto render_pixel = [&]() {
    Pixel pix;
    if (!setup_pixel(sc.cams, out, s, w, cidx, px, py, tm.ray_offset[cidx], &pix)) return;
    float exposure = sc.env.exposure[w % sc.env.exposure_n];
    PtCtx c{sc, s, sv, w, exposure, &rays};
    const int centre = kSubpixel / 2;
    if (!pix.want_rgb) {
      V3 o, d, dc;
      float md;
      primary_ray(pix, s, kSubpixel, centre, centre, &o, &d, &md, &dc);
      ++rays.segments;
      Hit hit = trace_closest(sv, o, d, md, s.cull_backfaces != 0);
      write_geometry(out, w, pix, hit, dc);
      return;
    }
    V3 head_dir = -V3{pix.xmat[2], pix.xmat[5], pix.xmat[8]};
    int pseed = seed + w * sc.cams.total_rays + pix.rayid;
    V3 hdr = v3s(0.0f), albedo_sum = v3s(0.0f), normal_sum = v3s(0.0f), emission_sum = v3s(0.0f);
    float first_depth = 0.0f, lum_sq = 0.0f;
    int nsample = s.path_samples;
    for (int smp = 0; smp < nsample; ++smp) {
      V3 o, d, dc;
      float md;
      path_primary(pix, s, pseed, smp, &o, &d, &md, &dc);
      Primary first;
      V3 radiance = trace_path(c, o, d, md, head_dir, pseed, smp, &first);
      if (smp == 0) {
        write_geometry(out, w, pix, first.hit, dc);
        if (s.render_guides) {
          int64_t a = (int64_t)w * out.guide_stride + out.rgb_adr[pix.camid] + pix.pixelid;
          st3(out.albedo + 3 * a, first.base_color);
          st3(out.normal + 3 * a, first.normal);
        }
        if (first.hit.geom_id != -1) first_depth = first.hit.dist * -dc.z;
      }
      hdr += radiance;
      albedo_sum += first.albedo;
      normal_sum += first.normal;
      emission_sum += first.emission;
      if (s.denoise) {
        float lum = luminance(radiance) * exposure;
        lum_sq += lum * lum;
      }
    }
    float fn = (float)nsample;
    hdr = hdr / fn;
    int64_t adr = out.rgb_adr[pix.camid] + pix.pixelid;
    if (s.denoise) {
      int64_t a = (int64_t)w * out.aov_stride + adr;
      st3(out.hdr + 3 * ((int64_t)w * out.hdr_stride + adr), hdr);
      V3 alb = albedo_sum / fn;
      out.aov_albedo[3 * a] = float_to_half(alb.x);
      out.aov_albedo[3 * a + 1] = float_to_half(alb.y);
      out.aov_albedo[3 * a + 2] = float_to_half(alb.z);
      st3(out.aov_emission + 3 * a, emission_sum * (exposure / fn));
      if (len_sq(normal_sum) > 0.0f) normal_sum = normalize(normal_sum);
      out.aov_normal[3 * a] = float_to_half(normal_sum.x);
      out.aov_normal[3 * a + 1] = float_to_half(normal_sum.y);
      out.aov_normal[3 * a + 2] = float_to_half(normal_sum.z);
      out.aov_depth[a] = first_depth;
      float mean_lum = luminance(hdr) * exposure;
      out.aov_variance[a] = fmaxf(lum_sq / fn - mean_lum * mean_lum, 0.0f) / fn;
      return;
    }
    if (s.render_hdr) st3(out.hdr + 3 * ((int64_t)w * out.hdr_stride + adr), hdr);
    out.rgb[(int64_t)w * out.rgb_stride + adr] = encode_display(display_transform(hdr * exposure, s.tone_mapper));
  };
  render_pixel();
  if (kCount) flush_ray_counts(out.ray_counts, rays);
}

}  // namespace mjpbr

using namespace mjpbr;

extern "C" int mjpbr_pt_render(const MjpbrScene* scene, const MjpbrSettings* s, MjpbrOutputs* out, int32_t seed,
                               void* stream) {
  TileMap tm;
  int nlaunch;
  int rc = make_tile_map(scene, &tm, &nlaunch);
  if (rc) return rc;
  if (s->path_samples < 1) return set_error(MJPBR_EINVAL, "path_samples must be at least 1");
  if (s->max_depth < 0 || s->max_depth + 4 * s->max_layers >= kJitterDepth)
    return set_error(MJPBR_EINVAL, "max_depth + 4 * max_layers must be below %d", kJitterDepth);
  bool global;
  if ((rc = global_records(scene, &global))) return rc;
  size_t lds = global ? 0 : host_stage_bytes(scene);
  lds += plan_mesh_cache(scene, s->has_meshes != 0, lds, &tm);
  if ((rc = launch_stage_prep(scene, (hipStream_t)stream))) return rc;
  using PtFn = void (*)(MjpbrScene, MjpbrSettings, MjpbrOutputs, int, TileMap);
  const bool count = out->ray_counts != nullptr;
  // [waves: 4 without meshes, 6 light meshes, 8 heavy meshes][global][count]
  static const PtFn kFns[3][2][2] = {
      {{pt_kernel<false, false, 4>, pt_kernel<false, true, 4>}, {pt_kernel<true, false, 4>, pt_kernel<true, true, 4>}},
      {{pt_kernel<false, false, 6>, pt_kernel<false, true, 6>}, {pt_kernel<true, false, 6>, pt_kernel<true, true, 6>}},
      {{pt_kernel<false, false, 8>, pt_kernel<false, true, 8>}, {pt_kernel<true, false, 8>, pt_kernel<true, true, 8>}}};
  const int occupancy = !s->has_meshes ? 0 : s->mesh_triangles > kHeavyMeshTriangles ? 2 : 1;
  const PtFn fn = kFns[occupancy][global][count];
  hipLaunchKernelGGL(fn, dim3(nlaunch), dim3(256), lds, (hipStream_t)stream, *scene, *s, *out, seed, tm);
  return check_launch("pt_kernel");
}
