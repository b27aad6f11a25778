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
#include "sample.hpp"
#include "stage.hpp"

namespace mjpbr {

constexpr int kSubpixel = 15;  // pathtrace.py:65
constexpr int kSlots = 16, kDepths = 64, kJitterDepth = kDepths - 1;
constexpr int kSlotLights = 4, kSlottedLights = 4;
constexpr int kExtraLights = 1 << 30;
constexpr float kRRMin = 0.05f, kRRMax = 0.95f;

MJ_DEV float prand(int seed, int sample, int depth, int slot) {
  return rand_at(seed, (sample * kDepths + depth) * kSlots + slot);
}

MJ_DEV V2 light_rand2(int seed, int sample, int depth, int nlight, int li) {
  if (li < kSlottedLights) {
    int slot = kSlotLights + 2 * li;
    return V2{prand(seed, sample, depth, slot), prand(seed, sample, depth, slot + 1)};
  }
  int counter = kExtraLights + ((sample * kDepths + depth) * nlight + li) * 2;
  return V2{rand_at(seed, counter), rand_at(seed, counter + 1)};
}

MJ_DEV float power_heuristic(float a, float b) {
  float a2 = a * a;
  return a2 / fmaxf(a2 + b * b, 1.0e-30f);
}

struct PtCtx {
  const MjpbrScene& sc;
  const MjpbrSettings& s;
  const SceneView& sv;
  int w;
  float exposure;
};

struct Primary {
  Hit hit;
  V3 albedo, normal, emission, base_color;
};

// pathtrace.py:128-162
MJ_DEV V3 light_sample(const PtCtx& c, Light L, const Lobes& lb, V3 point, V3 normal, V3 t1, V3 t2, V3 v, V2 u) {
  if (L.type != MJPBR_LIGHT_DIRECTIONAL && L.bulbradius > 0.0f) {
    V3 b1, b2;
    orthonormal_basis(normalize(point - L.pos), &b1, &b2);
    float r = L.bulbradius * sqrtf(u.x);
    float sp, cp;
    sincos_turns(u.y, &sp, &cp);
    L.pos = L.pos + (b1 * cp + b2 * sp) * r;
  }
  V3 l, ds, ss;
  float dist;
  eval_light(L, c.s.light_model, c.s.has_spot != 0, point, &l, &dist, &ds, &ss);
  V3 ll = V3{dot(l, t1), dot(l, t2), dot(l, normal)};
  if (ll.z <= 0.0f || (len_sq(ds) == 0.0f && len_sq(ss) == 0.0f)) return v3s(0.0f);
  if (c.s.use_shadows && L.castshadow) {
    float max_t = L.type != MJPBR_LIGHT_DIRECTIONAL ? dist - kShadowGap : kUnboundedShadow;
    if (trace_any(c.sv, point + normal * kRayOffset, l, max_t, c.s.cull_backfaces != 0)) return v3s(0.0f);
  }
  V3 diffuse = eval_bsdf_cos(lb.diffuse_color, v3s(0.0f), lb.alpha, v3s(0.0f), v, ll);
  V3 specular = eval_bsdf_cos(v3s(0.0f), lb.f0, lb.alpha, lb.energy_comp, v, ll);
  return cw(diffuse, ds) + cw(specular, ss);
}

// pathtrace.py:164-192
MJ_DEV V3 direct_lights(const PtCtx& c, const Lobes& lb, V3 point, V3 normal, V3 t1, V3 t2, V3 v, V3 head_dir, int seed,
                        int sample, int depth) {
  V3 radiance = v3s(0.0f);
  const MjpbrLights& lt = c.sc.lit;
  for (int li = 0; li < lt.nlight; ++li) {
    if (!bat(lt.light_active, lt.light_active_n, c.w, lt.nlight)[li]) continue;
    V2 u = light_rand2(seed, sample, depth, lt.nlight, li);
    radiance += light_sample(c, fetch_light(lt, c.w, li), lb, point, normal, t1, t2, v, u);
  }
  if (c.s.headlight_active) {
    Light H = headlight(c.s, c.s.headlight_intensity / c.exposure, head_dir);
    radiance += light_sample(c, H, lb, point, normal, t1, t2, v, V2{0.0f, 0.0f});
  }
  return radiance;
}

MJ_DEV V3 environment(const PtCtx& c, V3 dir) {
  int e;
  float yaw, scale;
  env_lookup(c.sc.env, c.w, &e, &yaw, &scale);
  return sample_radiance(c.sc.env, e, dir, yaw, 0.0f) * scale;
}

// pathtrace.py:200-338
MJ_DEV V3 trace_path(const PtCtx& c, V3 origin, V3 dir, float max_dist, V3 head_dir, int seed, int sample,
                     Primary* first) {
  const MjpbrSettings& s = c.s;
  const MjpbrEnvironment& env = c.sc.env;
  bool cull = s.cull_backfaces != 0;
  V3 ambient = ld3(env.ambient + 3 * (c.w % env.ambient_n));
  V3 radiance = v3s(0.0f), beta = v3s(1.0f);
  first->hit.geom_id = -1;
  first->hit.dist = 0.0f;
  first->albedo = v3s(1.0f);
  first->normal = v3s(0.0f);
  first->emission = v3s(0.0f);
  first->base_color = v3s(0.0f);
  bool camera_ray = true, delta_bounce = true;
  float pdf_prev = 0.0f;
  int depth = 0, passes = 0;
  const float footprint = 0.0f;
  while (depth <= s.max_depth) {
    Hit hit = trace_closest(c.sv, origin, dir, max_dist, cull);
    if (camera_ray && passes == 0) first->hit = hit;
    if (hit.geom_id == -1) {
      if (camera_ray) {
        if (s.env_background)
          radiance += cw(beta, environment(c, dir));
        else
          radiance += cw(beta, ld3(env.background) / c.exposure);
      } else {
        V3 env_light = ambient;
        if (s.has_env) {
          float weight = 1.0f;
          if (!delta_bounce) {
            int e;
            float yaw, scale;
            env_lookup(env, c.w, &e, &yaw, &scale);
            weight = power_heuristic(pdf_prev, env_direction_pdf(env, e, yaw, dir));
          }
          env_light = env_light + environment(c, dir) * weight;
        }
        radiance += cw(beta, env_light);
      }
      break;
    }
    V3 point = origin + dir * hit.dist;
    V3 normal;
    Surface sf = eval_surface(c.sc, s, c.w, hit, point, dir, footprint, &normal);
    if (s.has_transparency) {
      if (prand(seed, sample, depth + passes, 0) >= sf.alpha && passes < 4 * s.max_layers) {
        origin = point + dir * kRayOffset;
        max_dist = max_dist - hit.dist;
        ++passes;
        continue;
      }
    }
    radiance += cw(beta, sf.emissive / c.exposure);
    if (depth == s.max_depth) break;
    max_dist = kMaxVal;
    if (camera_ray) {
      first->albedo = sf.base_color;
      first->normal = normal;
      first->base_color = sf.base_color;
      first->emission = sf.emissive / c.exposure;
    }
    camera_ray = false;
    if (s.has_reflection) {
      if (prand(seed, sample, depth, 1) < mirror_reflectance(c.sc, c.w, hit.geom_id)) {
        dir = dir - normal * (2.0f * dot(dir, normal));
        origin = point + normal * kRayOffset;
        delta_bounce = true;
        ++depth;
        continue;
      }
    }
    V3 t1, t2;
    orthonormal_basis(normal, &t1, &t2);
    V3 view = -dir;
    V3 v = V3{dot(view, t1), dot(view, t2), dot(view, normal)};
    float nov = fmaxf(v.z, 1.0e-4f);
    Lobes lb = shading_lobes(env, sf, nov);
    float alpha = lb.alpha;
    V3 spec_weight = cw(lb.spec_albedo, lb.energy_comp);
    float p_spec = specular_probability(lb.diffuse_color, spec_weight);
    if (depth == 0) first->albedo = lb.diffuse_color + spec_weight;
    radiance += cw(beta, direct_lights(c, lb, point, normal, t1, t2, v, head_dir, seed, sample, depth));
    if (s.has_env) {
      int e;
      float yaw, scale;
      env_lookup(env, c.w, &e, &yaw, &scale);
      V2 u_env = V2{prand(seed, sample, depth, 2), prand(seed, sample, depth, 3)};
      float pdf_env;
      V3 d_env = env_sample_direction(env, e, yaw, u_env, &pdf_env);
      V3 l_env = V3{dot(d_env, t1), dot(d_env, t2), dot(d_env, normal)};
      if (l_env.z > 0.0f && pdf_env > 0.0f) {
        V3 f = eval_bsdf_cos(lb.diffuse_color, lb.f0, alpha, lb.energy_comp, v, l_env);
        if (!trace_any(c.sv, point + normal * kRayOffset, d_env, kUnboundedShadow, cull)) {
          float weight = power_heuristic(pdf_env, pdf_bsdf(alpha, p_spec, v, l_env));
          radiance += cw(beta, cw(f, environment(c, d_env))) * (weight / pdf_env);
        }
      }
    }
    V2 u = V2{prand(seed, sample, depth, 13), prand(seed, sample, depth, 14)};
    V3 l = sample_bsdf(alpha, p_spec, v, prand(seed, sample, depth, 12), u);
    if (l.z <= 0.0f) break;
    float pdf = pdf_bsdf(alpha, p_spec, v, l);
    if (pdf <= 0.0f) break;
    beta = cw(beta, eval_bsdf_cos(lb.diffuse_color, lb.f0, alpha, lb.energy_comp, v, l) / pdf);
    dir = normalize(t1 * l.x + t2 * l.y + normal * l.z);
    origin = point + normal * kRayOffset;
    pdf_prev = pdf;
    delta_bounce = false;
    ++depth;
    if (depth >= s.rr_depth) {
      float q = clampf(fmaxf(beta.x, fmaxf(beta.y, beta.z)), kRRMin, kRRMax);
      if (prand(seed, sample, depth, 15) >= q) break;
      beta = beta / q;
    }
  }
  return radiance;
}

// 8 waves per SIMD with spills beat 2 without: the path loop is latency bound (measured 1.8x on
// the full-feature scene at 1024 x 128^2, 1 sample).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
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
  SceneView sv = stage_world(sc, w, lds);
  Pixel pix;
  if (!setup_pixel(sc.cams, out, s, w, cidx, px, py, tm.ray_offset[cidx], &pix)) return;
  float exposure = sc.env.exposure[w % sc.env.exposure_n];
  PtCtx c{sc, s, sv, w, exposure};
  const int centre = kSubpixel / 2;
  if (!pix.want_rgb) {
    V3 o, d, dc;
    float md;
    primary_ray(pix, s, kSubpixel, centre, centre, &o, &d, &md, &dc);
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
    int sx = centre, sy = centre;
    if (smp > 0) {
      sx = (int)(prand(pseed, smp, kJitterDepth, 0) * (float)kSubpixel);
      sy = (int)(prand(pseed, smp, kJitterDepth, 1) * (float)kSubpixel);
    }
    V3 o, d, dc;
    float md;
    primary_ray(pix, s, kSubpixel, sx, sy, &o, &d, &md, &dc);
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
  size_t lds = host_stage_bytes(&scene->geo);
  if (lds > 64 * 1024) return set_error(MJPBR_EINVAL, "scene too large for LDS staging (%d geoms)", scene->geo.ngeom);
  hipLaunchKernelGGL(pt_kernel, dim3(nlaunch), dim3(256), lds, (hipStream_t)stream, *scene, *s, *out, seed, tm);
  return check_launch("pt_kernel");
}
