// path.hpp — one path of the path-traced tier as a resumable state machine
// (reference pathtrace.py:117-338).
//
// Next-event estimation of every light (bulb discs for soft shadows) and of the
// importance-sampled environment with power-heuristic MIS against BSDF sampling (GGX visible
// normals + cosine mixture), stochastic mirror and alpha, Russian roulette. Random numbers are
// counter-based per (sample, depth, slot), so a path is a deterministic function of the pixel
// seed and its sample index, whichever kernel schedules it.
//
// `path_step` runs one segment of the bounce loop of pathtrace.py:220-338 on a PathState. The
// per-pixel megakernel loops it to completion per sample; the persistent wavefront kernel
// interleaves paths of many pixels on one wave and refills lanes whose path ended. Both run the
// same arithmetic, so their images are bit-identical.
#pragma once
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

// Rays one lane traced. The counting kernel instances add them up (flush_ray_counts); elsewhere
// the counts are never read and the compiler drops them.
struct RayCount {
  uint32_t segments, lights, env;
};

struct PtCtx {
  const MjpbrScene& sc;
  const MjpbrSettings& s;
  const SceneView& sv;
  int w;
  float exposure;
  RayCount* rays;
};

// Adds the wave's counts to counts[0..2] (segments, light shadow rays, environment shadow rays).
// Every lane of the wave must be active.
MJ_DEV void flush_ray_counts(uint64_t* counts, const RayCount& r) {
  const uint32_t v[3] = {r.segments, r.lights, r.env};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    unsigned long long x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(reinterpret_cast<unsigned long long*>(counts) + k, x);
  }
}

// First-hit record of a path (pathtrace.py:98-108).
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
    ++c.rays->lights;
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
    if (!staged_light_active(c.sv.lights, li)) continue;
    V2 u = light_rand2(seed, sample, depth, lt.nlight, li);
    radiance += light_sample(c, staged_light(c.sv.lights, li), lb, point, normal, t1, t2, v, u);
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

// State of one path between segments (the locals of pathtrace.py:200-338).
struct PathState {
  V3 origin, dir;
  float max_dist;
  V3 radiance, beta;
  float pdf_prev;
  int depth, passes;
  bool camera_ray, delta_bounce;
};

MJ_DEV void path_begin(PathState* st, Primary* first, V3 origin, V3 dir, float max_dist) {
  st->origin = origin;
  st->dir = dir;
  st->max_dist = max_dist;
  st->radiance = v3s(0.0f);
  st->beta = v3s(1.0f);
  st->pdf_prev = 0.0f;
  st->depth = 0;
  st->passes = 0;
  st->camera_ray = true;
  st->delta_bounce = true;
  first->hit.geom_id = -1;
  first->hit.dist = 0.0f;
  first->albedo = v3s(1.0f);
  first->normal = v3s(0.0f);
  first->emission = v3s(0.0f);
  first->base_color = v3s(0.0f);
}

// One segment of the bounce loop (pathtrace.py:220-338): returns false once the path ended, with
// its radiance in st->radiance. Requires st->depth <= max_depth on entry.
MJ_DEV bool path_step(const PtCtx& c, PathState* st, V3 head_dir, int seed, int sample, Primary* first) {
  const MjpbrSettings& s = c.s;
  const MjpbrEnvironment& env = c.sc.env;
  const bool cull = s.cull_backfaces != 0;
  const float footprint = 0.0f;
  ++c.rays->segments;
  Hit hit = trace_closest(c.sv, st->origin, st->dir, st->max_dist, cull);
  if (st->camera_ray && st->passes == 0) first->hit = hit;
  if (hit.geom_id == -1) {
    if (st->camera_ray) {
      if (s.env_background)
        st->radiance += cw(st->beta, environment(c, st->dir));
      else
        st->radiance += cw(st->beta, ld3(env.background) / c.exposure);
    } else {
      V3 env_light = ld3(env.ambient + 3 * (c.w % env.ambient_n));
      if (s.has_env) {
        float weight = 1.0f;
        if (!st->delta_bounce) {
          int e;
          float yaw, scale;
          env_lookup(env, c.w, &e, &yaw, &scale);
          weight = power_heuristic(st->pdf_prev, env_direction_pdf(env, e, yaw, st->dir));
        }
        env_light = env_light + environment(c, st->dir) * weight;
      }
      st->radiance += cw(st->beta, env_light);
    }
    return false;
  }
  const int depth = st->depth;
  V3 point = st->origin + st->dir * hit.dist;
  V3 normal;
  Surface sf = eval_surface(c.sc, c.sv, s, c.w, hit, point, st->dir, footprint, &normal);
  if (s.has_transparency) {
    if (prand(seed, sample, depth + st->passes, 0) >= sf.alpha && st->passes < 4 * s.max_layers) {
      st->origin = point + st->dir * kRayOffset;
      st->max_dist = st->max_dist - hit.dist;
      ++st->passes;
      return true;
    }
  }
  st->radiance += cw(st->beta, sf.emissive / c.exposure);
  if (depth == s.max_depth) return false;
  st->max_dist = kMaxVal;
  if (st->camera_ray) {
    first->albedo = sf.base_color;
    first->normal = normal;
    first->base_color = sf.base_color;
    first->emission = sf.emissive / c.exposure;
  }
  st->camera_ray = false;
  if (s.has_reflection) {
    if (prand(seed, sample, depth, 1) < c.sv.surfs[hit.geom_id * kSurfRec + 12]) {
      st->dir = st->dir - normal * (2.0f * dot(st->dir, normal));
      st->origin = point + normal * kRayOffset;
      st->delta_bounce = true;
      ++st->depth;
      return st->depth <= s.max_depth;
    }
  }
  V3 t1, t2;
  orthonormal_basis(normal, &t1, &t2);
  V3 view = -st->dir;
  V3 v = V3{dot(view, t1), dot(view, t2), dot(view, normal)};
  float nov = fmaxf(v.z, 1.0e-4f);
  Lobes lb = shading_lobes(env, sf, nov);
  float alpha = lb.alpha;
  V3 spec_weight = cw(lb.spec_albedo, lb.energy_comp);
  float p_spec = specular_probability(lb.diffuse_color, spec_weight);
  if (depth == 0) first->albedo = lb.diffuse_color + spec_weight;
  st->radiance += cw(st->beta, direct_lights(c, lb, point, normal, t1, t2, v, head_dir, seed, sample, depth));
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
      ++c.rays->env;
      if (!trace_any(c.sv, point + normal * kRayOffset, d_env, kUnboundedShadow, cull)) {
        float weight = power_heuristic(pdf_env, pdf_bsdf(alpha, p_spec, v, l_env));
        st->radiance += cw(st->beta, cw(f, environment(c, d_env))) * (weight / pdf_env);
      }
    }
  }
  V2 u = V2{prand(seed, sample, depth, 13), prand(seed, sample, depth, 14)};
  V3 l = sample_bsdf(alpha, p_spec, v, prand(seed, sample, depth, 12), u);
  if (l.z <= 0.0f) return false;
  float pdf = pdf_bsdf(alpha, p_spec, v, l);
  if (pdf <= 0.0f) return false;
  st->beta = cw(st->beta, eval_bsdf_cos(lb.diffuse_color, lb.f0, alpha, lb.energy_comp, v, l) / pdf);
  st->dir = normalize(t1 * l.x + t2 * l.y + normal * l.z);
  st->origin = point + normal * kRayOffset;
  st->pdf_prev = pdf;
  st->delta_bounce = false;
  ++st->depth;
  if (st->depth >= s.rr_depth) {
    float q = clampf(fmaxf(st->beta.x, fmaxf(st->beta.y, st->beta.z)), kRRMin, kRRMax);
    if (prand(seed, sample, st->depth, 15) >= q) return false;
    st->beta = st->beta / q;
  }
  return st->depth <= s.max_depth;
}

// A whole path (pathtrace.py:200-338).
MJ_DEV V3 trace_path(const PtCtx& c, V3 origin, V3 dir, float max_dist, V3 head_dir, int seed, int sample, Primary* first) {
  PathState st;
  path_begin(&st, first, origin, dir, max_dist);
  while (path_step(c, &st, head_dir, seed, sample, first)) {
  }
  return st.radiance;
}

// Primary ray of sample `smp` of a pixel: the pixel centre of the kSubpixel^2 grid for sample 0,
// a jittered cell otherwise (pathtrace.py:360-375).
MJ_DEV void path_primary(const Pixel& pix, const MjpbrSettings& s, int pseed, int smp, V3* o, V3* d, float* md, V3* dc) {
  const int centre = kSubpixel / 2;
  int sx = centre, sy = centre;
  if (smp > 0) {
    sx = (int)(prand(pseed, smp, kJitterDepth, 0) * (float)kSubpixel);
    sy = (int)(prand(pseed, smp, kJitterDepth, 1) * (float)kSubpixel);
  }
  primary_ray(pix, s, kSubpixel, sx, sy, o, d, md, dc);
}

}  // namespace mjpbr
