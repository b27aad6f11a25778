// rt_common.hpp — shading pieces of the real-time tier (reference render.py:67-247): shadow
// rays, direct lights, split-sum IBL, background, feature mask of the kernel instances.
#pragma once
#pragma clang fp contract(off)
#include "stage.hpp"

namespace mjpbr {

constexpr float kMinThroughput = 1.0f / 255.0f;  // render.py:62

// Compile-time feature mask of a kernel instance (the reference bakes the same choices into its
// kernel key, kernel_inputs.py:321-337 and io.py:300-330). A bit that is clear removes code
// whose runtime condition is false for every launch the instance is picked for:
//   kFMesh    mesh / SDF / height-field geoms are rendered (MjpbrSettings.has_meshes)
//   kFSS      samples_per_axis > 1
//   kFSoft    shadows with shadow_samples > 1
//   kFMirror  has_reflection
//   kFLayer   has_transparency
//   kFTex     textures
// and two launch forms: kFGlobal (world records read in place from geo.world_stage) and kFQuad
// (samples_per_axis == 2: one lane a sample, a pixel's four samples on adjacent lanes, four
// workgroups a 16x16 tile).
enum : unsigned { kFMesh = 1, kFSS = 2, kFSoft = 4, kFMirror = 8, kFLayer = 16, kFTex = 32, kFAll = 63, kFGlobal = 64, kFQuad = 128 };

struct RtCtx {
  const MjpbrScene& sc;
  const MjpbrSettings& s;
  const SceneView& sv;
  int w;
  int seed;
  float exposure;
};

// render.py:67-75
MJ_DEV V2 shadow_rand2(int seed, int stream, int nlight, int li, int nsample, int k) {
  int counter = ((stream * nlight + li) * nsample + k) * 2;
  return V2{rand_at(seed, counter), rand_at(seed, counter + 1)};
}

// render.py:95-127. Directional lights and lights sampled once take one any-hit ray; soft shadows
// trace `shadow_samples` rays to points of the bulb's disc, four at a time as one packet query
// (they share the origin). With one ray, v is 0 or 1 and v + (1 - v) * fraction equals the
// reference's `hit ? fraction : 1` exactly; `visible` counts unoccluded rays, an exact sum.
constexpr int kShadowPacket = 4;

template <unsigned F>
MJ_DEV float light_visibility(const RtCtx& c, const Light& L, V3 point, V3 normal, V3 l, float dist, int stream,
                              int nlight, int li) {
  constexpr bool kMesh = F & kFMesh;
  const MjpbrSettings& s = c.s;
  V3 origin = point + normal * kRayOffset;
  bool cull = s.cull_backfaces != 0;
  const bool directional = L.type == MJPBR_LIGHT_DIRECTIONAL;
  const bool soft = (F & kFSoft) && !directional && s.shadow_samples > 1;
  if (!soft && !kMesh) {
    const bool hit = trace_any<kMesh, F == 0u>(c.sv, origin, l, directional ? kUnboundedShadow : dist - kShadowGap, cull);
    return hit ? s.shadow_light_fraction : 1.0f;
  }
  // One packet call site for both cases (with meshes, one ray goes through it too, so its mesh
  // tests are shared out over the wave's lanes as well).
  const int n = soft ? s.shadow_samples : 1;
  V3 t1 = v3s(0.0f), t2 = v3s(0.0f);
  if (soft) orthonormal_basis(l, &t1, &t2);
  float visible = 0.0f;
  for (int k0 = 0; k0 < n; k0 += kShadowPacket) {
    V3 dirs[kShadowPacket];
    float tmax[kShadowPacket];
    if (soft) {
#pragma unroll
      for (int j = 0; j < kShadowPacket; ++j) {
        V2 u = shadow_rand2(c.seed, stream, nlight, li, n, min(k0 + j, n - 1));
        float r = L.bulbradius * sqrtf(u.x);
        float sp, cp;
        sincos_turns(u.y, &sp, &cp);
        V3 to_target = L.pos + (t1 * cp + t2 * sp) * r - origin;
        float d = len(to_target);
        dirs[j] = to_target / d;
        tmax[j] = d - kShadowGap;
      }
    } else {
#pragma unroll
      for (int j = 0; j < kShadowPacket; ++j) {
        dirs[j] = l;
        tmax[j] = directional ? kUnboundedShadow : dist - kShadowGap;
      }
    }
    const int m = min(kShadowPacket, n - k0);
    const uint32_t hit = trace_any_packet<kMesh, kShadowPacket>(c.sv, origin, dirs, tmax, m, cull);
    visible += (float)(m - __builtin_popcount(hit));
  }
  float v = visible / (float)n;
  return v + (1.0f - v) * s.shadow_light_fraction;
}

// The unshadowed part of render.py:129-161: direction and distance to the light, the BRDF
// times the light's diffuse / specular scales (`term`) and n.l. False when the light adds nothing
// (back-facing or zero scales).
MJ_DEV bool light_term(const RtCtx& c, const Light& L, const Lobes& lb, V3 point, V3 normal, V3 view, float nov, V3* l,
                       float* dist, V3* term, float* nol) {
  V3 ds, ss;
  eval_light(L, c.s.light_model, c.s.has_spot != 0, point, l, dist, &ds, &ss);
  *nol = dot(normal, *l);
  if (*nol <= 0.0f || (len_sq(ds) == 0.0f && len_sq(ss) == 0.0f)) return false;
  V3 h = normalize(*l + view);
  float noh = fmaxf(dot(normal, h), 0.0f);
  float loh = fmaxf(dot(*l, h), 0.0f);
  V3 fr = cw(specular_lobe(lb.f0, lb.alpha, noh, nov, *nol, loh), lb.energy_comp);
  V3 fd = lb.diffuse_color * kInvPi;
  *term = cw(fd, ds) + cw(fr, ss);
  return true;
}

// render.py:129-161
template <unsigned F>
MJ_DEV V3 shade_light(const RtCtx& c, const Light& L, const Lobes& lb, V3 point, V3 normal, V3 view, float nov,
                      int stream, int nlight, int li) {
  V3 l, term;
  float dist, nol;
  if (!light_term(c, L, lb, point, normal, view, nov, &l, &dist, &term, &nol)) return v3s(0.0f);
  float vis = 1.0f;
  if (c.s.use_shadows && L.castshadow) vis = light_visibility<F>(c, L, point, normal, l, dist, stream, nlight, li);
  return term * (nol * vis);
}

// The image-based part of render.py:184-221: split-sum diffuse and specular with Lagarde
// specular occlusion; also returns the lobes and n.v the lights need. kStagedYaw: the sky
// rotation's cosine and sine come from the world's staged record (the same values, evaluated once
// per world) instead of being evaluated here; instances pick it where it measured faster.
template <bool kStagedYaw>
MJ_DEV V3 ambient_radiance(const RtCtx& c, const Surface& sf, V3 normal, V3 view, Lobes* lobes, float* nov_out) {
  const MjpbrEnvironment& env = c.sc.env;
  float nov = fmaxf(dot(normal, view), 1.0e-4f);
  Lobes lb = shading_lobes(env, sf, nov);
  float alpha = lb.alpha;
  V3 ambient = ld3(env.ambient + 3 * (c.w % env.ambient_n));
  V3 irradiance = ambient, prefiltered = ambient;
  if (c.s.has_env) {
    int e;
    float yaw, scale;
    env_lookup(env, c.w, &e, &yaw, &scale);
    const V3 rn = kStagedYaw ? rotate_cs(normal, c.sv.env[0], c.sv.env[1]) : rotate_yaw(normal, -yaw);
    irradiance = ambient + eval_sh9(env.env_sh + e * 27, rn) * scale;
    V3 r = view - normal * (2.0f * dot(view, normal));
    V3 dominant = normalize(-r * (1.0f - alpha * alpha) + normal * (alpha * alpha));
    prefiltered = ambient + sample_specular(env, e, dominant, yaw, lb.perceptual_roughness) * scale;
  }
  float ao = sf.occlusion;
  float spec_ao = clampf(pow_hw(nov + ao, exp2f(-16.0f * alpha - 1.0f)) - 1.0f + ao, 0.0f, 1.0f);
  *lobes = lb;
  *nov_out = nov;
  return cw(irradiance, lb.diffuse_color) * ao + cw(prefiltered, cw(lb.spec_albedo, lb.energy_comp)) * spec_ao;
}

// render.py:184-239
template <unsigned F>
MJ_DEV V3 shade_surface(const RtCtx& c, const Surface& sf, V3 point, V3 normal, V3 view, V3 head_dir, int stream) {
  Lobes lb;
  float nov;
  // staged sky rotation: 1.5 % faster without meshes, 4.5 % slower with them (1024 x 128^2)
  V3 radiance = ambient_radiance<(F & kFMesh) == 0>(c, sf, normal, view, &lb, &nov);
  const MjpbrLights& lt = c.sc.lit;
  int nl = lt.nlight + (c.s.headlight_active ? 1 : 0);
  for (int li = 0; li < nl; ++li) {
    Light L;
    int nlight = lt.nlight, lidx = li;
    if (li < lt.nlight) {
      if (!staged_light_active(c.sv.lights, li)) continue;
      L = staged_light(c.sv.lights, li);
    } else {
      L = headlight(c.s, c.s.headlight_intensity / c.exposure, head_dir);
      nlight = 1;
      lidx = 0;
    }
    radiance += shade_light<F>(c, L, lb, point, normal, view, nov, stream, nlight, lidx);
  }
  return radiance + sf.emissive / c.exposure;
}

// render.py:241-247
MJ_DEV V3 background(const RtCtx& c, V3 dir, float spread) {
  const MjpbrEnvironment& env = c.sc.env;
  if (c.s.env_background) {
    int e;
    float yaw, scale;
    env_lookup(env, c.w, &e, &yaw, &scale);
    float lod = log2_hw(fmaxf(spread * c.s.inv_texel_angle, kMinTexels));
    return sample_radiance(env, e, dir, yaw, lod) * scale;
  }
  return ld3(env.background) / c.exposure;
}

MJ_DEV void write_guides(const MjpbrOutputs& out, int w, const Pixel& pix, V3 albedo, V3 normal) {
  int64_t a = (int64_t)w * out.guide_stride + out.rgb_adr[pix.camid] + pix.pixelid;
  st3(out.albedo + 3 * a, albedo);
  st3(out.normal + 3 * a, normal);
}

}  // namespace mjpbr
