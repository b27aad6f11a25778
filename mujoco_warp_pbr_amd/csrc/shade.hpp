// shade.hpp — surface evaluation, textures, environment lookups, lights and the Filament BRDF
// (reference: material.py, texture.py, envmap.py, ibl.py, lights.py, brdf.py, color.py).
#pragma once
#include "trace.hpp"

namespace mjpbr {

// Texture roles (MuJoCo mjtTextureRole).
constexpr int kRoleRGB = 1, kRoleOcclusion = 2, kRoleRoughness = 3, kRoleMetallic = 4, kRoleNormal = 5,
              kRoleOpacity = 6, kRoleEmissive = 7, kRoleRGBA = 8, kRoleORM = 9;
constexpr float kMinPerceptualRoughness = 0.045f;  // brdf.py:29
constexpr float kSpecAAVariance = 0.15f, kSpecAAThreshold = 0.2f;  // material.py:58-59
constexpr float kMinTexels = 1.0e-8f;  // render.py:64

struct Surface {
  V3 base_color;
  float alpha;
  float metallic;
  float perceptual_roughness;
  float reflectance;
  V3 emissive;
  float occlusion;
};

struct Lobes {
  float perceptual_roughness;
  float alpha;
  V3 f0;
  V3 energy_comp;
  V3 spec_albedo;
  V3 diffuse_color;
};

struct Light {
  int type;
  bool castshadow;
  V3 pos, dir, diffuse, specular, attenuation;
  float cutoff, exponent, intensity, range, bulbradius;
};

template <class T>
MJ_DEV const T* bat(const T* p, int n, int w, int stride) {
  return p + (int64_t)(w % n) * stride;
}

// ---------------------------------------------------------------- textures (own sampler)
// Bilinear filtering at texel centres with fp32 weights; trilinear across mip levels.
MJ_DEV float4 load_texel(const uint8_t* base, int fmt, int idx) {
  if (fmt == MJPBR_TEX_RGBA8) {
    uchar4 t = reinterpret_cast<const uchar4*>(base)[idx];
    const float k = 1.0f / 255.0f;
    return make_float4(t.x * k, t.y * k, t.z * k, t.w * k);
  }
  ushort4 t = reinterpret_cast<const ushort4*>(base)[idx];
  return make_float4(half_to_float(t.x), half_to_float(t.y), half_to_float(t.z), half_to_float(t.w));
}

MJ_DEV float4 lerp4(float4 a, float4 b, float t) {
  float s = 1.0f - t;
  return make_float4(a.x * s + b.x * t, a.y * s + b.y * t, a.z * s + b.z * t, a.w * s + b.w * t);
}

// wrap_v: repeat (material textures) or clamp (equirect rows).
MJ_DEV float4 bilerp(const uint8_t* base, int fmt, int w, int h, float u, float v, bool wrap_v) {
  float uu = u - floorf(u);
  float x = uu * (float)w - 0.5f;
  float x0f = floorf(x);
  float fx = x - x0f;
  int x0 = (int)x0f;
  int i0 = x0 < 0 ? w - 1 : x0;
  int i1 = x0 + 1 >= w ? 0 : x0 + 1;
  float vv = wrap_v ? v - floorf(v) : v;
  float y = vv * (float)h - 0.5f;
  float y0f = floorf(y);
  float fy = y - y0f;
  int y0 = (int)y0f;
  int j0, j1;
  if (wrap_v) {
    j0 = y0 < 0 ? h - 1 : y0;
    j1 = y0 + 1 >= h ? 0 : y0 + 1;
  } else {
    j0 = min(max(y0, 0), h - 1);
    j1 = min(max(y0 + 1, 0), h - 1);
  }
  float4 t00 = load_texel(base, fmt, j0 * w + i0);
  float4 t10 = load_texel(base, fmt, j0 * w + i1);
  float4 t01 = load_texel(base, fmt, j1 * w + i0);
  float4 t11 = load_texel(base, fmt, j1 * w + i1);
  return lerp4(lerp4(t00, t10, fx), lerp4(t01, t11, fx), fy);
}

MJ_DEV float4 sample_mip(const uint8_t* pool, const int64_t* level_adr, const int* desc, int lvl_scale, V2 uv,
                         float lod, bool wrap_v) {
  int w0 = desc[0], h0 = desc[1], nl = desc[2], fmt = desc[3];
  lod = clampf(lod, 0.0f, (float)(nl - 1));
  float l0f = floorf(lod);
  float f = lod - l0f;
  int l0 = (int)l0f;
  int l1 = min(l0 + 1, nl - 1);
  float4 c0 = bilerp(pool + level_adr[l0] * lvl_scale, fmt, max(1, w0 >> l0), max(1, h0 >> l0), uv.x, uv.y, wrap_v);
  if (f > 0.0f) {
    float4 c1 =
        bilerp(pool + level_adr[l1] * lvl_scale, fmt, max(1, w0 >> l1), max(1, h0 >> l1), uv.x, uv.y, wrap_v);
    c0 = lerp4(c0, c1, f);
  }
  return c0;
}

MJ_DEV float4 sample_texture(const MjpbrMaterials& m, int tex, V2 uv, float lod) {
  return sample_mip(m.tex_data, m.tex_level_adr + tex * MJPBR_MAX_LEVELS, m.tex_desc + 4 * tex, 1, uv,
                    fmaxf(lod, 0.0f), true);
}

// texture.py:224-228
MJ_DEV float texture_lod(const MjpbrMaterials& m, int tex, float footprint, float uv_per_length) {
  float size = (float)max(m.tex_desc[4 * tex], m.tex_desc[4 * tex + 1]);
  float texels = footprint * uv_per_length * size;
  return log2_hw(fmaxf(texels, 1.0e-8f));
}

// ---------------------------------------------------------------- environment
// envmap.py:145-151
MJ_DEV V2 direction_to_uv(V3 d, float yaw) {
  float phi = atan2f(d.y, d.x) - yaw;
  float u = 0.5f + phi * (0.5f * kInvPi);
  float v = acosf(clampf(d.z, -1.0f, 1.0f)) * kInvPi;
  return V2{u - floorf(u), v};
}

MJ_DEV V3 uv_to_direction(V2 uv) {
  float sp, cp, st, ct;
  sincos_turns(uv.x - 0.5f, &sp, &cp);
  sincos_turns(uv.y * 0.5f, &st, &ct);
  return V3{st * cp, st * sp, ct};
}

// d rotated about z by the angle whose cosine and sine are c, s.
MJ_DEV V3 rotate_cs(V3 d, float c, float s) { return V3{c * d.x - s * d.y, s * d.x + c * d.y, d.z}; }

MJ_DEV V3 rotate_yaw(V3 d, float yaw) {
  float c = cosf(yaw), s = sinf(yaw);
  return V3{c * d.x - s * d.y, s * d.x + c * d.y, d.z};
}

MJ_DEV V3 sample_radiance(const MjpbrEnvironment& env, int e, V3 dir, float yaw, float lod) {
  V2 uv = direction_to_uv(dir, yaw);
  float4 c = sample_mip(reinterpret_cast<const uint8_t*>(env.env_radiance), env.env_level_adr + e * MJPBR_MAX_LEVELS,
                        env.env_desc + 4 * e, 2, uv, fmaxf(lod, 0.0f), false);
  return V3{c.x, c.y, c.z};
}

MJ_DEV V3 sample_prefiltered(const MjpbrEnvironment& env, int e, V3 dir, float yaw, float pr) {
  V2 uv = direction_to_uv(dir, yaw);
  int L = env.pf_levels, H = env.pf_height, W = env.pf_width;
  float z = clampf(pr, 0.0f, 1.0f) * (float)(L - 1);
  float z0f = floorf(z);
  float fz = z - z0f;
  int z0 = min((int)z0f, L - 1);
  int z1 = min(z0 + 1, L - 1);
  const uint8_t* base = reinterpret_cast<const uint8_t*>(env.env_prefiltered + (int64_t)e * L * H * W * 4);
  float4 c = bilerp(base + (int64_t)z0 * H * W * 8, MJPBR_TEX_RGBA16F, W, H, uv.x, uv.y, false);
  if (fz > 0.0f) {
    float4 c1 = bilerp(base + (int64_t)z1 * H * W * 8, MJPBR_TEX_RGBA16F, W, H, uv.x, uv.y, false);
    c = lerp4(c, c1, fz);
  }
  return V3{c.x, c.y, c.z};
}

// envmap.py:281-289
MJ_DEV V3 sample_specular(const MjpbrEnvironment& env, int e, V3 dir, float yaw, float pr) {
  V3 blurred = sample_prefiltered(env, e, dir, yaw, pr);
  float t = pr * (float)(env.pf_levels - 1);
  if (t >= 1.0f) return blurred;
  V3 sharp = sample_radiance(env, e, dir, yaw, 0.0f);
  return sharp * (1.0f - t) + blurred * t;
}

// envmap.py:252-266
MJ_DEV V3 eval_sh9(const float* sh, V3 n) {
  float x = n.x, y = n.y, z = n.z;
  float b[9] = {0.282095f,
                0.488603f * y,
                0.488603f * z,
                0.488603f * x,
                1.092548f * x * y,
                1.092548f * y * z,
                0.315392f * (3.0f * z * z - 1.0f),
                1.092548f * x * z,
                0.546274f * (x * x - y * y)};
  V3 c = v3s(0.0f);
#pragma unroll
  for (int k = 0; k < 9; ++k) c = c + ld3(sh + 3 * k) * b[k];
  return vmax(c, 0.0f);
}

// ibl.py:117-119: bilinear, clamped, at (NoV, perceptual roughness).
MJ_DEV V2 sample_dfg(const MjpbrEnvironment& env, float nov, float pr) {
  int N = env.dfg_size;
  float x = nov * (float)N - 0.5f, y = pr * (float)N - 0.5f;
  float x0f = floorf(x), y0f = floorf(y);
  float fx = x - x0f, fy = y - y0f;
  int x0 = (int)x0f, y0 = (int)y0f;
  int i0 = min(max(x0, 0), N - 1), i1 = min(max(x0 + 1, 0), N - 1);
  int j0 = min(max(y0, 0), N - 1), j1 = min(max(y0 + 1, 0), N - 1);
  const float2* lut = reinterpret_cast<const float2*>(env.dfg_lut);
  float2 a = lut[j0 * N + i0], b = lut[j0 * N + i1], c = lut[j1 * N + i0], d = lut[j1 * N + i1];
  float sx = 1.0f - fx, sy = 1.0f - fy;
  return V2{(a.x * sx + b.x * fx) * sy + (c.x * sx + d.x * fx) * fy, (a.y * sx + b.y * fx) * sy + (c.y * sx + d.y * fx) * fy};
}

// ---------------------------------------------------------------- BRDF (brdf.py)
MJ_DEV float d_ggx(float noh, float alpha) {
  float a2 = alpha * alpha;
  float f = (noh * a2 - noh) * noh + 1.0f;
  return a2 / (kPi * f * f);
}
MJ_DEV float v_smith_ggx_correlated(float nov, float nol, float alpha) {
  float a2 = alpha * alpha;
  float gv = nol * sqrtf(nov * nov * (1.0f - a2) + a2);
  float gl = nov * sqrtf(nol * nol * (1.0f - a2) + a2);
  return 0.5f / (gv + gl);
}
MJ_DEV V3 f_schlick(V3 f0, float f90, float voh) {
  float x = 1.0f - voh;
  float x2 = x * x;
  float fc = x2 * x2 * x;
  return f0 * (1.0f - fc) + v3s(f90 * fc);
}
MJ_DEV V3 specular_lobe(V3 f0, float alpha, float noh, float nov, float nol, float loh) {
  return f_schlick(f0, 1.0f, loh) * (d_ggx(noh, alpha) * v_smith_ggx_correlated(nov, nol, alpha));
}

MJ_DEV Lobes shading_lobes(const MjpbrEnvironment& env, const Surface& s, float nov) {
  Lobes L;
  float pr = clampf(s.perceptual_roughness, kMinPerceptualRoughness, 1.0f);
  L.perceptual_roughness = pr;
  L.alpha = pr * pr;
  float dielectric = 0.16f * s.reflectance * s.reflectance * (1.0f - s.metallic);
  L.f0 = s.base_color * s.metallic + v3s(dielectric);
  V2 dfg = sample_dfg(env, nov, pr);
  L.energy_comp = v3s(1.0f) + L.f0 * (1.0f / fmaxf(dfg.y, 1.0e-4f) - 1.0f);
  L.spec_albedo = (v3s(1.0f) - L.f0) * dfg.x + L.f0 * dfg.y;
  L.diffuse_color = cw(s.base_color * (1.0f - s.metallic), v3s(1.0f) - L.spec_albedo);
  return L;
}

// ---------------------------------------------------------------- lights (lights.py)
MJ_DEV Light fetch_light(const MjpbrLights& lt, int w, int li) {
  Light L;
  int n = lt.nlight;
  L.type = bat(lt.light_type, lt.light_type_n, w, n)[li];
  L.castshadow = bat(lt.light_castshadow, lt.light_castshadow_n, w, n)[li] != 0;
  L.pos = ld3(lt.light_xpos + ((int64_t)w * n + li) * 3);
  L.dir = ld3(lt.light_xdir + ((int64_t)w * n + li) * 3);
  L.diffuse = ld3(bat(lt.light_diffuse, lt.light_diffuse_n, w, 3 * n) + 3 * li);
  L.specular = ld3(bat(lt.light_specular, lt.light_specular_n, w, 3 * n) + 3 * li);
  L.attenuation = ld3(bat(lt.light_attenuation, lt.light_attenuation_n, w, 3 * n) + 3 * li);
  L.cutoff = bat(lt.light_cutoff, lt.light_cutoff_n, w, n)[li];
  L.exponent = bat(lt.light_exponent, lt.light_exponent_n, w, n)[li];
  L.intensity = bat(lt.light_intensity, lt.light_intensity_n, w, n)[li];
  L.range = bat(lt.light_range, lt.light_range_n, w, n)[li];
  L.bulbradius = bat(lt.light_bulbradius, lt.light_bulbradius_n, w, n)[li];
  return L;
}

// Staged light record (kLightRec floats): [0] type | castshadow << 8 | active << 9 (bits),
// [1..3] pos, [4..6] dir, [7..9] diffuse, [10..12] specular, [13..15] attenuation, [16] cutoff,
// [17] exponent, [18] intensity, [19] range, [20] bulbradius, [21..23] unused.
MJ_DEV void stage_light(const MjpbrLights& lt, int w, int li, float* r) {
  Light L = fetch_light(lt, w, li);
  int active = bat(lt.light_active, lt.light_active_n, w, lt.nlight)[li] != 0;
  r[0] = __int_as_float(L.type | (L.castshadow ? 0x100 : 0) | (active << 9));
  V3 v[5] = {L.pos, L.dir, L.diffuse, L.specular, L.attenuation};
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    r[1 + 3 * k] = v[k].x;
    r[2 + 3 * k] = v[k].y;
    r[3 + 3 * k] = v[k].z;
  }
  r[16] = L.cutoff;
  r[17] = L.exponent;
  r[18] = L.intensity;
  r[19] = L.range;
  r[20] = L.bulbradius;
  r[21] = r[22] = r[23] = 0.0f;
}

MJ_DEV bool staged_light_active(const float* lights, int li) {
  return (__float_as_int(lights[li * kLightRec]) >> 9) & 1;
}

MJ_DEV Light staged_light(const float* lights, int li) {
  const float4* q = reinterpret_cast<const float4*>(lights + li * kLightRec);
  float4 a = q[0], b = q[1], c = q[2], d = q[3], e = q[4], f = q[5];
  Light L;
  int bits = __float_as_int(a.x);
  L.type = bits & 0xff;
  L.castshadow = (bits & 0x100) != 0;
  L.pos = V3{a.y, a.z, a.w};
  L.dir = V3{b.x, b.y, b.z};
  L.diffuse = V3{b.w, c.x, c.y};
  L.specular = V3{c.z, c.w, d.x};
  L.attenuation = V3{d.y, d.z, d.w};
  L.cutoff = e.x;
  L.exponent = e.y;
  L.intensity = e.z;
  L.range = e.w;
  L.bulbradius = f.x;
  return L;
}

MJ_DEV Light headlight(const MjpbrSettings& s, float intensity, V3 dir) {
  Light L;
  L.type = MJPBR_LIGHT_DIRECTIONAL;
  L.castshadow = false;
  L.pos = v3s(0.0f);
  L.dir = dir;
  L.diffuse = ld3(s.headlight_diffuse);
  L.specular = ld3(s.headlight_specular);
  L.attenuation = V3{1.0f, 0.0f, 0.0f};
  L.cutoff = 0.0f;
  L.exponent = 0.0f;
  L.intensity = intensity;
  L.range = 1.0f;
  L.bulbradius = 0.0f;
  return L;
}

// lights.py:151-198: (direction to light, distance, diffuse scale, specular scale).
MJ_DEV void eval_light(const Light& L, int light_model, bool has_spot, V3 point, V3* l, float* dist, V3* ds,
                       V3* ss) {
  bool physical = light_model == 0;
  float d = 1.0e10f;
  float att = 1.0f;
  if (L.type == MJPBR_LIGHT_DIRECTIONAL) {
    *l = -L.dir;
  } else {
    V3 to = L.pos - point;
    d = len(to);
    *l = to / fmaxf(d, 1.0e-9f);
    if (physical) {
      float d2 = d * d;
      float r = fmaxf(L.range, 1.0e-6f);
      float factor = d2 / (r * r);
      float window = clampf(1.0f - factor * factor, 0.0f, 1.0f);
      att = window * window / fmaxf(d2, 1.0e-4f);
    } else {
      att = 1.0f / fmaxf(dot(L.attenuation, V3{1.0f, d, d * d}), 1.0e-9f);
    }
    if (has_spot && L.type == MJPBR_LIGHT_SPOT) {
      float cos_theta = dot(-*l, L.dir);
      float outer = L.cutoff * (kPi / 180.0f);
      float cos_outer = cosf(outer);
      if (physical) {
        float cos_inner = cosf(outer * 0.75f);
        float cone_scale = 1.0f / fmaxf(cos_inner - cos_outer, 1.0f / 1024.0f);
        float sc = clampf((cos_theta - cos_outer) * cone_scale, 0.0f, 1.0f);
        att = att * sc * sc;
      } else {
        att = cos_theta < cos_outer ? 0.0f : att * powf(fmaxf(cos_theta, 0.0f), L.exponent);
      }
    }
  }
  *dist = d;
  if (physical) {
    V3 sc = L.diffuse * (L.intensity * att);
    *ds = sc;
    *ss = sc;
  } else {
    float k = kPi * att;
    *ds = L.diffuse * k;
    *ss = L.specular * k;
  }
}

// ---------------------------------------------------------------- colour (color.py)
MJ_DEV V3 tonemap_pbr_neutral(V3 c) {
  const float start = 0.8f - 0.04f;
  const float desat = 0.15f;
  float x = fminf(c.x, fminf(c.y, c.z));
  float offset = x < 0.08f ? x - 6.25f * x * x : 0.04f;
  c = c - v3s(offset);
  float peak = fmaxf(c.x, fmaxf(c.y, c.z));
  if (peak < start) return c;
  float d = 1.0f - start;
  float new_peak = 1.0f - d * d / (peak + d - start);
  c = c * (new_peak / peak);
  float g = 1.0f - 1.0f / (desat * (peak - new_peak) + 1.0f);
  return c * (1.0f - g) + v3s(new_peak * g);
}

MJ_DEV V3 tonemap_aces(V3 c) {
  V3 v = V3{0.59719f * c.x + 0.35458f * c.y + 0.04823f * c.z, 0.07600f * c.x + 0.90834f * c.y + 0.01566f * c.z,
            0.02840f * c.x + 0.13383f * c.y + 0.83777f * c.z};
  V3 num = cw(v, v + v3s(0.0245786f)) - v3s(0.000090537f);
  V3 den = cw(v, v * 0.983729f + v3s(0.4329510f)) + v3s(0.238081f);
  v = V3{num.x / den.x, num.y / den.y, num.z / den.z};
  return V3{1.60475f * v.x - 0.53108f * v.y - 0.07367f * v.z, -0.10208f * v.x + 1.10813f * v.y - 0.00605f * v.z,
            -0.00327f * v.x - 0.07276f * v.y + 1.07602f * v.z};
}

MJ_DEV V3 display_transform(V3 c, int tone_mapper) {
  c = vmax(c, 0.0f);
  if (tone_mapper == 0)
    c = tonemap_pbr_neutral(c);
  else if (tone_mapper == 1)
    c = tonemap_aces(c);
  return vmin(vmax(c, 0.0f), 1.0f);
}

MJ_DEV float srgb_oetf(float x) { return x <= 0.0031308f ? 12.92f * x : 1.055f * pow_hw(x, 1.0f / 2.4f) - 0.055f; }

MJ_DEV uint32_t encode_display(V3 c) {
  uint32_t r = (uint32_t)(srgb_oetf(c.x) * 255.0f + 0.5f);
  uint32_t g = (uint32_t)(srgb_oetf(c.y) * 255.0f + 0.5f);
  uint32_t b = (uint32_t)(srgb_oetf(c.z) * 255.0f + 0.5f);
  return (255u << 24) | (r << 16) | (g << 8) | b;  // MuJoCo Warp's 0xAARRGGBB
}

// ---------------------------------------------------------------- surfaces (material.py)
MJ_DEV float widen_roughness(float pr, float a2_extra) {
  pr = clampf(pr, 0.0f, 1.0f);
  float a2 = fminf(pr * pr * pr * pr + a2_extra, 1.0f);
  return sqrtf(sqrtf(a2));
}

// MuJoCo Warp's _geom_rgba: a geom with a material and MuJoCo's default rgba shows the
// material's colour.
MJ_DEV float4 geom_rgba(const MjpbrGeometry& g, int w, int gi, int* mat_id) {
  int mid = bat(g.geom_matid, g.geom_matid_n, w, g.ngeom)[gi];
  const float* r = bat(g.geom_rgba, g.geom_rgba_n, w, 4 * g.ngeom) + 4 * gi;
  float4 c = make_float4(r[0], r[1], r[2], r[3]);
  if (mid >= 0 && c.x == 0.5f && c.y == 0.5f && c.z == 0.5f && c.w == 1.0f) {
    const float* m = bat(g.mat_rgba, g.mat_rgba_n, w, 4 * g.nmat) + 4 * mid;
    c = make_float4(m[0], m[1], m[2], m[3]);
  }
  *mat_id = mid;
  return c;
}

MJ_DEV V2 wrap_delta(V2 d) { return V2{d.x - rintf(d.x), d.y - rintf(d.y)}; }

// Analytic UVs of primitives (own convention: unrepeated uv), at the three points of the
// finite-difference UV Jacobian (material.py:412-428) in one switch over the geom type, so a
// wave whose lanes hit different primitive types branches once, not three times.
MJ_DEV void primitive_uv3(int type, const V3 (&p)[3], V3 ln, V3 size, bool uniform, V2 (&uv)[3]) {
  switch (type) {
    case MJPBR_GEOM_PLANE:
    case MJPBR_GEOM_HFIELD: {
      const bool plain = uniform || size.x <= 0.0f || size.y <= 0.0f;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        uv[k] = plain ? V2{p[k].x, -p[k].y} : V2{p[k].x / (2.0f * size.x) + 0.5f, 0.5f - p[k].y / (2.0f * size.y)};
      break;
    }
    case MJPBR_GEOM_SPHERE:
    case MJPBR_GEOM_ELLIPSOID: {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        V3 q = type == MJPBR_GEOM_SPHERE ? p[k] : V3{p[k].x / size.x, p[k].y / size.y, p[k].z / size.z};
        float r = len(q);
        float u = atan2f(q.y, q.x) * (0.5f * kInvPi) + 0.5f;
        float v = acosf(clampf(r > 0.0f ? q.z / r : 1.0f, -1.0f, 1.0f)) * kInvPi;
        uv[k] = V2{u, v};
      }
      break;
    }
    case MJPBR_GEOM_CYLINDER: {
      const bool cap = fabsf(ln.z) > 0.5f;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (cap)
          uv[k] = V2{p[k].x / (2.0f * size.x) + 0.5f, 0.5f - p[k].y / (2.0f * size.x)};
        else
          uv[k] = V2{atan2f(p[k].y, p[k].x) * (0.5f * kInvPi) + 0.5f, 0.5f - p[k].z / (2.0f * size.y)};
      }
      break;
    }
    case MJPBR_GEOM_CAPSULE: {
#pragma unroll
      for (int k = 0; k < 3; ++k)
        uv[k] = V2{atan2f(p[k].y, p[k].x) * (0.5f * kInvPi) + 0.5f, 0.5f - p[k].z / (2.0f * (size.y + size.x))};
      break;
    }
    case MJPBR_GEOM_BOX: {
      // the face is chosen by the local normal, the same for the three points
      float ax = fabsf(ln.x), ay = fabsf(ln.y), az = fabsf(ln.z);
      const int ia = (ax >= ay && ax >= az) ? 1 : 0;
      const int ib = (ax >= ay && ax >= az) || ay >= az ? 2 : 1;
      const float sa = ia == 1 ? size.y : size.x;
      const float sb = ib == 2 ? size.z : size.y;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float a = ia == 1 ? p[k].y : p[k].x;
        float b = ib == 2 ? p[k].z : p[k].y;
        uv[k] = uniform ? V2{a, -b} : V2{a / (2.0f * sa) + 0.5f, 0.5f - b / (2.0f * sb)};
      }
      break;
    }
    default:
#pragma unroll
      for (int k = 0; k < 3; ++k) uv[k] = V2{0.0f, 0.0f};
  }
}

// texture.py:161-175
MJ_DEV void frame_from_uv_jacobian(V3 t1, V3 t2, V2 du_dt, V2 dv_dt, V3* tangent, V3* bitangent, float* upl) {
  float det = du_dt.x * dv_dt.y - du_dt.y * dv_dt.x;
  float inv = 1.0f / (fabsf(det) > 1.0e-20f ? det : 1.0e-20f);
  *tangent = (t1 * dv_dt.y - t2 * dv_dt.x) * inv;
  *bitangent = (t2 * du_dt.x - t1 * du_dt.y) * inv;
  *upl = sqrtf(fabsf(det));
}

// texture.py:178-221
MJ_DEV bool mesh_triangle_frame(const MjpbrMaterials& m, int mesh, int face, float bu, float bv, V2 repeat, V2* uv,
                                V3* tangent, V3* bitangent, float* upl) {
  int tc_adr = m.mesh_texcoordadr[mesh];
  if (tc_adr < 0) return false;
  int face_adr = m.mesh_faceadr[mesh] + face;
  const int* tc = m.mesh_facetexcoord + 3 * face_adr;
  if (tc[0] < 0 || tc[1] < 0 || tc[2] < 0) return false;
  const float* T = m.mesh_texcoord;
  V2 uv0 = V2{T[2 * (tc_adr + tc[0])] * repeat.x, T[2 * (tc_adr + tc[0]) + 1] * repeat.y};
  V2 uv1 = V2{T[2 * (tc_adr + tc[1])] * repeat.x, T[2 * (tc_adr + tc[1]) + 1] * repeat.y};
  V2 uv2 = V2{T[2 * (tc_adr + tc[2])] * repeat.x, T[2 * (tc_adr + tc[2]) + 1] * repeat.y};
  float bw = 1.0f - bu - bv;
  *uv = V2{uv0.x * bu + uv1.x * bv + uv2.x * bw, uv0.y * bu + uv1.y * bv + uv2.y * bw};
  int vadr = m.mesh_vertadr[mesh];
  const int* f = m.mesh_face + 3 * face_adr;
  V3 p0 = ld3(m.mesh_vert + 3 * (vadr + f[0]));
  V3 e1 = ld3(m.mesh_vert + 3 * (vadr + f[1])) - p0;
  V3 e2 = ld3(m.mesh_vert + 3 * (vadr + f[2])) - p0;
  V2 d1 = V2{uv1.x - uv0.x, uv1.y - uv0.y};
  V2 d2 = V2{uv2.x - uv0.x, uv2.y - uv0.y};
  float det = d1.x * d2.y - d2.x * d1.y;
  float inv = 1.0f / (fabsf(det) > 1.0e-20f ? det : 1.0e-20f);
  *tangent = (e1 * d2.y - e2 * d1.y) * inv;
  *bitangent = (e2 * d1.x - e1 * d2.x) * inv;
  float area = len(cross(e1, e2));
  *upl = sqrtf(fabsf(det) / fmaxf(area, 1.0e-20f));
  return true;
}

MJ_DEV V3 vertex_normal(V3 n, V3 face_normal) { return dot(n, face_normal) < 0.0f ? -n : n; }

MJ_DEV float4 fetch_role(const MjpbrMaterials& m, int tex, V2 uv, float cone, float upl) {
  return sample_texture(m, tex, uv, texture_lod(m, tex, cone, upl));
}

// material.py:300-501. Returns the surface and the viewer-facing world shading normal.
// The hit geom's material constants, frame and size come from the world's staged records (LDS).
// kMesh = false / kTex = false drop the mesh-attribute and texture branches (scenes without
// mesh-like geoms / settings without textures).
template <bool kMesh = true, bool kTex = true>
MJ_DEV Surface eval_surface(const MjpbrScene& sc, const SceneView& sv, const MjpbrSettings& s, int w, const Hit& hit,
                            V3 point, V3 dir, float footprint, V3* shading_world) {
  const MjpbrGeometry& g = sc.geo;
  const MjpbrMaterials& m = sc.mats;
  Surface sf;
  sf.occlusion = 1.0f;
  sf.emissive = v3s(0.0f);
  int gid = hit.geom_id;
  const float4* sr = reinterpret_cast<const float4*>(sv.surfs + gid * kSurfRec);
  float4 r0 = sr[0], r1 = sr[1], r2 = sr[2], r3 = sr[3];
  sf.base_color = V3{r0.x, r0.y, r0.z};
  sf.alpha = r0.w;
  sf.metallic = r1.x;
  sf.perceptual_roughness = r1.y;
  sf.reflectance = r1.z;
  float emission = r1.w;
  int mat_id = __float_as_int(r2.x);
  int gtype = __float_as_int(r3.y);
  const float* grec = sv.geoms + kGeomRec * gid;
  float R[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k] = grec[4 + k];
  V3 local_dir = mat_tmul(R, dir);
  V3 local_normal = mat_tmul(R, hit.normal);
  if (dot(local_normal, local_dir) > 0.0f) local_normal = -local_normal;
  V3 shading = local_normal;
  float curvature = 0.0f;
  bool is_mesh = kMesh && (gtype == MJPBR_GEOM_MESH || gtype == MJPBR_GEOM_SDF);
  int mesh_id = hit.mesh_id, face_id = hit.face_id;

  if (is_mesh && s.vertex_normals && mesh_id >= 0 && face_id >= 0) {
    int face_adr = m.mesh_faceadr[mesh_id] + face_id;
    const int* tri = m.mesh_facenormal + 3 * face_adr;
    int nadr = m.mesh_normaladr[mesh_id];
    V3 n0 = vertex_normal(ld3(m.mesh_normal + 3 * (nadr + tri[0])), local_normal);
    V3 n1 = vertex_normal(ld3(m.mesh_normal + 3 * (nadr + tri[1])), local_normal);
    V3 n2 = vertex_normal(ld3(m.mesh_normal + 3 * (nadr + tri[2])), local_normal);
    shading = normalize(n0 * hit.bary_u + n1 * hit.bary_v + n2 * (1.0f - hit.bary_u - hit.bary_v));
    if (dot(shading, local_dir) > 0.0f) shading = -shading;
    if (s.specular_aa) {
      const int* f = m.mesh_face + 3 * face_adr;
      int vadr = m.mesh_vertadr[mesh_id];
      V3 p0 = ld3(m.mesh_vert + 3 * (vadr + f[0]));
      float edge = sqrtf(len(cross(ld3(m.mesh_vert + 3 * (vadr + f[1])) - p0, ld3(m.mesh_vert + 3 * (vadr + f[2])) - p0)));
      float spread = fmaxf(len(n0 - n1), fmaxf(len(n1 - n2), len(n2 - n0)));
      curvature = spread / fmaxf(edge, 1.0e-9f);
    }
  }
  V3 size = V3{0.0f, 0.0f, 0.0f};
  if (!is_mesh) size = V3{grec[3], grec[13], grec[14]};
  if (!is_mesh && s.specular_aa) {
    if (gtype == MJPBR_GEOM_SPHERE || gtype == MJPBR_GEOM_CAPSULE || gtype == MJPBR_GEOM_CYLINDER)
      curvature = 1.0f / fmaxf(size.x, 1.0e-9f);
    else if (gtype == MJPBR_GEOM_ELLIPSOID)
      curvature = 1.0f / fmaxf(fminf(size.x, fminf(size.y, size.z)), 1.0e-9f);
  }

  float cone = footprint / fmaxf(fabsf(dot(local_normal, local_dir)), 0.05f);

  if (kTex && s.textures && mat_id >= 0) {
    int mw = w % m.mat_texid_n;
    V2 repeat = V2{r2.z, r2.w};
    bool uniform = __float_as_int(r2.y) != 0;
    bool has_uv = false;
    V2 uv = V2{0.0f, 0.0f};
    V3 tangent = v3s(0.0f), bitangent = v3s(0.0f);
    float upl = 0.0f;
    if (is_mesh) {
      if (mesh_id >= 0 && face_id >= 0)
        has_uv = mesh_triangle_frame(m, mesh_id, face_id, hit.bary_u, hit.bary_v, repeat, &uv, &tangent, &bitangent, &upl);
    } else {
      V3 gpos = V3{grec[0], grec[1], grec[2]};
      V3 local = mat_tmul(R, point - gpos);
      V3 t1, t2;
      orthonormal_basis(local_normal, &t1, &t2);
      float eps = 1.0e-3f * fmaxf(fmaxf(size.x, size.y), fmaxf(size.z, 1.0e-3f));
      const V3 pts[3] = {local, local + t1 * eps, local + t2 * eps};
      V2 uvs[3];
      primitive_uv3(gtype, pts, local_normal, size, uniform, uvs);
      const V2 uv0 = uvs[0], uv1 = uvs[1], uv2 = uvs[2];
      const V2 rep = repeat, off = V2{0.0f, 0.0f};
      V2 d1 = wrap_delta(V2{uv1.x - uv0.x, uv1.y - uv0.y});
      V2 d2 = wrap_delta(V2{uv2.x - uv0.x, uv2.y - uv0.y});
      d1 = V2{d1.x * rep.x, d1.y * rep.y};
      d2 = V2{d2.x * rep.x, d2.y * rep.y};
      float inv_eps = 1.0f / eps;
      frame_from_uv_jacobian(t1, t2, V2{d1.x * inv_eps, d2.x * inv_eps}, V2{d1.y * inv_eps, d2.y * inv_eps}, &tangent,
                             &bitangent, &upl);
      uv = V2{uv0.x * rep.x + off.x, uv0.y * rep.y + off.y};
      has_uv = true;
    }
    if (has_uv) {
      const int* tid = m.mat_texid + ((int64_t)mw * g.nmat + mat_id) * MJPBR_NTEXROLE;
      if (s.base_color_map) {
        int t = tid[kRoleRGB];
        if (t < 0) t = tid[kRoleRGBA];
        if (t >= 0) {
          float4 tx = fetch_role(m, t, uv, cone, upl);
          sf.base_color = cw(sf.base_color, V3{tx.x, tx.y, tx.z});
          sf.alpha = sf.alpha * tx.w;
        }
      }
      if (s.orm_map) {
        int t = tid[kRoleORM];
        if (t >= 0) {
          float4 tx = fetch_role(m, t, uv, cone, upl);
          sf.occlusion = tx.x;
          sf.perceptual_roughness = sf.perceptual_roughness * tx.y;
          sf.metallic = sf.metallic * tx.z;
        } else {
          t = tid[kRoleOcclusion];
          if (t >= 0) sf.occlusion = fetch_role(m, t, uv, cone, upl).x;
          t = tid[kRoleRoughness];
          if (t >= 0) sf.perceptual_roughness = sf.perceptual_roughness * fetch_role(m, t, uv, cone, upl).x;
          t = tid[kRoleMetallic];
          if (t >= 0) sf.metallic = sf.metallic * fetch_role(m, t, uv, cone, upl).x;
        }
      }
      if (s.opacity_map) {
        int t = tid[kRoleOpacity];
        if (t >= 0) sf.alpha = sf.alpha * fetch_role(m, t, uv, cone, upl).x;
      }
      if (s.emissive_map) {
        int t = tid[kRoleEmissive];
        if (t >= 0) {
          float4 tx = fetch_role(m, t, uv, cone, upl);
          sf.emissive = V3{tx.x, tx.y, tx.z} * emission;
          emission = 0.0f;
        }
      }
      if (s.normal_map) {
        int t = tid[kRoleNormal];
        if (t >= 0) {
          float4 tx = fetch_role(m, t, uv, cone, upl);
          V3 nts = V3{tx.x, tx.y, tx.z} * 2.0f - v3s(1.0f);
          if (s.specular_aa) {
            float l = clampf(len(nts), 1.0e-3f, 1.0f);
            sf.perceptual_roughness = widen_roughness(sf.perceptual_roughness, (1.0f - l) / l);
          }
          V3 tt = tangent - shading * dot(shading, tangent);
          if (len_sq(tt) > 1.0e-20f) {
            tt = normalize(tt);
            V3 b = cross(shading, tt);
            if (dot(b, bitangent) > 0.0f) b = -b;
            V3 mapped = normalize(tt * nts.x + b * nts.y + shading * nts.z);
            if (dot(mapped, local_normal) > 0.0f) shading = mapped;
          }
        }
      }
    }
  }
  if (mat_id >= 0 && emission > 0.0f) sf.emissive = sf.base_color * emission;
  if (s.specular_aa) {
    float dn = cone * curvature;
    float kern = fminf(2.0f * kSpecAAVariance * dn * dn, kSpecAAThreshold);
    sf.perceptual_roughness = widen_roughness(sf.perceptual_roughness, kern);
  }
  *shading_world = normalize(mat_mul(R, shading));
  return sf;
}

MJ_DEV float mirror_reflectance(const MjpbrScene& sc, int w, int gid) {
  if (gid < 0) return 0.0f;
  int mid = bat(sc.geo.geom_matid, sc.geo.geom_matid_n, w, sc.geo.ngeom)[gid];
  if (mid < 0) return 0.0f;
  return bat(sc.mats.mat_reflectance, sc.mats.mat_reflectance_n, w, sc.geo.nmat)[mid];
}

// Staged surface record of geom gi in world w (kSurfRec floats): [0..3] rgba (geom_rgba),
// [4] metallic, [5] perceptual roughness, [6] reflectance, [7] material emission, [8] material
// id (bits), [9] texuniform (bits), [10..11] texrepeat, [12] mirror reflectance, [13] geom type
// (bits), [14..15] unused.
MJ_DEV void stage_surface(const MjpbrScene& sc, int w, int gi, float* r) {
  const MjpbrGeometry& g = sc.geo;
  const MjpbrMaterials& m = sc.mats;
  int mid;
  float4 rgba = geom_rgba(g, w, gi, &mid);
  r[0] = rgba.x;
  r[1] = rgba.y;
  r[2] = rgba.z;
  r[3] = rgba.w;
  r[4] = bat(m.geom_metallic, m.geom_metallic_n, w, g.ngeom)[gi];
  r[5] = bat(m.geom_roughness, m.geom_roughness_n, w, g.ngeom)[gi];
  r[6] = bat(m.geom_reflectance, m.geom_reflectance_n, w, g.ngeom)[gi];
  r[7] = mid >= 0 ? bat(m.mat_emission, m.mat_emission_n, w, g.nmat)[mid] : 0.0f;
  r[8] = __int_as_float(mid);
  r[9] = __int_as_float(mid >= 0 ? bat(m.mat_texuniform, m.mat_texuniform_n, w, g.nmat)[mid] : 0);
  const float* rp = mid >= 0 ? bat(m.mat_texrepeat, m.mat_texrepeat_n, w, 2 * g.nmat) + 2 * mid : nullptr;
  r[10] = rp ? rp[0] : 0.0f;
  r[11] = rp ? rp[1] : 0.0f;
  r[12] = mirror_reflectance(sc, w, gi);
  r[13] = __int_as_float(g.geom_type[gi]);
  r[14] = r[15] = 0.0f;
}

MJ_DEV void env_lookup(const MjpbrEnvironment& env, int w, int* e, float* yaw, float* scale) {
  *e = env.env_id[w % env.env_id_n];
  *yaw = env.env_yaw[w % env.env_yaw_n];
  *scale = env.env_intensity[w % env.env_intensity_n];
}

}  // namespace mjpbr
