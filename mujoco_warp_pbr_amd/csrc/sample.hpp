// sample.hpp — BSDF and environment importance sampling of the path tracer
// (reference brdf.py:116-191, envmap.py:433-490).
#pragma once
#include "shade.hpp"

namespace mjpbr {

MJ_DEV float smith_g1(float nov, float alpha) {
  float a2 = alpha * alpha;
  return 2.0f * nov / (nov + sqrtf(a2 + (1.0f - a2) * nov * nov));
}

// Half vector distributed as GGX's visible normals for view v (Heitz 2018), local frame.
MJ_DEV V3 sample_ggx_vndf(V3 v, float alpha, V2 u) {
  V3 vh = normalize(V3{alpha * v.x, alpha * v.y, v.z});
  float lensq = vh.x * vh.x + vh.y * vh.y;
  V3 t1 = V3{1.0f, 0.0f, 0.0f};
  if (lensq > 0.0f) t1 = V3{-vh.y, vh.x, 0.0f} / sqrtf(lensq);
  V3 t2 = cross(vh, t1);
  float r = sqrtf(u.x);
  float sp, cp;
  sincos_turns(u.y, &sp, &cp);
  float p1 = r * cp;
  float p2 = r * sp;
  float s = 0.5f * (1.0f + vh.z);
  p2 = (1.0f - s) * sqrtf(fmaxf(1.0f - p1 * p1, 0.0f)) + s * p2;
  V3 nh = t1 * p1 + t2 * p2 + vh * sqrtf(fmaxf(1.0f - p1 * p1 - p2 * p2, 0.0f));
  return normalize(V3{alpha * nh.x, alpha * nh.y, fmaxf(nh.z, 0.0f)});
}

// BRDF x cos(theta_l), local frame (brdf.py:140-156).
MJ_DEV V3 eval_bsdf_cos(V3 diffuse, V3 f0, float alpha, V3 ec, V3 v, V3 l) {
  float nol = l.z;
  float nov = fmaxf(v.z, 1.0e-4f);
  if (nol <= 0.0f) return v3s(0.0f);
  V3 h = normalize(l + v);
  V3 spec = cw(specular_lobe(f0, alpha, fmaxf(h.z, 0.0f), nov, nol, fmaxf(dot(l, h), 0.0f)), ec);
  return (diffuse * kInvPi + spec) * nol;
}

// Solid-angle pdf of sample_bsdf (brdf.py:159-168).
MJ_DEV float pdf_bsdf(float alpha, float p_spec, V3 v, V3 l) {
  if (l.z <= 0.0f) return 0.0f;
  float nov = fmaxf(v.z, 1.0e-4f);
  V3 h = normalize(l + v);
  float pdf_spec = smith_g1(nov, alpha) * d_ggx(fmaxf(h.z, 0.0f), alpha) / (4.0f * nov);
  float pdf_diff = l.z * kInvPi;
  return p_spec * pdf_spec + (1.0f - p_spec) * pdf_diff;
}

MJ_DEV float specular_probability(V3 diffuse_color, V3 spec_albedo) {
  float ws = luminance(spec_albedo);
  float wd = luminance(diffuse_color);
  if (wd <= 0.0f) return 1.0f;
  return clampf(ws / fmaxf(ws + wd, 1.0e-6f), 0.1f, 0.9f);
}

MJ_DEV V3 sample_bsdf(float alpha, float p_spec, V3 v, float u_lobe, V2 u) {
  if (u_lobe < p_spec) {
    V3 h = sample_ggx_vndf(v, alpha, u);
    return h * (2.0f * dot(v, h)) - v;
  }
  float r = sqrtf(u.x);
  float sp, cp;
  sincos_turns(u.y, &sp, &cp);
  return V3{r * cp, r * sp, sqrtf(fmaxf(1.0f - u.x, 0.0f))};
}

// Yaw rotation about +z through the exact-reduction sincos (yaw in radians).
MJ_DEV V3 rotate_yaw_exact(V3 d, float yaw) {
  float s, c;
  sincos_turns(yaw * 0.159154943091895335768f, &s, &c);
  return V3{c * d.x - s * d.y, s * d.x + c * d.y, d.z};
}

// Largest i in [0, n) with cdf[i] <= u (envmap.py:433-444).
MJ_DEV int cdf_search(const float* cdf, int n, float u) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (cdf[mid] <= u)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// World direction drawn proportional to the environment's luminance and its solid-angle pdf
// (envmap.py:447-478).
MJ_DEV V3 env_sample_direction(const MjpbrEnvironment& env, int e, float yaw, V2 u, float* pdf) {
  int h = env.cdf_h, w = env.cdf_w;
  const float* marginal = env.env_marginal + (int64_t)e * (h + 1);
  int r = cdf_search(marginal, h, u.x);
  float c0 = marginal[r];
  float dv = (u.x - c0) / fmaxf(marginal[r + 1] - c0, 1.0e-12f);
  const float* cond = env.env_conditional + ((int64_t)e * h + r) * (w + 1);
  int c = cdf_search(cond, w, u.y);
  float k0 = cond[c];
  float du = (u.y - k0) / fmaxf(cond[c + 1] - k0, 1.0e-12f);
  V2 uv = V2{((float)c + clampf(du, 0.0f, 1.0f)) / (float)w, ((float)r + clampf(dv, 0.0f, 1.0f)) / (float)h};
  V3 d = rotate_yaw_exact(uv_to_direction(uv), yaw);
  float st, ct;
  sincos_turns(uv.y * 0.5f, &st, &ct);
  float sin_theta = fmaxf(st, 1.0e-6f);
  *pdf = env.env_prob[((int64_t)e * h + r) * w + c] * (float)(w * h) / (2.0f * kPi * kPi * sin_theta);
  return d;
}

MJ_DEV float env_direction_pdf(const MjpbrEnvironment& env, int e, float yaw, V3 dir) {
  int h = env.cdf_h, w = env.cdf_w;
  V2 uv = direction_to_uv(dir, yaw);
  int c = min((int)(uv.x * (float)w), w - 1);
  int r = min((int)(uv.y * (float)h), h - 1);
  float st, ct;
  sincos_turns(uv.y * 0.5f, &st, &ct);
  float sin_theta = fmaxf(st, 1.0e-6f);
  return env.env_prob[((int64_t)e * h + r) * w + c] * (float)(w * h) / (2.0f * kPi * kPi * sin_theta);
}

}  // namespace mjpbr
