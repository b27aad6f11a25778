// geom.hpp — ray / primitive and ray / triangle intersection (geom frame math of MuJoCo's
// ray functions, as called by the reference's traverse.py:131-192 `intersect`).
//
// Each returns the distance of the nearest intersection at distance >= 0 (or -1) and the
// outward world normal there. A ray starting inside a closed primitive returns its exit point,
// whose normal faces along the ray (the caller culls such back-face hits).
#pragma once
#include "dmath.hpp"

namespace mjpbr {

// Smallest root >= 0 of a x^2 + 2 b x + c = 0, else +inf (MuJoCo's ray_quad convention).
MJ_DEV void ray_quad(float a, float b, float c, float* x0, float* x1) {
  float det = b * b - a * c;
  if (det < kMinVal) {
    *x0 = INFINITY;
    *x1 = INFINITY;
    return;
  }
  float sd = sqrtf(det);
  float r0 = (-b - sd) / a;
  float r1 = (-b + sd) / a;
  *x0 = r0 >= 0.0f ? r0 : INFINITY;
  *x1 = r1 >= 0.0f ? r1 : INFINITY;
}

MJ_DEV float ray_plane(V3 pos, const float* mat, V3 size, V3 pnt, V3 vec, V3* n) {
  V3 lp = mat_tmul(mat, pnt - pos);
  V3 lv = mat_tmul(mat, vec);
  if (lv.z > -kMinVal) return -1.0f;
  float x = -lp.z / lv.z;
  if (x < 0.0f) return -1.0f;
  float px = lp.x + x * lv.x;
  float py = lp.y + x * lv.y;
  if ((size.x <= 0.0f || fabsf(px) <= size.x) && (size.y <= 0.0f || fabsf(py) <= size.y)) {
    *n = V3{mat[2], mat[5], mat[8]};
    return x;
  }
  return -1.0f;
}

MJ_DEV float ray_sphere(V3 pos, float r2, V3 pnt, V3 vec, V3* n) {
  V3 dif = pnt - pos;
  float a = dot(vec, vec);
  float b = dot(vec, dif);
  float c = dot(dif, dif) - r2;
  float x0, x1;
  ray_quad(a, b, c, &x0, &x1);
  float x = fminf(x0, x1);
  if (x == INFINITY) return -1.0f;
  *n = normalize(pnt + vec * x - pos);
  return x;
}

MJ_DEV float ray_ellipsoid(V3 pos, const float* mat, V3 size, V3 pnt, V3 vec, V3* n) {
  V3 lp = mat_tmul(mat, pnt - pos);
  V3 lv = mat_tmul(mat, vec);
  V3 s = V3{1.0f / size.x, 1.0f / size.y, 1.0f / size.z};
  V3 p = cw(lp, s);
  V3 v = cw(lv, s);
  float a = dot(v, v);
  float b = dot(v, p);
  float c = dot(p, p) - 1.0f;
  float x0, x1;
  ray_quad(a, b, c, &x0, &x1);
  float x = fminf(x0, x1);
  if (x == INFINITY) return -1.0f;
  V3 h = lp + lv * x;
  V3 g = V3{h.x / (size.x * size.x), h.y / (size.y * size.y), h.z / (size.z * size.z)};
  *n = normalize(mat_mul(mat, g));
  return x;
}

MJ_DEV float ray_capsule(V3 pos, const float* mat, V3 size, V3 pnt, V3 vec, V3* n) {
  float r = size.x, h = size.y;
  float ssz = r + h;
  V3 nb;
  if (ray_sphere(pos, ssz * ssz, pnt, vec, &nb) < 0.0f) return -1.0f;
  V3 lp = mat_tmul(mat, pnt - pos);
  V3 lv = mat_tmul(mat, vec);
  float best = INFINITY;
  V3 bn = V3{0.0f, 0.0f, 0.0f};
  // round side
  float a = lv.x * lv.x + lv.y * lv.y;
  float b = lv.x * lp.x + lv.y * lp.y;
  float c = lp.x * lp.x + lp.y * lp.y - r * r;
  float x0, x1;
  ray_quad(a, b, c, &x0, &x1);
  if (x0 < best && fabsf(lp.z + x0 * lv.z) <= h) {
    best = x0;
    bn = V3{lp.x + x0 * lv.x, lp.y + x0 * lv.y, 0.0f};
  }
  if (x1 < best && fabsf(lp.z + x1 * lv.z) <= h) {
    best = x1;
    bn = V3{lp.x + x1 * lv.x, lp.y + x1 * lv.y, 0.0f};
  }
  // hemispherical caps
  for (int side = -1; side <= 1; side += 2) {
    V3 cen = V3{0.0f, 0.0f, (float)side * h};
    V3 dif = lp - cen;
    float ca = dot(lv, lv);
    float cb = dot(lv, dif);
    float cc = dot(dif, dif) - r * r;
    float y0, y1;
    ray_quad(ca, cb, cc, &y0, &y1);
    if (y0 < best && (float)side * (lp.z + y0 * lv.z) >= h) {
      best = y0;
      bn = lp + lv * y0 - cen;
    }
    if (y1 < best && (float)side * (lp.z + y1 * lv.z) >= h) {
      best = y1;
      bn = lp + lv * y1 - cen;
    }
  }
  if (best == INFINITY) return -1.0f;
  *n = normalize(mat_mul(mat, bn));
  return best;
}

MJ_DEV float ray_cylinder(V3 pos, const float* mat, V3 size, V3 pnt, V3 vec, V3* n) {
  float r = size.x, h = size.y;
  V3 lp = mat_tmul(mat, pnt - pos);
  V3 lv = mat_tmul(mat, vec);
  float best = INFINITY;
  V3 bn = V3{0.0f, 0.0f, 0.0f};
  float a = lv.x * lv.x + lv.y * lv.y;
  float b = lv.x * lp.x + lv.y * lp.y;
  float c = lp.x * lp.x + lp.y * lp.y - r * r;
  float x0, x1;
  ray_quad(a, b, c, &x0, &x1);
  if (x0 < best && fabsf(lp.z + x0 * lv.z) <= h) {
    best = x0;
    bn = V3{lp.x + x0 * lv.x, lp.y + x0 * lv.y, 0.0f};
  }
  if (x1 < best && fabsf(lp.z + x1 * lv.z) <= h) {
    best = x1;
    bn = V3{lp.x + x1 * lv.x, lp.y + x1 * lv.y, 0.0f};
  }
  if (fabsf(lv.z) > kMinVal) {
    for (int side = -1; side <= 1; side += 2) {
      float x = ((float)side * h - lp.z) / lv.z;
      if (x >= 0.0f && x < best) {
        float px = lp.x + x * lv.x;
        float py = lp.y + x * lv.y;
        if (px * px + py * py <= r * r) {
          best = x;
          bn = V3{0.0f, 0.0f, (float)side};
        }
      }
    }
  }
  if (best == INFINITY) return -1.0f;
  *n = normalize(mat_mul(mat, bn));
  return best;
}

MJ_DEV float ray_box(V3 pos, const float* mat, V3 size, V3 pnt, V3 vec, V3* n) {
  V3 lp = mat_tmul(mat, pnt - pos);
  V3 lv = mat_tmul(mat, vec);
  float sz[3] = {size.x, size.y, size.z};
  float p[3] = {lp.x, lp.y, lp.z};
  float v[3] = {lv.x, lv.y, lv.z};
  float best = INFINITY;
  int axis = -1;
  float sgn = 0.0f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (fabsf(v[i]) > kMinVal) {
#pragma unroll
      for (int side = -1; side <= 1; side += 2) {
        float x = ((float)side * sz[i] - p[i]) / v[i];
        if (x >= 0.0f && x < best) {
          int j = (i + 1) % 3, k = (i + 2) % 3;
          if (fabsf(p[j] + x * v[j]) <= sz[j] && fabsf(p[k] + x * v[k]) <= sz[k]) {
            best = x;
            axis = i;
            sgn = (float)side;
          }
        }
      }
    }
  }
  if (axis < 0) return -1.0f;
  V3 ln = V3{axis == 0 ? sgn : 0.0f, axis == 1 ? sgn : 0.0f, axis == 2 ? sgn : 0.0f};
  *n = mat_mul(mat, ln);
  return best;
}

// Moller-Trumbore (JGT 1997) on a triangle stored as v0 and the edges e1 = v1 - v0,
// e2 = v2 - v0; accepts 0 <= t < tmax. b1, b2 are the weights of v1 and v2.
MJ_DEV bool ray_triangle(V3 o, V3 d, V3 v0, V3 e1, V3 e2, float tmax, float* t, float* b1, float* b2) {
  V3 p = cross(d, e2);
  float det = dot(e1, p);
  if (fabsf(det) < 1.0e-20f) return false;
  // Edge tests on the unscaled numerators (sign-normalised by det), so a rejected triangle costs
  // no division; an accepted one divides once. t and the barycentrics are those of dividing
  // each numerator by det.
  V3 s = o - v0;
  float us = dot(s, p);
  float adet = fabsf(det);
  if (det < 0.0f) us = -us;
  if (us < 0.0f || us > adet) return false;
  V3 q = cross(s, e1);
  float vs = dot(d, q), ts = dot(e2, q);
  if (det < 0.0f) {
    vs = -vs;
    ts = -ts;
  }
  if (vs < 0.0f || us + vs > adet || ts < 0.0f) return false;
  float inv = 1.0f / adet;
  float tt = ts * inv;
  if (tt >= tmax) return false;
  *t = tt;
  *b1 = us * inv;
  *b2 = vs * inv;
  return true;
}

// Entry distance of the ray into a box (>= 0), or -1 when missed or not before max_t
// (traverse.py:114-124).
MJ_DEV float ray_aabb(V3 o, V3 rcp, const float* lo, const float* hi, float max_t) {
  float t1x = (lo[0] - o.x) * rcp.x, t2x = (hi[0] - o.x) * rcp.x;
  float t1y = (lo[1] - o.y) * rcp.y, t2y = (hi[1] - o.y) * rcp.y;
  float t1z = (lo[2] - o.z) * rcp.z, t2z = (hi[2] - o.z) * rcp.z;
  float tmin = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
  float tmax = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
  float entry = fmaxf(tmin, 0.0f);
  return (tmax >= entry && entry < max_t) ? entry : -1.0f;
}

// ray_aabb with the slab distances as fma(bound, rcp, -o * rcp): one instruction a slab plane
// instead of two. The caller passes rcp clamped to +-kSlabRcpMax (an infinite reciprocal of a
// zero direction component would make o * rcp - bound * rcp an inf - inf NaN) and orc = o * rcp,
// once per ray. The rounding differs from (bound - o) * rcp by about an ulp of o * rcp, far inside
// the boxes' padding (1e-5 of the extent plus 1e-6 of the coordinates); a clamped reciprocal keeps
// every slab distance's sign and puts a slab the ray does not cross at least 1e30 * (distance to
// the slab) away, beyond every cut.
constexpr float kSlabRcpMax = 1.0e30f;
MJ_DEV V3 slab_rcp(V3 rcp) {
  return V3{fminf(fmaxf(rcp.x, -kSlabRcpMax), kSlabRcpMax), fminf(fmaxf(rcp.y, -kSlabRcpMax), kSlabRcpMax),
            fminf(fmaxf(rcp.z, -kSlabRcpMax), kSlabRcpMax)};
}
MJ_DEV float ray_aabb_fma(V3 orc, V3 rcp, const float* lo, const float* hi, float max_t) {
  float t1x = fmaf(lo[0], rcp.x, -orc.x), t2x = fmaf(hi[0], rcp.x, -orc.x);
  float t1y = fmaf(lo[1], rcp.y, -orc.y), t2y = fmaf(hi[1], rcp.y, -orc.y);
  float t1z = fmaf(lo[2], rcp.z, -orc.z), t2z = fmaf(hi[2], rcp.z, -orc.z);
  float tmin = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
  float tmax = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
  float entry = fmaxf(tmin, 0.0f);
  return (tmax >= entry && entry < max_t) ? entry : -1.0f;
}

}  // namespace mjpbr
