/* rt_oracle.c — CPU oracle of the real-time tier (reference mujoco_warp_pbr/_src/render.py,
 * traverse.py, trace.py, camera.py, material.py, texture.py, envmap.py, ibl.py, lights.py,
 * brdf.py, color.py, kernel_inputs.py).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle_common.h). Scalar, one pixel at a time, and every ray
 * query is brute force: each enabled geom in index order, each mesh triangle in face order.
 * No BVH is read. Entry points take the same ABI structs as libmjpbr_hip.so with host
 * pointers.
 */
#include <stdlib.h>
#include "oracle_common.h"

/* ------------------------------------------------------------------ primitives (MuJoCo ray_*) */
static void quad_o(float a, float b, float c, float* x0, float* x1) {
  float det = b * b - a * c;
  if (det < O_MINVAL) { *x0 = INFINITY; *x1 = INFINITY; return; }
  float sd = sqrtf(det);
  float r0 = (-b - sd) / a, r1 = (-b + sd) / a;
  *x0 = r0 >= 0.0f ? r0 : INFINITY;
  *x1 = r1 >= 0.0f ? r1 : INFINITY;
}

static float plane_o(v3 pos, const float* mat, v3 size, v3 pnt, v3 vec, v3* n) {
  v3 lp = mtmul(mat, sub3(pnt, pos)), lv = mtmul(mat, vec);
  if (lv.z > -O_MINVAL) return -1.0f;
  float x = -lp.z / lv.z;
  if (x < 0.0f) return -1.0f;
  float px = lp.x + x * lv.x, py = lp.y + x * lv.y;
  if ((size.x <= 0.0f || fabsf(px) <= size.x) && (size.y <= 0.0f || fabsf(py) <= size.y)) {
    *n = mk3(mat[2], mat[5], mat[8]);
    return x;
  }
  return -1.0f;
}

static float sphere_o(v3 pos, float r2, v3 pnt, v3 vec, v3* n) {
  v3 dif = sub3(pnt, pos);
  float x0, x1;
  quad_o(dot3(vec, vec), dot3(vec, dif), dot3(dif, dif) - r2, &x0, &x1);
  float x = fminf(x0, x1);
  if (x == INFINITY) return -1.0f;
  *n = norm3(sub3(add3(pnt, mul3(vec, x)), pos));
  return x;
}

static float ellipsoid_o(v3 pos, const float* mat, v3 size, v3 pnt, v3 vec, v3* n) {
  v3 lp = mtmul(mat, sub3(pnt, pos)), lv = mtmul(mat, vec);
  v3 s = mk3(1.0f / size.x, 1.0f / size.y, 1.0f / size.z);
  v3 p = cwm(lp, s), v = cwm(lv, s);
  float x0, x1;
  quad_o(dot3(v, v), dot3(v, p), dot3(p, p) - 1.0f, &x0, &x1);
  float x = fminf(x0, x1);
  if (x == INFINITY) return -1.0f;
  v3 h = add3(lp, mul3(lv, x));
  v3 g = mk3(h.x / (size.x * size.x), h.y / (size.y * size.y), h.z / (size.z * size.z));
  *n = norm3(mmul(mat, g));
  return x;
}

static float capsule_o(v3 pos, const float* mat, v3 size, v3 pnt, v3 vec, v3* n) {
  float r = size.x, h = size.y, ssz = r + h;
  v3 tmp;
  if (sphere_o(pos, ssz * ssz, pnt, vec, &tmp) < 0.0f) return -1.0f;
  v3 lp = mtmul(mat, sub3(pnt, pos)), lv = mtmul(mat, vec);
  float best = INFINITY;
  v3 bn = mk3(0, 0, 0);
  float x[2];
  quad_o(lv.x * lv.x + lv.y * lv.y, lv.x * lp.x + lv.y * lp.y, lp.x * lp.x + lp.y * lp.y - r * r, &x[0], &x[1]);
  for (int i = 0; i < 2; ++i)
    if (x[i] < best && fabsf(lp.z + x[i] * lv.z) <= h) {
      best = x[i];
      bn = mk3(lp.x + x[i] * lv.x, lp.y + x[i] * lv.y, 0.0f);
    }
  for (int side = -1; side <= 1; side += 2) {
    v3 cen = mk3(0.0f, 0.0f, (float)side * h);
    v3 dif = sub3(lp, cen);
    float y[2];
    quad_o(dot3(lv, lv), dot3(lv, dif), dot3(dif, dif) - r * r, &y[0], &y[1]);
    for (int i = 0; i < 2; ++i)
      if (y[i] < best && (float)side * (lp.z + y[i] * lv.z) >= h) {
        best = y[i];
        bn = sub3(add3(lp, mul3(lv, y[i])), cen);
      }
  }
  if (best == INFINITY) return -1.0f;
  *n = norm3(mmul(mat, bn));
  return best;
}

static float cylinder_o(v3 pos, const float* mat, v3 size, v3 pnt, v3 vec, v3* n) {
  float r = size.x, h = size.y;
  v3 lp = mtmul(mat, sub3(pnt, pos)), lv = mtmul(mat, vec);
  float best = INFINITY;
  v3 bn = mk3(0, 0, 0);
  float x[2];
  quad_o(lv.x * lv.x + lv.y * lv.y, lv.x * lp.x + lv.y * lp.y, lp.x * lp.x + lp.y * lp.y - r * r, &x[0], &x[1]);
  for (int i = 0; i < 2; ++i)
    if (x[i] < best && fabsf(lp.z + x[i] * lv.z) <= h) {
      best = x[i];
      bn = mk3(lp.x + x[i] * lv.x, lp.y + x[i] * lv.y, 0.0f);
    }
  if (fabsf(lv.z) > O_MINVAL) {
    for (int side = -1; side <= 1; side += 2) {
      float t = ((float)side * h - lp.z) / lv.z;
      if (t >= 0.0f && t < best) {
        float px = lp.x + t * lv.x, py = lp.y + t * lv.y;
        if (px * px + py * py <= r * r) { best = t; bn = mk3(0.0f, 0.0f, (float)side); }
      }
    }
  }
  if (best == INFINITY) return -1.0f;
  *n = norm3(mmul(mat, bn));
  return best;
}

static float box_o(v3 pos, const float* mat, v3 size, v3 pnt, v3 vec, v3* n) {
  v3 lp = mtmul(mat, sub3(pnt, pos)), lv = mtmul(mat, vec);
  float sz[3] = {size.x, size.y, size.z}, p[3] = {lp.x, lp.y, lp.z}, v[3] = {lv.x, lv.y, lv.z};
  float best = INFINITY, sgn = 0.0f;
  int axis = -1;
  for (int i = 0; i < 3; ++i) {
    if (!(fabsf(v[i]) > O_MINVAL)) continue;
    for (int side = -1; side <= 1; side += 2) {
      float x = ((float)side * sz[i] - p[i]) / v[i];
      if (x >= 0.0f && x < best) {
        int j = (i + 1) % 3, k = (i + 2) % 3;
        if (fabsf(p[j] + x * v[j]) <= sz[j] && fabsf(p[k] + x * v[k]) <= sz[k]) { best = x; axis = i; sgn = (float)side; }
      }
    }
  }
  if (axis < 0) return -1.0f;
  *n = mmul(mat, mk3(axis == 0 ? sgn : 0.0f, axis == 1 ? sgn : 0.0f, axis == 2 ? sgn : 0.0f));
  return best;
}

/* Moller & Trumbore (1997), "Fast, minimum storage ray-triangle intersection", as published:
 * solve o + t d = v0 + b1 e1 + b2 e2 by Cramer's rule with one division by the determinant.
 * Double-sided (culling is the caller's); edges and vertices count as hits; 0 <= t < tmax. */
static int tri_o(v3 o, v3 d, v3 v0, v3 e1, v3 e2, float tmax, float* t, float* b1, float* b2) {
  v3 p = cross3(d, e2);
  float det = dot3(e1, p);
  if (fabsf(det) < 1.0e-20f) return 0; /* ray parallel to the triangle's plane */
  float inv_det = 1.0f / det;
  v3 s = sub3(o, v0);
  float u = dot3(s, p) * inv_det;
  if (u < 0.0f || u > 1.0f) return 0;
  v3 q = cross3(s, e1);
  float v = dot3(d, q) * inv_det;
  if (v < 0.0f || u + v > 1.0f) return 0;
  float tt = dot3(e2, q) * inv_det;
  if (tt < 0.0f || tt >= tmax) return 0;
  *t = tt; *b1 = u; *b2 = v;
  return 1;
}

/* ------------------------------------------------------------------ brute-force ray queries */
typedef struct {
  int geom_id;
  float dist;
  v3 normal;
  float bary_u, bary_v;
  int face_id, mesh_id;
} HitO;

typedef struct {
  const MjpbrScene* sc;
  const MjpbrSettings* s;
  int w;
  int seed;
  float exposure;
} Ctx;

static int bidx(int w, int n) { return w % n; }

static void geom_rgba_o(const MjpbrGeometry* g, int w, int gi, float rgba[4], int* mat_id) {
  int mid = g->geom_matid[bidx(w, g->geom_matid_n) * g->ngeom + gi];
  const float* r = g->geom_rgba + ((int64_t)bidx(w, g->geom_rgba_n) * g->ngeom + gi) * 4;
  for (int k = 0; k < 4; ++k) rgba[k] = r[k];
  if (mid >= 0 && r[0] == 0.5f && r[1] == 0.5f && r[2] == 0.5f && r[3] == 1.0f) {
    const float* m = g->mat_rgba + ((int64_t)bidx(w, g->mat_rgba_n) * g->nmat + mid) * 4;
    for (int k = 0; k < 4; ++k) rgba[k] = m[k];
  }
  *mat_id = mid;
}

static float cut_o(float dist) { return dist * (1.0f + O_TOL_REL) + O_TOL_ABS; }

/* Rotation of geom gi as traced: geom_xmat, or geom_xmat R^T for a mesh / SDF geom whose mesh
 * BVH is built in frame R (reference trace.py:150-165, geom_xmat_trace). */
static const float* traced_mat_o(const MjpbrGeometry* g, int w, int gi, float* tmp) {
  const float* M = g->geom_xmat + ((int64_t)w * g->ngeom + gi) * 9;
  int type = g->geom_type[gi];
  int dataid = g->geom_dataid[bidx(w, g->geom_dataid_n) * g->ngeom + gi];
  if (!g->mesh_frame || dataid < 0 || (type != MJPBR_GEOM_MESH && type != MJPBR_GEOM_SDF)) return M;
  const float* R = g->mesh_frame + 9 * dataid;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) tmp[3 * i + j] = M[3 * i] * R[3 * j] + M[3 * i + 1] * R[3 * j + 1] + M[3 * i + 2] * R[3 * j + 2];
  return tmp;
}

/* One geom (traverse.py:131-192). any_hit: meshes report 0 on any triangle before max_t. */
static float intersect_o(const Ctx* c, int gi, v3 o, v3 d, float max_t, int any_hit, v3* n, float* bu, float* bv,
                         int* face, int* mesh) {
  const MjpbrGeometry* g = &c->sc->geo;
  const MjpbrMaterials* m = &c->sc->mats;
  int w = c->w;
  int type = g->geom_type[gi];
  v3 pos = load3(g->geom_xpos + ((int64_t)w * g->ngeom + gi) * 3);
  float framed[9];
  const float* mat = traced_mat_o(g, w, gi, framed);
  v3 size = load3(g->geom_size + ((int64_t)bidx(w, g->geom_size_n) * g->ngeom + gi) * 3);
  switch (type) {
    case MJPBR_GEOM_PLANE: return plane_o(pos, mat, size, o, d, n);
    case MJPBR_GEOM_SPHERE: return sphere_o(pos, size.x * size.x, o, d, n);
    case MJPBR_GEOM_ELLIPSOID: return ellipsoid_o(pos, mat, size, o, d, n);
    case MJPBR_GEOM_CAPSULE: return capsule_o(pos, mat, size, o, d, n);
    case MJPBR_GEOM_CYLINDER: return cylinder_o(pos, mat, size, o, d, n);
    case MJPBR_GEOM_BOX: return box_o(pos, mat, size, o, d, n);
    case MJPBR_GEOM_MESH:
    case MJPBR_GEOM_SDF:
    case MJPBR_GEOM_HFIELD: {
      /* SDF geoms trace their mesh; height field h traces its triangle surface, mesh nmesh + h,
         and shades as a primitive (traverse.py:161-165, 185-196). */
      int mid = g->geom_dataid[bidx(w, g->geom_dataid_n) * g->ngeom + gi];
      if (mid < 0) return -1.0f;
      if (type == MJPBR_GEOM_HFIELD) mid += g->nmesh;
      v3 lo = mtmul(mat, sub3(o, pos)), ld = mtmul(mat, d);
      int fadr = m->mesh_faceadr[mid], vadr = m->mesh_vertadr[mid];
      /* Conservative early out: a ray missing the (inflated) mesh-frame bounds hits no triangle. */
      {
        const float* a = g->mesh_aabb + 6 * mid;
        float lo3[3] = {lo.x, lo.y, lo.z}, ld3[3] = {ld.x, ld.y, ld.z}, tmin = 0.0f, tmax = INFINITY;
        int miss = 0;
        for (int k = 0; k < 3 && !miss; ++k) {
          float ext = a[3 + k] - a[k];
          float pad = 1.0e-3f * ext + 1.0e-4f * (fabsf(a[k]) + fabsf(a[3 + k])) + 1.0e-5f;
          float l = a[k] - pad, h = a[3 + k] + pad;
          if (ld3[k] == 0.0f) {
            if (lo3[k] < l || lo3[k] > h) miss = 1;
          } else {
            double t1 = ((double)l - lo3[k]) / ld3[k], t2 = ((double)h - lo3[k]) / ld3[k];
            if (t1 > t2) { double t = t1; t1 = t2; t2 = t; }
            if (t1 > tmin) tmin = (float)t1;
            if (t2 < tmax) tmax = (float)t2;
            if (tmax < tmin) miss = 1;
          }
        }
        if (miss) return -1.0f;
      }
      int nf = m->mesh_facenum[mid];
      float best = INFINITY, bb1 = 0, bb2 = 0;
      int best_face = -1;
      v3 be1 = mk3(0, 0, 0), be2 = mk3(0, 0, 0);
      for (int f = 0; f < nf; ++f) {
        const int* fv = m->mesh_face + 3 * (fadr + f);
        v3 p0 = load3(m->mesh_vert + 3 * (vadr + fv[0]));
        v3 e1 = sub3(load3(m->mesh_vert + 3 * (vadr + fv[1])), p0);
        v3 e2 = sub3(load3(m->mesh_vert + 3 * (vadr + fv[2])), p0);
        float t, b1, b2;
        if (!tri_o(lo, ld, p0, e1, e2, max_t, &t, &b1, &b2)) continue;
        if (any_hit) { *n = mk3(0, 0, 0); return 0.0f; }
        if (t < best || (t == best && f < best_face)) { best = t; best_face = f; bb1 = b1; bb2 = b2; be1 = e1; be2 = e2; }
      }
      if (any_hit || best_face < 0) return -1.0f;
      *n = norm3(mmul(mat, norm3(cross3(be1, be2))));
      *bu = 1.0f - bb1 - bb2;
      *bv = bb1;
      *face = best_face;
      *mesh = mid;
      return best;
    }
    default: return -1.0f;
  }
}

static int geom_skipped(const Ctx* c, int gi) {
  float rgba[4];
  int mid;
  geom_rgba_o(&c->sc->geo, c->w, gi, rgba, &mid);
  return rgba[3] == 0.0f;
}

/* The k-th enabled geom in index order: the enabled set is bvh_geom and direct_geom together,
 * both ascending (the renderer tests the direct geoms before its BVH; brute force keeps index
 * order). */
static int enabled_geom_o(const MjpbrGeometry* g, int k, int* ib, int* id) {
  if (*id >= g->ndirect || (*ib < g->nbvh && g->bvh_geom[*ib] < g->direct_geom[*id])) return g->bvh_geom[(*ib)++];
  return g->direct_geom[(*id)++];
}

/* traverse.py:194-323 with first_hit = False, all geoms in index order. */
static HitO closest_o(const Ctx* c, v3 o, v3 d, float max_dist) {
  const MjpbrGeometry* g = &c->sc->geo;
  int cull = c->s->cull_backfaces;
  HitO h = {-1, max_dist, {0, 0, 0}, 0, 0, -1, -1};
  int ib = 0, id = 0;
  for (int k = 0; k < g->nbvh + g->ndirect; ++k) {
    int gi = enabled_geom_o(g, k, &ib, &id);
    if (geom_skipped(c, gi)) continue;
    v3 n = mk3(0, 0, 0);
    float bu = 0, bv = 0;
    int face = -1, mesh = -1;
    float dd = intersect_o(c, gi, o, d, cut_o(h.dist), 0, &n, &bu, &bv, &face, &mesh);
    if (cull && dd >= 0.0f && dot3(d, n) > 0.0f) dd = -1.0f;
    float tol = O_TOL_REL * fmaxf(h.dist, 1.0f);
    int closer = (dd < h.dist - tol) || (fabsf(dd - h.dist) <= tol && (h.geom_id < 0 || gi > h.geom_id));
    if (dd >= 0.0f && closer) {
      h.dist = dd; h.normal = n; h.geom_id = gi; h.bary_u = bu; h.bary_v = bv; h.face_id = face; h.mesh_id = mesh;
    }
  }
  return h;
}

static int any_o(const Ctx* c, v3 o, v3 d, float max_dist) {
  const MjpbrGeometry* g = &c->sc->geo;
  int cull = c->s->cull_backfaces;
  int ib = 0, id = 0;
  for (int k = 0; k < g->nbvh + g->ndirect; ++k) {
    int gi = enabled_geom_o(g, k, &ib, &id);
    if (geom_skipped(c, gi)) continue;
    v3 n = mk3(0, 0, 0);
    float bu, bv;
    int face, mesh;
    float dd = intersect_o(c, gi, o, d, max_dist, 1, &n, &bu, &bv, &face, &mesh);
    if (cull && dd >= 0.0f && dot3(d, n) > 0.0f) dd = -1.0f;
    if (dd >= 0.0f && dd < max_dist) return 1;
  }
  return 0;
}

/* ------------------------------------------------------------------ textures (bilinear/trilinear) */
typedef struct { float x, y, z, w; } f4;

static f4 texel_o(const uint8_t* base, int fmt, int idx) {
  f4 r;
  if (fmt == MJPBR_TEX_RGBA8) {
    const uint8_t* t = base + 4 * idx;
    const float k = 1.0f / 255.0f;
    r.x = t[0] * k; r.y = t[1] * k; r.z = t[2] * k; r.w = t[3] * k;
  } else {
    const uint16_t* t = (const uint16_t*)(base) + 4 * idx;
    r.x = half2f(t[0]); r.y = half2f(t[1]); r.z = half2f(t[2]); r.w = half2f(t[3]);
  }
  return r;
}

static f4 lerp4o(f4 a, f4 b, float t) {
  float s = 1.0f - t;
  f4 r = {a.x * s + b.x * t, a.y * s + b.y * t, a.z * s + b.z * t, a.w * s + b.w * t};
  return r;
}

static int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

static f4 bilerp_o(const uint8_t* base, int fmt, int w, int h, float u, float v, int wrap_v) {
  float uu = u - floorf(u);
  float x = uu * (float)w - 0.5f;
  float x0f = floorf(x), fx = x - x0f;
  int x0 = (int)x0f;
  int i0 = x0 < 0 ? w - 1 : x0, i1 = x0 + 1 >= w ? 0 : x0 + 1;
  float vv = wrap_v ? v - floorf(v) : v;
  float y = vv * (float)h - 0.5f;
  float y0f = floorf(y), fy = y - y0f;
  int y0 = (int)y0f, j0, j1;
  if (wrap_v) { j0 = y0 < 0 ? h - 1 : y0; j1 = y0 + 1 >= h ? 0 : y0 + 1; }
  else { j0 = clampi(y0, 0, h - 1); j1 = clampi(y0 + 1, 0, h - 1); }
  f4 t00 = texel_o(base, fmt, j0 * w + i0), t10 = texel_o(base, fmt, j0 * w + i1);
  f4 t01 = texel_o(base, fmt, j1 * w + i0), t11 = texel_o(base, fmt, j1 * w + i1);
  return lerp4o(lerp4o(t00, t10, fx), lerp4o(t01, t11, fx), fy);
}

static f4 mip_o(const uint8_t* pool, const int64_t* level_adr, const int32_t* desc, int scale, float u, float v,
                float lod, int wrap_v) {
  int w0 = desc[0], h0 = desc[1], nl = desc[2], fmt = desc[3];
  lod = clampf_(lod, 0.0f, (float)(nl - 1));
  float l0f = floorf(lod), f = lod - l0f;
  int l0 = (int)l0f, l1 = l0 + 1 < nl ? l0 + 1 : nl - 1;
  int w = w0 >> l0, h = h0 >> l0;
  f4 c0 = bilerp_o(pool + level_adr[l0] * scale, fmt, w > 1 ? w : 1, h > 1 ? h : 1, u, v, wrap_v);
  if (f > 0.0f) {
    int w1 = w0 >> l1, h1 = h0 >> l1;
    f4 c1 = bilerp_o(pool + level_adr[l1] * scale, fmt, w1 > 1 ? w1 : 1, h1 > 1 ? h1 : 1, u, v, wrap_v);
    c0 = lerp4o(c0, c1, f);
  }
  return c0;
}

static f4 tex_sample_o(const MjpbrMaterials* m, int tex, v2 uv, float lod) {
  return mip_o(m->tex_data, m->tex_level_adr + tex * MJPBR_MAX_LEVELS, m->tex_desc + 4 * tex, 1, uv.x, uv.y,
               fmaxf(lod, 0.0f), 1);
}

/* texture.py:224-228 */
static float tex_lod_o(const MjpbrMaterials* m, int tex, float footprint, float upl) {
  int a = m->tex_desc[4 * tex], b = m->tex_desc[4 * tex + 1];
  float texels = footprint * upl * (float)(a > b ? a : b);
  return log2f(fmaxf(texels, 1.0e-8f));
}

/* ------------------------------------------------------------------ environment (envmap.py) */
static v2 dir_to_uv_o(v3 d, float yaw) {
  float phi = atan2f(d.y, d.x) - yaw;
  float u = 0.5f + phi * (0.5f * O_INV_PI);
  float v = acosf(clampf_(d.z, -1.0f, 1.0f)) * O_INV_PI;
  v2 r = {u - floorf(u), v};
  return r;
}

static v3 uv_to_dir_o(v2 uv) {
  float sp, cp, st, ct;
  sincos_turns_o(uv.x - 0.5f, &sp, &cp);
  sincos_turns_o(uv.y * 0.5f, &st, &ct);
  return mk3(st * cp, st * sp, ct);
}

static v3 rot_yaw_o(v3 d, float yaw) {
  float c = cosf(yaw), s = sinf(yaw);
  return mk3(c * d.x - s * d.y, s * d.x + c * d.y, d.z);
}

static v3 radiance_o(const MjpbrEnvironment* env, int e, v3 dir, float yaw, float lod) {
  v2 uv = dir_to_uv_o(dir, yaw);
  f4 c = mip_o((const uint8_t*)env->env_radiance, env->env_level_adr + e * MJPBR_MAX_LEVELS, env->env_desc + 4 * e, 2,
               uv.x, uv.y, fmaxf(lod, 0.0f), 0);
  return mk3(c.x, c.y, c.z);
}

static v3 prefiltered_o(const MjpbrEnvironment* env, int e, v3 dir, float yaw, float pr) {
  v2 uv = dir_to_uv_o(dir, yaw);
  int L = env->pf_levels, H = env->pf_height, W = env->pf_width;
  float z = clampf_(pr, 0.0f, 1.0f) * (float)(L - 1);
  float z0f = floorf(z), fz = z - z0f;
  int z0 = (int)z0f < L - 1 ? (int)z0f : L - 1;
  int z1 = z0 + 1 < L - 1 ? z0 + 1 : L - 1;
  const uint8_t* base = (const uint8_t*)(env->env_prefiltered + (int64_t)e * L * H * W * 4);
  f4 c = bilerp_o(base + (int64_t)z0 * H * W * 8, MJPBR_TEX_RGBA16F, W, H, uv.x, uv.y, 0);
  if (fz > 0.0f) c = lerp4o(c, bilerp_o(base + (int64_t)z1 * H * W * 8, MJPBR_TEX_RGBA16F, W, H, uv.x, uv.y, 0), fz);
  return mk3(c.x, c.y, c.z);
}

/* envmap.py:281-289 */
static v3 specular_env_o(const MjpbrEnvironment* env, int e, v3 dir, float yaw, float pr) {
  v3 blurred = prefiltered_o(env, e, dir, yaw, pr);
  float t = pr * (float)(env->pf_levels - 1);
  if (t >= 1.0f) return blurred;
  v3 sharp = radiance_o(env, e, dir, yaw, 0.0f);
  return add3(mul3(sharp, 1.0f - t), mul3(blurred, t));
}

/* envmap.py:252-266 */
static v3 sh9_o(const float* sh, v3 n) {
  float x = n.x, y = n.y, z = n.z;
  float b[9] = {0.282095f, 0.488603f * y, 0.488603f * z, 0.488603f * x, 1.092548f * x * y, 1.092548f * y * z,
                0.315392f * (3.0f * z * z - 1.0f), 1.092548f * x * z, 0.546274f * (x * x - y * y)};
  v3 c = mk3(0, 0, 0);
  for (int k = 0; k < 9; ++k) c = add3(c, mul3(load3(sh + 3 * k), b[k]));
  return max3s(c, 0.0f);
}

/* ibl.py:117-119 */
static v2 dfg_o(const MjpbrEnvironment* env, float nov, float pr) {
  int N = env->dfg_size;
  float x = nov * (float)N - 0.5f, y = pr * (float)N - 0.5f;
  float x0f = floorf(x), y0f = floorf(y), fx = x - x0f, fy = y - y0f;
  int x0 = (int)x0f, y0 = (int)y0f;
  int i0 = clampi(x0, 0, N - 1), i1 = clampi(x0 + 1, 0, N - 1), j0 = clampi(y0, 0, N - 1), j1 = clampi(y0 + 1, 0, N - 1);
  const float* L = env->dfg_lut;
  const float *a = L + 2 * (j0 * N + i0), *b = L + 2 * (j0 * N + i1), *c = L + 2 * (j1 * N + i0), *d = L + 2 * (j1 * N + i1);
  float sx = 1.0f - fx, sy = 1.0f - fy;
  v2 r = {(a[0] * sx + b[0] * fx) * sy + (c[0] * sx + d[0] * fx) * fy, (a[1] * sx + b[1] * fx) * sy + (c[1] * sx + d[1] * fx) * fy};
  return r;
}

static void env_lookup_o(const MjpbrEnvironment* env, int w, int* e, float* yaw, float* scale) {
  *e = env->env_id[w % env->env_id_n];
  *yaw = env->env_yaw[w % env->env_yaw_n];
  *scale = env->env_intensity[w % env->env_intensity_n];
}

/* ------------------------------------------------------------------ BRDF (brdf.py:32-110) */
static float d_ggx_o(float noh, float alpha) {
  float a2 = alpha * alpha;
  float f = (noh * a2 - noh) * noh + 1.0f;
  return a2 / (O_PI * f * f);
}
static float v_smith_o(float nov, float nol, float alpha) {
  float a2 = alpha * alpha;
  float gv = nol * sqrtf(nov * nov * (1.0f - a2) + a2);
  float gl = nov * sqrtf(nol * nol * (1.0f - a2) + a2);
  return 0.5f / (gv + gl);
}
static v3 schlick_o(v3 f0, float f90, float voh) {
  float x = 1.0f - voh, x2 = x * x, fc = x2 * x2 * x;
  return add3(mul3(f0, 1.0f - fc), mk3(f90 * fc, f90 * fc, f90 * fc));
}
static v3 spec_lobe_o(v3 f0, float alpha, float noh, float nov, float nol, float loh) {
  return mul3(schlick_o(f0, 1.0f, loh), d_ggx_o(noh, alpha) * v_smith_o(nov, nol, alpha));
}

typedef struct {
  v3 base_color;
  float alpha, metallic, pr, reflectance;
  v3 emissive;
  float occlusion;
} SurfO;

typedef struct {
  float pr, alpha;
  v3 f0, energy_comp, spec_albedo, diffuse_color;
} LobesO;

/* material.py:527-539 */
static LobesO lobes_o(const MjpbrEnvironment* env, const SurfO* s, float nov) {
  LobesO L;
  L.pr = clampf_(s->pr, 0.045f, 1.0f);
  L.alpha = L.pr * L.pr;
  float dielectric = 0.16f * s->reflectance * s->reflectance * (1.0f - s->metallic);
  L.f0 = add3(mul3(s->base_color, s->metallic), mk3(dielectric, dielectric, dielectric));
  v2 dfg = dfg_o(env, nov, L.pr);
  float k = 1.0f / fmaxf(dfg.y, 1.0e-4f) - 1.0f;
  L.energy_comp = add3(mk3(1, 1, 1), mul3(L.f0, k));
  L.spec_albedo = add3(mul3(sub3(mk3(1, 1, 1), L.f0), dfg.x), mul3(L.f0, dfg.y));
  L.diffuse_color = cwm(mul3(s->base_color, 1.0f - s->metallic), sub3(mk3(1, 1, 1), L.spec_albedo));
  return L;
}

/* ------------------------------------------------------------------ lights (lights.py) */
typedef struct {
  int type, castshadow;
  v3 pos, dir, diffuse, specular, attenuation;
  float cutoff, exponent, intensity, range, bulbradius;
} LightO;

static LightO fetch_light_o(const MjpbrLights* lt, int w, int li) {
  LightO L;
  int n = lt->nlight;
#define B(name, k) (lt->name[(int64_t)(w % lt->name##_n) * n * (k) + (int64_t)li * (k)])
  L.type = B(light_type, 1);
  L.castshadow = B(light_castshadow, 1) != 0;
  L.pos = load3(lt->light_xpos + ((int64_t)w * n + li) * 3);
  L.dir = load3(lt->light_xdir + ((int64_t)w * n + li) * 3);
  L.diffuse = load3(&B(light_diffuse, 3));
  L.specular = load3(&B(light_specular, 3));
  L.attenuation = load3(&B(light_attenuation, 3));
  L.cutoff = B(light_cutoff, 1);
  L.exponent = B(light_exponent, 1);
  L.intensity = B(light_intensity, 1);
  L.range = B(light_range, 1);
  L.bulbradius = B(light_bulbradius, 1);
#undef B
  return L;
}

static void eval_light_o(const LightO* L, int model, int has_spot, v3 point, v3* l, float* dist, v3* ds, v3* ss) {
  int physical = model == 0;
  float d = 1.0e10f, att = 1.0f;
  if (L->type == MJPBR_LIGHT_DIRECTIONAL) {
    *l = neg3(L->dir);
  } else {
    v3 to = sub3(L->pos, point);
    d = length3(to);
    *l = div3(to, fmaxf(d, 1.0e-9f));
    if (physical) {
      float d2 = d * d, r = fmaxf(L->range, 1.0e-6f);
      float factor = d2 / (r * r);
      float window = clampf_(1.0f - factor * factor, 0.0f, 1.0f);
      att = window * window / fmaxf(d2, 1.0e-4f);
    } else {
      att = 1.0f / fmaxf(dot3(L->attenuation, mk3(1.0f, d, d * d)), 1.0e-9f);
    }
    if (has_spot && L->type == MJPBR_LIGHT_SPOT) {
      float cos_theta = dot3(neg3(*l), L->dir);
      float outer = L->cutoff * (O_PI / 180.0f);
      float cos_outer = cosf(outer);
      if (physical) {
        float cos_inner = cosf(outer * 0.75f);
        float cs = 1.0f / fmaxf(cos_inner - cos_outer, 1.0f / 1024.0f);
        float s = clampf_((cos_theta - cos_outer) * cs, 0.0f, 1.0f);
        att = att * s * s;
      } else {
        att = cos_theta < cos_outer ? 0.0f : att * powf(fmaxf(cos_theta, 0.0f), L->exponent);
      }
    }
  }
  *dist = d;
  if (physical) {
    v3 sc = mul3(L->diffuse, L->intensity * att);
    *ds = sc; *ss = sc;
  } else {
    float k = O_PI * att;
    *ds = mul3(L->diffuse, k); *ss = mul3(L->specular, k);
  }
}

/* ------------------------------------------------------------------ colour (color.py) */
static v3 pbr_neutral_o(v3 c) {
  const float start = 0.8f - 0.04f, desat = 0.15f;
  float x = fminf(c.x, fminf(c.y, c.z));
  float offset = x < 0.08f ? x - 6.25f * x * x : 0.04f;
  c = sub3(c, mk3(offset, offset, offset));
  float peak = fmaxf(c.x, fmaxf(c.y, c.z));
  if (peak < start) return c;
  float d = 1.0f - start;
  float np = 1.0f - d * d / (peak + d - start);
  c = mul3(c, np / peak);
  float g = 1.0f - 1.0f / (desat * (peak - np) + 1.0f);
  return add3(mul3(c, 1.0f - g), mk3(np * g, np * g, np * g));
}

static v3 aces_o(v3 c) {
  v3 v = mk3(0.59719f * c.x + 0.35458f * c.y + 0.04823f * c.z, 0.07600f * c.x + 0.90834f * c.y + 0.01566f * c.z,
             0.02840f * c.x + 0.13383f * c.y + 0.83777f * c.z);
  v3 num = sub3(cwm(v, add3(v, mk3(0.0245786f, 0.0245786f, 0.0245786f))), mk3(0.000090537f, 0.000090537f, 0.000090537f));
  v3 den = add3(cwm(v, add3(mul3(v, 0.983729f), mk3(0.4329510f, 0.4329510f, 0.4329510f))), mk3(0.238081f, 0.238081f, 0.238081f));
  v = mk3(num.x / den.x, num.y / den.y, num.z / den.z);
  return mk3(1.60475f * v.x - 0.53108f * v.y - 0.07367f * v.z, -0.10208f * v.x + 1.10813f * v.y - 0.00605f * v.z,
             -0.00327f * v.x - 0.07276f * v.y + 1.07602f * v.z);
}

static v3 display_o(v3 c, int tm) {
  c = max3s(c, 0.0f);
  if (tm == 0) c = pbr_neutral_o(c);
  else if (tm == 1) c = aces_o(c);
  return min3s(max3s(c, 0.0f), 1.0f);
}

static float oetf_o(float x) { return x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f; }

static uint32_t encode_o(v3 c) {
  uint32_t r = (uint32_t)(oetf_o(c.x) * 255.0f + 0.5f), g = (uint32_t)(oetf_o(c.y) * 255.0f + 0.5f),
           b = (uint32_t)(oetf_o(c.z) * 255.0f + 0.5f);
  return (255u << 24) | (r << 16) | (g << 8) | b; /* 0xAARRGGBB (MuJoCo Warp pack_rgba_to_uint32) */
}

/* ------------------------------------------------------------------ surfaces (material.py:300-501) */
static float widen_o(float pr, float extra) {
  pr = clampf_(pr, 0.0f, 1.0f);
  float a2 = fminf(pr * pr * pr * pr + extra, 1.0f);
  return sqrtf(sqrtf(a2));
}

static void basis_o(v3 n, v3* t1, v3* t2) {
  float sign = n.z >= 0.0f ? 1.0f : -1.0f;
  float a = -1.0f / (sign + n.z);
  float b = n.x * n.y * a;
  *t1 = mk3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  *t2 = mk3(b, sign + n.y * n.y * a, -n.y);
}

static v2 wrapd_o(v2 d) { v2 r = {d.x - rintf(d.x), d.y - rintf(d.y)}; return r; }

static void prim_uv_o(int type, v3 p, v3 ln, v3 size, v2 repeat, int uniform, v2* uv, v2* rep, v2* off) {
  *rep = repeat;
  off->x = 0.0f; off->y = 0.0f;
  v2 r = {0.0f, 0.0f};
  if (type == MJPBR_GEOM_PLANE || type == MJPBR_GEOM_HFIELD) {
    if (uniform || size.x <= 0.0f || size.y <= 0.0f) { r.x = p.x; r.y = -p.y; }
    else { r.x = p.x / (2.0f * size.x) + 0.5f; r.y = 0.5f - p.y / (2.0f * size.y); }
  } else if (type == MJPBR_GEOM_SPHERE || type == MJPBR_GEOM_ELLIPSOID) {
    v3 q = type == MJPBR_GEOM_SPHERE ? p : mk3(p.x / size.x, p.y / size.y, p.z / size.z);
    float rr = length3(q);
    r.x = atan2f(q.y, q.x) * (0.5f * O_INV_PI) + 0.5f;
    r.y = acosf(clampf_(rr > 0.0f ? q.z / rr : 1.0f, -1.0f, 1.0f)) * O_INV_PI;
  } else if (type == MJPBR_GEOM_CYLINDER) {
    if (fabsf(ln.z) > 0.5f) { r.x = p.x / (2.0f * size.x) + 0.5f; r.y = 0.5f - p.y / (2.0f * size.x); }
    else { r.x = atan2f(p.y, p.x) * (0.5f * O_INV_PI) + 0.5f; r.y = 0.5f - p.z / (2.0f * size.y); }
  } else if (type == MJPBR_GEOM_CAPSULE) {
    r.x = atan2f(p.y, p.x) * (0.5f * O_INV_PI) + 0.5f;
    r.y = 0.5f - p.z / (2.0f * (size.y + size.x));
  } else if (type == MJPBR_GEOM_BOX) {
    float ax = fabsf(ln.x), ay = fabsf(ln.y), az = fabsf(ln.z), a, b, sa, sb;
    if (ax >= ay && ax >= az) { a = p.y; b = p.z; sa = size.y; sb = size.z; }
    else if (ay >= az) { a = p.x; b = p.z; sa = size.x; sb = size.z; }
    else { a = p.x; b = p.y; sa = size.x; sb = size.y; }
    if (uniform) { r.x = a; r.y = -b; }
    else { r.x = a / (2.0f * sa) + 0.5f; r.y = 0.5f - b / (2.0f * sb); }
  }
  *uv = r;
}

static f4 role_o(const MjpbrMaterials* m, int tex, v2 uv, float cone, float upl) {
  return tex_sample_o(m, tex, uv, tex_lod_o(m, tex, cone, upl));
}

static SurfO eval_surface_o(const Ctx* c, const HitO* hit, v3 point, v3 dir, float footprint, v3* shading_world) {
  const MjpbrGeometry* g = &c->sc->geo;
  const MjpbrMaterials* m = &c->sc->mats;
  const MjpbrSettings* s = c->s;
  int w = c->w, gid = hit->geom_id, ng = g->ngeom;
  SurfO sf;
  sf.occlusion = 1.0f;
  sf.emissive = mk3(0, 0, 0);
  float rgba[4];
  int mat_id;
  geom_rgba_o(g, w, gid, rgba, &mat_id);
  sf.base_color = mk3(rgba[0], rgba[1], rgba[2]);
  sf.alpha = rgba[3];
  sf.metallic = m->geom_metallic[(w % m->geom_metallic_n) * ng + gid];
  sf.pr = m->geom_roughness[(w % m->geom_roughness_n) * ng + gid];
  sf.reflectance = m->geom_reflectance[(w % m->geom_reflectance_n) * ng + gid];
  float emission = mat_id >= 0 ? m->mat_emission[(w % m->mat_emission_n) * g->nmat + mat_id] : 0.0f;
  int gtype = g->geom_type[gid];
  float framed[9];
  const float* R = traced_mat_o(g, w, gid, framed);
  v3 ldir = mtmul(R, dir);
  v3 lnrm = mtmul(R, hit->normal);
  if (dot3(lnrm, ldir) > 0.0f) lnrm = neg3(lnrm);
  v3 shading = lnrm;
  float curvature = 0.0f;
  int is_mesh = gtype == MJPBR_GEOM_MESH || gtype == MJPBR_GEOM_SDF;
  int mesh_id = hit->mesh_id, face_id = hit->face_id;
  if (is_mesh && s->vertex_normals && mesh_id >= 0 && face_id >= 0) {
    int fadr = m->mesh_faceadr[mesh_id] + face_id;
    const int* tri = m->mesh_facenormal + 3 * fadr;
    int nadr = m->mesh_normaladr[mesh_id];
    v3 n[3];
    for (int k = 0; k < 3; ++k) {
      n[k] = load3(m->mesh_normal + 3 * (nadr + tri[k]));
      if (dot3(n[k], lnrm) < 0.0f) n[k] = neg3(n[k]);
    }
    shading = norm3(add3(add3(mul3(n[0], hit->bary_u), mul3(n[1], hit->bary_v)), mul3(n[2], 1.0f - hit->bary_u - hit->bary_v)));
    if (dot3(shading, ldir) > 0.0f) shading = neg3(shading);
    if (s->specular_aa) {
      const int* f = m->mesh_face + 3 * fadr;
      int vadr = m->mesh_vertadr[mesh_id];
      v3 p0 = load3(m->mesh_vert + 3 * (vadr + f[0]));
      float edge = sqrtf(length3(cross3(sub3(load3(m->mesh_vert + 3 * (vadr + f[1])), p0), sub3(load3(m->mesh_vert + 3 * (vadr + f[2])), p0))));
      float spread = fmaxf(length3(sub3(n[0], n[1])), fmaxf(length3(sub3(n[1], n[2])), length3(sub3(n[2], n[0]))));
      curvature = spread / fmaxf(edge, 1.0e-9f);
    }
  }
  v3 size = mk3(0, 0, 0);
  if (!is_mesh) size = load3(g->geom_size + ((int64_t)(w % g->geom_size_n) * ng + gid) * 3);
  if (!is_mesh && s->specular_aa) {
    if (gtype == MJPBR_GEOM_SPHERE || gtype == MJPBR_GEOM_CAPSULE || gtype == MJPBR_GEOM_CYLINDER)
      curvature = 1.0f / fmaxf(size.x, 1.0e-9f);
    else if (gtype == MJPBR_GEOM_ELLIPSOID)
      curvature = 1.0f / fmaxf(fminf(size.x, fminf(size.y, size.z)), 1.0e-9f);
  }
  float cone = footprint / fmaxf(fabsf(dot3(lnrm, ldir)), 0.05f);
  if (s->textures && mat_id >= 0) {
    int mw = w % m->mat_texid_n;
    const float* rp = m->mat_texrepeat + ((int64_t)(w % m->mat_texrepeat_n) * g->nmat + mat_id) * 2;
    v2 repeat = {rp[0], rp[1]};
    int uniform = m->mat_texuniform[(w % m->mat_texuniform_n) * g->nmat + mat_id] != 0;
    int has_uv = 0;
    v2 uv = {0, 0};
    v3 tangent = mk3(0, 0, 0), bitangent = mk3(0, 0, 0);
    float upl = 0.0f;
    if (is_mesh) {
      int tc_adr = mesh_id >= 0 ? m->mesh_texcoordadr[mesh_id] : -1;
      if (mesh_id >= 0 && face_id >= 0 && tc_adr >= 0) {
        int fadr = m->mesh_faceadr[mesh_id] + face_id;
        const int* tc = m->mesh_facetexcoord + 3 * fadr;
        if (tc[0] >= 0 && tc[1] >= 0 && tc[2] >= 0) {
          const float* T = m->mesh_texcoord;
          v2 u0 = {T[2 * (tc_adr + tc[0])] * repeat.x, T[2 * (tc_adr + tc[0]) + 1] * repeat.y};
          v2 u1 = {T[2 * (tc_adr + tc[1])] * repeat.x, T[2 * (tc_adr + tc[1]) + 1] * repeat.y};
          v2 u2 = {T[2 * (tc_adr + tc[2])] * repeat.x, T[2 * (tc_adr + tc[2]) + 1] * repeat.y};
          float bw = 1.0f - hit->bary_u - hit->bary_v;
          uv.x = u0.x * hit->bary_u + u1.x * hit->bary_v + u2.x * bw;
          uv.y = u0.y * hit->bary_u + u1.y * hit->bary_v + u2.y * bw;
          int vadr = m->mesh_vertadr[mesh_id];
          const int* f = m->mesh_face + 3 * fadr;
          v3 p0 = load3(m->mesh_vert + 3 * (vadr + f[0]));
          v3 e1 = sub3(load3(m->mesh_vert + 3 * (vadr + f[1])), p0);
          v3 e2 = sub3(load3(m->mesh_vert + 3 * (vadr + f[2])), p0);
          v2 d1 = {u1.x - u0.x, u1.y - u0.y}, d2 = {u2.x - u0.x, u2.y - u0.y};
          float det = d1.x * d2.y - d2.x * d1.y;
          float inv = 1.0f / (fabsf(det) > 1.0e-20f ? det : 1.0e-20f);
          tangent = mul3(sub3(mul3(e1, d2.y), mul3(e2, d1.y)), inv);
          bitangent = mul3(sub3(mul3(e2, d1.x), mul3(e1, d2.x)), inv);
          float area = length3(cross3(e1, e2));
          upl = sqrtf(fabsf(det) / fmaxf(area, 1.0e-20f));
          has_uv = 1;
        }
      }
    } else {
      v3 gpos = load3(g->geom_xpos + ((int64_t)w * ng + gid) * 3);
      v3 local = mtmul(R, sub3(point, gpos));
      v3 t1, t2;
      basis_o(lnrm, &t1, &t2);
      float eps = 1.0e-3f * fmaxf(fmaxf(size.x, size.y), fmaxf(size.z, 1.0e-3f));
      v2 uv0, rep, off, uv1, uv2, r1, o1;
      prim_uv_o(gtype, local, lnrm, size, repeat, uniform, &uv0, &rep, &off);
      prim_uv_o(gtype, add3(local, mul3(t1, eps)), lnrm, size, repeat, uniform, &uv1, &r1, &o1);
      prim_uv_o(gtype, add3(local, mul3(t2, eps)), lnrm, size, repeat, uniform, &uv2, &r1, &o1);
      v2 dd1 = {uv1.x - uv0.x, uv1.y - uv0.y}, dd2 = {uv2.x - uv0.x, uv2.y - uv0.y};
      v2 d1 = wrapd_o(dd1), d2 = wrapd_o(dd2);
      d1.x *= rep.x; d1.y *= rep.y; d2.x *= rep.x; d2.y *= rep.y;
      float inv_eps = 1.0f / eps;
      v2 du = {d1.x * inv_eps, d2.x * inv_eps}, dv = {d1.y * inv_eps, d2.y * inv_eps};
      float det = du.x * dv.y - du.y * dv.x;
      float inv = 1.0f / (fabsf(det) > 1.0e-20f ? det : 1.0e-20f);
      tangent = mul3(sub3(mul3(t1, dv.y), mul3(t2, dv.x)), inv);
      bitangent = mul3(sub3(mul3(t2, du.x), mul3(t1, du.y)), inv);
      upl = sqrtf(fabsf(det));
      uv.x = uv0.x * rep.x + off.x;
      uv.y = uv0.y * rep.y + off.y;
      has_uv = 1;
    }
    if (has_uv) {
      const int* tid = m->mat_texid + ((int64_t)mw * g->nmat + mat_id) * MJPBR_NTEXROLE;
      if (s->base_color_map) {
        int t = tid[1] >= 0 ? tid[1] : tid[8];
        if (t >= 0) {
          f4 tx = role_o(m, t, uv, cone, upl);
          sf.base_color = cwm(sf.base_color, mk3(tx.x, tx.y, tx.z));
          sf.alpha = sf.alpha * tx.w;
        }
      }
      if (s->orm_map) {
        int t = tid[9];
        if (t >= 0) {
          f4 tx = role_o(m, t, uv, cone, upl);
          sf.occlusion = tx.x;
          sf.pr = sf.pr * tx.y;
          sf.metallic = sf.metallic * tx.z;
        } else {
          if (tid[2] >= 0) sf.occlusion = role_o(m, tid[2], uv, cone, upl).x;
          if (tid[3] >= 0) sf.pr = sf.pr * role_o(m, tid[3], uv, cone, upl).x;
          if (tid[4] >= 0) sf.metallic = sf.metallic * role_o(m, tid[4], uv, cone, upl).x;
        }
      }
      if (s->opacity_map && tid[6] >= 0) sf.alpha = sf.alpha * role_o(m, tid[6], uv, cone, upl).x;
      if (s->emissive_map && tid[7] >= 0) {
        f4 tx = role_o(m, tid[7], uv, cone, upl);
        sf.emissive = mul3(mk3(tx.x, tx.y, tx.z), emission);
        emission = 0.0f;
      }
      if (s->normal_map && tid[5] >= 0) {
        f4 tx = role_o(m, tid[5], uv, cone, upl);
        v3 nts = sub3(mul3(mk3(tx.x, tx.y, tx.z), 2.0f), mk3(1, 1, 1));
        if (s->specular_aa) {
          float l = clampf_(length3(nts), 1.0e-3f, 1.0f);
          sf.pr = widen_o(sf.pr, (1.0f - l) / l);
        }
        v3 tt = sub3(tangent, mul3(shading, dot3(shading, tangent)));
        if (dot3(tt, tt) > 1.0e-20f) {
          tt = norm3(tt);
          v3 b = cross3(shading, tt);
          if (dot3(b, bitangent) > 0.0f) b = neg3(b);
          v3 mapped = norm3(add3(add3(mul3(tt, nts.x), mul3(b, nts.y)), mul3(shading, nts.z)));
          if (dot3(mapped, lnrm) > 0.0f) shading = mapped;
        }
      }
    }
  }
  if (mat_id >= 0 && emission > 0.0f) sf.emissive = mul3(sf.base_color, emission);
  if (s->specular_aa) {
    float dn = cone * curvature;
    float kern = fminf(2.0f * 0.15f * dn * dn, 0.2f);
    sf.pr = widen_o(sf.pr, kern);
  }
  *shading_world = norm3(mmul(R, shading));
  return sf;
}

static float mirror_o(const Ctx* c, int gid) {
  const MjpbrGeometry* g = &c->sc->geo;
  if (gid < 0) return 0.0f;
  int mid = g->geom_matid[(c->w % g->geom_matid_n) * g->ngeom + gid];
  if (mid < 0) return 0.0f;
  const MjpbrMaterials* m = &c->sc->mats;
  return m->mat_reflectance[(c->w % m->mat_reflectance_n) * g->nmat + mid];
}

/* ------------------------------------------------------------------ RT shading (render.py) */
static float visibility_o(const Ctx* c, const LightO* L, v3 point, v3 normal, v3 l, float dist, int stream, int nlight, int li) {
  const MjpbrSettings* s = c->s;
  v3 origin = add3(point, mul3(normal, O_RAY_OFFSET));
  if (L->type == MJPBR_LIGHT_DIRECTIONAL) return any_o(c, origin, l, O_UNBOUNDED) ? s->shadow_light_fraction : 1.0f;
  if (s->shadow_samples <= 1) return any_o(c, origin, l, dist - O_SHADOW_GAP) ? s->shadow_light_fraction : 1.0f;
  v3 t1, t2;
  basis_o(l, &t1, &t2);
  float visible = 0.0f;
  for (int k = 0; k < s->shadow_samples; ++k) {
    int counter = ((stream * nlight + li) * s->shadow_samples + k) * 2;
    float u0 = rand_o(c->seed, counter), u1 = rand_o(c->seed, counter + 1);
    float r = L->bulbradius * sqrtf(u0);
    float sp, cp;
    sincos_turns_o(u1, &sp, &cp);
    v3 target = sub3(add3(L->pos, mul3(add3(mul3(t1, cp), mul3(t2, sp)), r)), origin);
    float d = length3(target);
    if (!any_o(c, origin, div3(target, d), d - O_SHADOW_GAP)) visible += 1.0f;
  }
  float v = visible / (float)s->shadow_samples;
  return v + (1.0f - v) * s->shadow_light_fraction;
}

static v3 shade_light_o(const Ctx* c, const LightO* L, const LobesO* lb, v3 point, v3 normal, v3 view, float nov, int stream,
                        int nlight, int li) {
  v3 l, ds, ss;
  float dist;
  eval_light_o(L, c->s->light_model, c->s->has_spot, point, &l, &dist, &ds, &ss);
  float nol = dot3(normal, l);
  if (nol <= 0.0f || (dot3(ds, ds) == 0.0f && dot3(ss, ss) == 0.0f)) return mk3(0, 0, 0);
  float vis = 1.0f;
  if (c->s->use_shadows && L->castshadow) vis = visibility_o(c, L, point, normal, l, dist, stream, nlight, li);
  v3 h = norm3(add3(l, view));
  float noh = fmaxf(dot3(normal, h), 0.0f), loh = fmaxf(dot3(l, h), 0.0f);
  v3 fr = cwm(spec_lobe_o(lb->f0, lb->alpha, noh, nov, nol, loh), lb->energy_comp);
  v3 fd = mul3(lb->diffuse_color, O_INV_PI);
  return mul3(add3(cwm(fd, ds), cwm(fr, ss)), nol * vis);
}

/* Lagarde & de Rousiers, "Moving Frostbite to PBR" (2014), specular occlusion from ambient
 * occlusion (reference render.py:217-219). */
static float spec_ao_o(float nov, float ao, float alpha) {
  return clampf_(powf(nov + ao, powf(2.0f, -16.0f * alpha - 1.0f)) - 1.0f + ao, 0.0f, 1.0f);
}

static v3 shade_surface_o(const Ctx* c, const SurfO* sf, v3 point, v3 normal, v3 view, v3 head_dir, int stream) {
  const MjpbrScene* sc = c->sc;
  const MjpbrSettings* s = c->s;
  const MjpbrEnvironment* env = &sc->env;
  float nov = fmaxf(dot3(normal, view), 1.0e-4f);
  LobesO lb = lobes_o(env, sf, nov);
  float alpha = lb.alpha;
  v3 ambient = load3(env->ambient + 3 * (c->w % env->ambient_n));
  v3 irradiance = ambient, prefiltered = ambient;
  if (s->has_env) {
    int e;
    float yaw, scale;
    env_lookup_o(env, c->w, &e, &yaw, &scale);
    irradiance = add3(ambient, mul3(sh9_o(env->env_sh + e * 27, rot_yaw_o(normal, -yaw)), scale));
    v3 r = sub3(view, mul3(normal, 2.0f * dot3(view, normal)));
    v3 dominant = norm3(add3(mul3(neg3(r), 1.0f - alpha * alpha), mul3(normal, alpha * alpha)));
    prefiltered = add3(ambient, mul3(specular_env_o(env, e, dominant, yaw, lb.pr), scale));
  }
  float ao = sf->occlusion;
  float spec_ao = spec_ao_o(nov, ao, alpha);
  v3 radiance = add3(mul3(cwm(irradiance, lb.diffuse_color), ao), mul3(cwm(prefiltered, cwm(lb.spec_albedo, lb.energy_comp)), spec_ao));
  const MjpbrLights* lt = &sc->lit;
  for (int li = 0; li < lt->nlight; ++li) {
    if (!lt->light_active[(c->w % lt->light_active_n) * lt->nlight + li]) continue;
    LightO L = fetch_light_o(lt, c->w, li);
    radiance = add3(radiance, shade_light_o(c, &L, &lb, point, normal, view, nov, stream, lt->nlight, li));
  }
  if (s->headlight_active) {
    LightO H;
    memset(&H, 0, sizeof(H));
    H.type = MJPBR_LIGHT_DIRECTIONAL;
    H.dir = head_dir;
    H.diffuse = load3(s->headlight_diffuse);
    H.specular = load3(s->headlight_specular);
    H.attenuation = mk3(1, 0, 0);
    H.intensity = s->headlight_intensity / c->exposure;
    H.range = 1.0f;
    radiance = add3(radiance, shade_light_o(c, &H, &lb, point, normal, view, nov, stream, 1, 0));
  }
  return add3(radiance, div3(sf->emissive, c->exposure));
}

static v3 background_o(const Ctx* c, v3 dir, float spread) {
  const MjpbrEnvironment* env = &c->sc->env;
  if (c->s->env_background) {
    int e;
    float yaw, scale;
    env_lookup_o(env, c->w, &e, &yaw, &scale);
    float lod = log2f(fmaxf(spread * c->s->inv_texel_angle, 1.0e-8f));
    return mul3(radiance_o(env, e, dir, yaw, lod), scale);
  }
  return div3(load3(env->background), c->exposure);
}

static v3 reflection_o(const Ctx* c, v3 point, v3 normal, v3 dir, float footprint, float spread, v3 head_dir, int stream) {
  v3 r = sub3(dir, mul3(normal, 2.0f * dot3(dir, normal)));
  v3 origin = add3(point, mul3(normal, O_RAY_OFFSET));
  HitO hit = closest_o(c, origin, r, O_MAXVAL);
  if (hit.geom_id == -1) return background_o(c, r, spread);
  v3 p = add3(origin, mul3(r, hit.dist));
  v3 n;
  SurfO sf = eval_surface_o(c, &hit, p, r, footprint + spread * hit.dist, &n);
  return shade_surface_o(c, &sf, p, n, neg3(r), head_dir, stream);
}

static v3 trace_pixel_o(const Ctx* c, v3 origin, v3 dir, float max_dist, float footprint, float spread, v3 head_dir, int stream,
                        HitO* first, v3* falb, v3* fnrm) {
  const MjpbrSettings* s = c->s;
  v3 radiance = mk3(0, 0, 0);
  float throughput = 1.0f, travelled = 0.0f;
  first->geom_id = -1;
  *falb = mk3(0, 0, 0);
  *fnrm = mk3(0, 0, 0);
  int nlayer = s->has_transparency ? s->max_layers : 1;
  for (int layer = 0; layer < nlayer; ++layer) {
    HitO hit = closest_o(c, origin, dir, max_dist - travelled);
    if (layer == 0) *first = hit;
    if (hit.geom_id == -1) {
      radiance = add3(radiance, mul3(background_o(c, dir, spread), throughput));
      break;
    }
    v3 point = add3(origin, mul3(dir, hit.dist));
    float cone = footprint + spread * (travelled + hit.dist);
    v3 normal;
    SurfO sf = eval_surface_o(c, &hit, point, dir, cone, &normal);
    if (layer == 0 && s->render_guides) { *falb = sf.base_color; *fnrm = normal; }
    v3 shaded = shade_surface_o(c, &sf, point, normal, neg3(dir), head_dir, stream);
    if (s->has_reflection) {
      float mirror = mirror_o(c, hit.geom_id);
      if (mirror > 0.0f) {
        v3 refl = reflection_o(c, point, normal, dir, cone, spread, head_dir, stream + 1);
        shaded = add3(mul3(shaded, 1.0f - mirror), mul3(refl, mirror));
      }
    }
    if (s->has_transparency) {
      float opacity = clampf_(sf.alpha, 0.0f, 1.0f);
      radiance = add3(radiance, mul3(shaded, throughput * opacity));
      throughput *= 1.0f - opacity;
      if (throughput < 1.0f / 255.0f) break;
      travelled += hit.dist;
      origin = add3(point, mul3(dir, O_RAY_OFFSET));
    } else {
      radiance = add3(radiance, shaded);
      break;
    }
  }
  return radiance;
}

/* ------------------------------------------------------------------ cameras (camera.py) */
typedef struct {
  int camid, cam, pixelid, px, py, width, height, rayid;
  int want_rgb, want_depth, want_seg, projection;
  float fovy, sensor[2], intr[4], xmat[9];
  v3 xpos;
  const float* rorg; /* ray bundle from the camera's first pixel, NULL: pinhole */
  const float* rdir;
} PixO;

/* Ray bundles (camera.py:194-347). */
static void bundle_axis_o(int i, int n, int nb, int* x0, int* x1, float* t) {
  int64_t num = (int64_t)(2 * i + 1) * nb - n, den = 2 * (int64_t)n;
  int a = (int)((num + den) / den) - 1;
  int hi = nb - 2 > 0 ? nb - 2 : 0;
  if (a < 0) a = 0;
  if (a > hi) a = hi;
  *x0 = a;
  *x1 = a + 1 < nb - 1 ? a + 1 : nb - 1;
  *t = (float)(num - (int64_t)a * den) / (float)den;
}

static v3 lerp_o(v3 a, v3 b, float t) { return add3(a, mul3(sub3(b, a), t)); }

static void bundle_ray_o(const PixO* p, int subdiv, int sx, int sy, v3* dir, v3* org) {
  int x0, x1, y0, y1;
  float tx, ty;
  bundle_axis_o(p->px * subdiv + sx, p->width * subdiv, p->width, &x0, &x1, &tx);
  bundle_axis_o(p->py * subdiv + sy, p->height * subdiv, p->height, &y0, &y1, &ty);
  int r0 = y0 * p->width, r1 = y1 * p->width;
  v3 d00 = load3(p->rdir + 3 * (r0 + x0)), d10 = load3(p->rdir + 3 * (r0 + x1));
  v3 d01 = load3(p->rdir + 3 * (r1 + x0)), d11 = load3(p->rdir + 3 * (r1 + x1));
  if (dot3(d00, d00) == 0.0f || dot3(d10, d10) == 0.0f || dot3(d01, d01) == 0.0f || dot3(d11, d11) == 0.0f) {
    int near = (ty < 0.5f ? r0 : r1) + (tx < 0.5f ? x0 : x1);
    v3 d = load3(p->rdir + 3 * near);
    *dir = dot3(d, d) > 0.0f ? norm3(d) : d;
    *org = load3(p->rorg + 3 * near);
    return;
  }
  v3 o00 = load3(p->rorg + 3 * (r0 + x0)), o10 = load3(p->rorg + 3 * (r0 + x1));
  v3 o01 = load3(p->rorg + 3 * (r1 + x0)), o11 = load3(p->rorg + 3 * (r1 + x1));
  if (d00.z < 0.0f && d10.z < 0.0f && d01.z < 0.0f && d11.z < 0.0f) {
    d00 = mul3(d00, -1.0f / d00.z);
    d10 = mul3(d10, -1.0f / d10.z);
    d01 = mul3(d01, -1.0f / d01.z);
    d11 = mul3(d11, -1.0f / d11.z);
  }
  v3 d = lerp_o(lerp_o(d00, d10, tx), lerp_o(d01, d11, tx), ty);
  *org = lerp_o(lerp_o(o00, o10, tx), lerp_o(o01, o11, tx), ty);
  *dir = dot3(d, d) > 0.0f ? norm3(d) : d;
}

static void bundle_step_o(const PixO* p, int a, int b, int extrapolate, float* spread, float* width) {
  v3 da = load3(p->rdir + 3 * a), db = load3(p->rdir + 3 * b);
  *spread = 0.0f;
  *width = 0.0f;
  if (dot3(da, da) == 0.0f || dot3(db, db) == 0.0f) return;
  da = norm3(da);
  db = norm3(db);
  *width = length3(sub3(load3(p->rorg + 3 * b), load3(p->rorg + 3 * a)));
  if (extrapolate) {
    if (da.z < 0.0f && db.z < 0.0f)
      *spread = length3(sub3(norm3(add3(mul3(db, -2.0f / db.z), mul3(da, 1.0f / da.z))), db));
    else
      *spread = length3(sub3(norm3(sub3(mul3(db, 2.0f), da)), db));
    return;
  }
  *spread = length3(sub3(db, da));
}

static void pixel_cone_o(const PixO* p, float* spread, float* width);

static void cam_ray_o(const PixO* p, int subdiv, int sx, int sy, v3* dir, v3* off) {
  int W = p->width * subdiv, H = p->height * subdiv;
  float x = (float)(p->px * subdiv + sx), y = (float)(p->py * subdiv + sy);
  float u = (x + 0.5f) / (float)W - 0.5f, v = (y + 0.5f) / (float)H - 0.5f;
  if (p->projection == 1) {
    float half_h = p->fovy * 0.5f, half_w = half_h * (float)W / (float)H;
    *dir = mk3(0, 0, -1);
    *off = mk3(2.0f * u * half_w, -2.0f * v * half_h, 0.0f);
    return;
  }
  v3 d;
  if (p->sensor[0] != 0.0f && p->sensor[1] != 0.0f) {
    float fx = p->intr[0] / p->sensor[0] * (float)W, fy = p->intr[1] / p->sensor[1] * (float)H;
    float cx = p->intr[2] / p->sensor[0] * (float)W, cy = p->intr[3] / p->sensor[1] * (float)H;
    d = mk3((x + 0.5f - 0.5f * (float)W - cx) / fx, -(y + 0.5f - 0.5f * (float)H - cy) / fy, -1.0f);
  } else {
    double h = tan((double)p->fovy * (3.14159265358979323846 / 360.0));
    float ax = (float)(2.0 * h * (double)W / (double)H), ay = (float)(2.0 * h);
    d = mk3(u * ax, -v * ay, -1.0f);
  }
  *dir = norm3(d);
  *off = mk3(0, 0, 0);
}

static void pixel_cone_o(const PixO* p, float* spread, float* width) {
  if (p->rdir) {
    *spread = 0.0f;
    *width = 0.0f;
    if (p->width < 2) return;
    int b = p->py * p->width + p->px;
    if (p->px + 1 < p->width)
      bundle_step_o(p, b, b + 1, 0, spread, width);
    else
      bundle_step_o(p, b - 1, b, 1, spread, width);
    return;
  }
  v3 d0, o0, d1, o1;
  cam_ray_o(p, 1, 0, 0, &d0, &o0);
  cam_ray_o(p, 1, 1, 0, &d1, &o1);
  *spread = length3(sub3(d1, d0));
  *width = length3(sub3(o1, o0));
}

static void primary_o(const PixO* p, const MjpbrSettings* s, int subdiv, int sx, int sy, v3* o, v3* d, float* md, v3* dc) {
  v3 dcam, ocam;
  if (p->rdir) {
    bundle_ray_o(p, subdiv, sx, sy, &dcam, &ocam);
    *md = s->zfar > 0.0f ? s->zfar : O_MAXVAL;
    if (dot3(dcam, dcam) == 0.0f) {
      *md = 0.0f;
      dcam = mk3(0, 0, -1);
    }
    *o = add3(p->xpos, mmul(p->xmat, ocam));
    *d = mmul(p->xmat, dcam);
    *dc = mk3(0, 0, -1); /* depth along the ray */
    return;
  }
  cam_ray_o(p, subdiv, sx, sy, &dcam, &ocam);
  *o = add3(p->xpos, mmul(p->xmat, ocam));
  *d = mmul(p->xmat, dcam);
  *dc = dcam;
  *md = s->zfar > 0.0f ? s->zfar / fmaxf(-dcam.z, 1.0e-6f) : O_MAXVAL;
}

/* Reference ray id of pixel (px, py) under the render order (camera.py:62-70, inverted). */
static int ray_index_o(int order, int px, int py, int w, int h) {
  if (order == 1) {
    int band = py / 8;
    int rows = h - band * 8 < 8 ? h - band * 8 : 8;
    return band * w * 8 + px * rows + (py - band * 8);
  }
  return py * w + px;
}

static void write_geometry_o(MjpbrOutputs* out, int w, const PixO* p, const HitO* hit, v3 dc) {
  if (p->want_seg) {
    int32_t* seg = out->seg + 2 * ((int64_t)w * out->seg_stride + out->seg_adr[p->camid] + p->pixelid);
    seg[0] = hit->geom_id != -1 ? hit->geom_id : -1;
    seg[1] = hit->geom_id != -1 ? 5 : -1;
  }
  if (p->want_depth)
    out->depth[(int64_t)w * out->depth_stride + out->depth_adr[p->camid] + p->pixelid] = hit->geom_id != -1 ? hit->dist * -dc.z : 0.0f;
}

static int setup_pixel_o(const MjpbrScene* sc, const MjpbrOutputs* out, const MjpbrSettings* s, int w, int c, int px, int py,
                         int ray_offset, PixO* p) {
  const MjpbrCameras* cams = &sc->cams;
  p->camid = c;
  p->want_rgb = s->any_rgb && out->render_rgb[c];
  p->want_depth = out->render_depth[c] != 0;
  p->want_seg = out->render_seg[c] != 0;
  if (!p->want_rgb && !p->want_depth && !p->want_seg) return 0;
  int cam = cams->cam_id_map[c];
  p->cam = cam;
  p->width = cams->cam_res[c][0];
  p->height = cams->cam_res[c][1];
  p->px = px;
  p->py = py;
  p->pixelid = py * p->width + px;
  p->rayid = ray_offset + ray_index_o(s->render_order, px, py, p->width, p->height);
  p->projection = cams->cam_projection[cam];
  p->fovy = cams->cam_fovy[(w % cams->cam_fovy_n) * cams->ncam + cam];
  p->sensor[0] = cams->cam_sensorsize[2 * cam];
  p->sensor[1] = cams->cam_sensorsize[2 * cam + 1];
  for (int k = 0; k < 4; ++k) p->intr[k] = cams->cam_intrinsic[((int64_t)(w % cams->cam_intrinsic_n) * cams->ncam + cam) * 4 + k];
  for (int k = 0; k < 9; ++k) p->xmat[k] = cams->cam_xmat[((int64_t)w * cams->ncam + cam) * 9 + k];
  p->xpos = load3(cams->cam_xpos + ((int64_t)w * cams->ncam + cam) * 3);
  p->rorg = cams->ray_direction ? cams->ray_origin + 3 * (int64_t)ray_offset : NULL;
  p->rdir = cams->ray_direction ? cams->ray_direction + 3 * (int64_t)ray_offset : NULL;
  return 1;
}

/* render.py:343-395 for one pixel */
static void render_pixel_o(const MjpbrScene* sc, const MjpbrSettings* s, MjpbrOutputs* out, int seed, int w, const PixO* pix) {
  Ctx c = {sc, s, w, 0, sc->env.exposure[w % sc->env.exposure_n]};
  if (!pix->want_rgb) {
    v3 o, d, dc;
    float md;
    primary_o(pix, s, 1, 0, 0, &o, &d, &md, &dc);
    HitO hit = closest_o(&c, o, d, md);
    write_geometry_o(out, w, pix, &hit, dc);
    return;
  }
  int n_aa = s->samples_per_axis;
  float inv_aa = 1.0f / (float)n_aa;
  float spread, width;
  pixel_cone_o(pix, &spread, &width);
  spread *= inv_aa;
  width *= inv_aa;
  v3 head_dir = neg3(mk3(pix->xmat[2], pix->xmat[5], pix->xmat[8]));
  c.seed = seed + w * sc->cams.total_rays + pix->rayid;
  v3 display = mk3(0, 0, 0), hdr = mk3(0, 0, 0);
  for (int smp = 0; smp < n_aa * n_aa; ++smp) {
    v3 o, d, dc;
    float md;
    primary_o(pix, s, n_aa, smp % n_aa, smp / n_aa, &o, &d, &md, &dc);
    HitO first;
    v3 alb, nrm;
    v3 radiance = trace_pixel_o(&c, o, d, md, width, spread, head_dir, 2 * smp, &first, &alb, &nrm);
    if (smp == 0) {
      write_geometry_o(out, w, pix, &first, dc);
      if (s->render_guides) {
        int64_t a = (int64_t)w * out->guide_stride + out->rgb_adr[pix->camid] + pix->pixelid;
        float* al = out->albedo + 3 * a;
        float* nr = out->normal + 3 * a;
        al[0] = alb.x; al[1] = alb.y; al[2] = alb.z;
        nr[0] = nrm.x; nr[1] = nrm.y; nr[2] = nrm.z;
      }
    }
    display = add3(display, display_o(mul3(radiance, c.exposure), s->tone_mapper));
    hdr = add3(hdr, radiance);
  }
  float inv2 = inv_aa * inv_aa;
  int64_t adr = out->rgb_adr[pix->camid] + pix->pixelid;
  if (s->render_hdr) {
    float* h = out->hdr + 3 * ((int64_t)w * out->hdr_stride + adr);
    h[0] = hdr.x * inv2; h[1] = hdr.y * inv2; h[2] = hdr.z * inv2;
  }
  out->rgb[(int64_t)w * out->rgb_stride + adr] = encode_o(mul3(display, inv2));
}

/* Renders worlds [w0, w1) of the real-time tier. Returns 0. */
int oracle_rt_render(const MjpbrScene* sc, const MjpbrSettings* s, MjpbrOutputs* out, int32_t seed, int w0, int w1) {
  for (int w = w0; w < w1; ++w) {
    int ray_offset = sc->cams.ray_base;  /* a launch of the context's later cameras */
    for (int c = 0; c < sc->cams.nrender; ++c) {
      int W = sc->cams.cam_res[c][0], H = sc->cams.cam_res[c][1];
      for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
          PixO pix;
          if (setup_pixel_o(sc, out, s, w, c, px, py, ray_offset, &pix)) render_pixel_o(sc, s, out, seed, w, &pix);
        }
      ray_offset += W * H;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ precompute (ibl.py, envmap.py) */
static v2 hammersley_o(int i, float inv_n) {
  uint32_t b = (uint32_t)i;
  b = (b << 16) | (b >> 16);
  b = ((b & 0x55555555u) << 1) | ((b & 0xAAAAAAAAu) >> 1);
  b = ((b & 0x33333333u) << 2) | ((b & 0xCCCCCCCCu) >> 2);
  b = ((b & 0x0F0F0F0Fu) << 4) | ((b & 0xF0F0F0F0u) >> 4);
  b = ((b & 0x00FF00FFu) << 8) | ((b & 0xFF00FF00u) >> 8);
  v2 r = {(float)i * inv_n, (float)b * (1.0f / 4294967296.0f)};
  return r;
}

static v3 ggx_sample_o(v2 u, float alpha) {
  float a2 = alpha * alpha, sp, cp;
  sincos_turns_o(u.x, &sp, &cp);
  float cos2 = (1.0f - u.y) / (1.0f + (a2 - 1.0f) * u.y);
  float ct = sqrtf(cos2), st = sqrtf(fmaxf(1.0f - cos2, 0.0f));
  return mk3(st * cp, st * sp, ct);
}

/* ibl.py:62-93: lut[j (roughness), i (NoV)] = (int V Fc, int V). */
int oracle_dfg_lut(float* lut, int n, int nsample) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      float nov = ((float)i + 0.5f) / (float)n, pr = ((float)j + 0.5f) / (float)n, alpha = pr * pr;
      v3 v = mk3(sqrtf(1.0f - nov * nov), 0.0f, nov);
      float inv_n = 1.0f / (float)nsample, x = 0.0f, y = 0.0f;
      for (int s = 0; s < nsample; ++s) {
        v3 h = ggx_sample_o(hammersley_o(s, inv_n), alpha);
        float voh = dot3(v, h);
        v3 l = sub3(mul3(h, 2.0f * voh), v);
        if (l.z > 0.0f && voh > 0.0f) {
          float gv = v_smith_o(nov, l.z, alpha) * 4.0f * voh * l.z / h.z;
          float t = 1.0f - voh;
          x += gv * (t * t * t * t * t);
          y += gv;
        }
      }
      lut[2 * (j * n + i)] = x * inv_n;
      lut[2 * (j * n + i) + 1] = y * inv_n;
    }
  return 0;
}

/* envmap.py:171-221: prefiltered stack texel (level, y, x) as float RGB. */
int oracle_prefilter_texel(const uint16_t* radiance, const int64_t* level_adr, const int32_t* desc, int L, int H, int W, int nsample,
                           int level, int y, int x, float* rgb) {
  v2 uv = {((float)x + 0.5f) / (float)W, ((float)y + 0.5f) / (float)H};
  v3 n = uv_to_dir_o(uv);
  float pr = (float)level / (float)(L - 1), alpha = fmaxf(pr * pr, 1.0e-4f);
  v3 up = fabsf(n.z) < 0.999f ? mk3(0, 0, 1) : mk3(1, 0, 0);
  v3 t = norm3(cross3(up, n)), b = cross3(n, t);
  float texel_sa = 4.0f * O_PI / (float)(desc[0] * desc[1]);
  v3 color = mk3(0, 0, 0);
  float weight = 0.0f, inv_n = 1.0f / (float)nsample;
  for (int s = 0; s < nsample; ++s) {
    v3 hl = ggx_sample_o(hammersley_o(s, inv_n), alpha);
    v3 h = add3(add3(mul3(t, hl.x), mul3(b, hl.y)), mul3(n, hl.z));
    float noh = hl.z;
    v3 l = sub3(mul3(h, 2.0f * noh), n);
    float nol = dot3(n, l);
    if (nol > 0.0f) {
      float pdf = d_ggx_o(noh, alpha) * 0.25f;
      float ssa = inv_n / fmaxf(pdf, 1.0e-8f);
      float lod = level == 0 ? 0.0f : 0.5f * log2f(ssa / texel_sa) + 1.0f;
      v2 luv = dir_to_uv_o(l, 0.0f);
      f4 c = mip_o((const uint8_t*)radiance, level_adr, desc, 2, luv.x, luv.y, fmaxf(lod, 0.0f), 0);
      color = add3(color, mul3(mk3(c.x, c.y, c.z), nol));
      weight += nol;
    }
  }
  color = div3(color, fmaxf(weight, 1.0e-8f));
  rgb[0] = color.x; rgb[1] = color.y; rgb[2] = color.z;
  return 0;
}

/* Direct access to the sampling primitives, for their tests. */
float oracle_rand(int32_t seed, int32_t counter) { return rand_o(seed, counter); }
void oracle_sincos_turns(float t, float* s, float* c) { sincos_turns_o(t, s, c); }

/* Brute-force closest hit of one ray in world w, for tests of the traversal. */
int oracle_closest(const MjpbrScene* sc, const MjpbrSettings* s, int w, const float* o, const float* d, float max_dist, int* geom,
                   float* dist) {
  Ctx c = {sc, s, w, 0, 1.0f};
  HitO h = closest_o(&c, load3(o), load3(d), max_dist);
  *geom = h.geom_id;
  *dist = h.dist;
  return 0;
}

/* ------------------------------------------------------------------ test entry points
 * The oracle's scalar building blocks, exported so CPU tests can pin each one against closed
 * forms and published values (the renderer paths above call the same static functions). */
float oracle_ray_primitive(int type, const float* pos, const float* mat, const float* size, const float* o, const float* d,
                           float* normal) {
  v3 n = mk3(0, 0, 0), p = load3(pos), sz = load3(size), po = load3(o), pd = load3(d);
  float x = -1.0f;
  switch (type) {
    case MJPBR_GEOM_PLANE: x = plane_o(p, mat, sz, po, pd, &n); break;
    case MJPBR_GEOM_SPHERE: x = sphere_o(p, sz.x * sz.x, po, pd, &n); break;
    case MJPBR_GEOM_ELLIPSOID: x = ellipsoid_o(p, mat, sz, po, pd, &n); break;
    case MJPBR_GEOM_CAPSULE: x = capsule_o(p, mat, sz, po, pd, &n); break;
    case MJPBR_GEOM_CYLINDER: x = cylinder_o(p, mat, sz, po, pd, &n); break;
    case MJPBR_GEOM_BOX: x = box_o(p, mat, sz, po, pd, &n); break;
    default: break;
  }
  normal[0] = n.x; normal[1] = n.y; normal[2] = n.z;
  return x;
}

/* out = {t, b1, b2}; returns 1 on a hit. */
int oracle_ray_triangle(const float* o, const float* d, const float* v0, const float* v1, const float* v2, float tmax, float* out) {
  v3 a = load3(v0);
  float t = 0, b1 = 0, b2 = 0;
  int hit = tri_o(load3(o), load3(d), a, sub3(load3(v1), a), sub3(load3(v2), a), tmax, &t, &b1, &b2);
  out[0] = t; out[1] = b1; out[2] = b2;
  return hit;
}

void oracle_display(const float* c, int tone_mapper, float* out) {
  v3 r = display_o(load3(c), tone_mapper);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

uint32_t oracle_encode(const float* c) { return encode_o(load3(c)); }

float oracle_spec_ao(float nov, float ao, float alpha) { return spec_ao_o(nov, ao, alpha); }

/* One light at `point`: out = {l.xyz, dist, diffuse scale.xyz, specular scale.xyz}. */
void oracle_eval_light(int type, int model, int has_spot, const float* pos, const float* dir, float cutoff, float exponent,
                       const float* attenuation, float intensity, float range, const float* point, float* out) {
  LightO L;
  memset(&L, 0, sizeof(L));
  L.type = type;
  L.pos = load3(pos);
  L.dir = load3(dir);
  L.diffuse = mk3(1, 1, 1);
  L.specular = mk3(1, 1, 1);
  L.attenuation = load3(attenuation);
  L.cutoff = cutoff;
  L.exponent = exponent;
  L.intensity = intensity;
  L.range = range;
  v3 l, ds, ss;
  float dist;
  eval_light_o(&L, model, has_spot, load3(point), &l, &dist, &ds, &ss);
  out[0] = l.x; out[1] = l.y; out[2] = l.z; out[3] = dist;
  out[4] = ds.x; out[5] = ds.y; out[6] = ds.z;
  out[7] = ss.x; out[8] = ss.y; out[9] = ss.z;
}

/* out = F D V of the specular lobe (brdf.py:32-72), no NoL cosine. */
void oracle_spec_lobe(const float* f0, float alpha, float noh, float nov, float nol, float loh, float* out) {
  v3 r = spec_lobe_o(load3(f0), alpha, noh, nov, nol, loh);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* Bilinear DFG lookup (ibl.py:117-119) in an n x n x 2 table. */
void oracle_dfg_lookup(const float* lut, int n, float nov, float pr, float* out) {
  MjpbrEnvironment env;
  memset(&env, 0, sizeof(env));
  env.dfg_size = n;
  env.dfg_lut = lut;
  v2 r = dfg_o(&env, nov, pr);
  out[0] = r.x; out[1] = r.y;
}

/* Shading lobes (material.py:506-539) of a surface against an n x n x 2 DFG table:
 * out = {pr, alpha, f0.xyz, energy_comp.xyz, spec_albedo.xyz, diffuse_color.xyz}. */
void oracle_lobes(const float* base_color, float metallic, float reflectance, float pr, float nov, const float* lut, int n, float* out) {
  MjpbrEnvironment env;
  memset(&env, 0, sizeof(env));
  env.dfg_size = n;
  env.dfg_lut = lut;
  SurfO sf;
  memset(&sf, 0, sizeof(sf));
  sf.base_color = load3(base_color);
  sf.metallic = metallic;
  sf.reflectance = reflectance;
  sf.pr = pr;
  LobesO L = lobes_o(&env, &sf, nov);
  float v[14] = {L.pr, L.alpha, L.f0.x, L.f0.y, L.f0.z, L.energy_comp.x, L.energy_comp.y, L.energy_comp.z,
                 L.spec_albedo.x, L.spec_albedo.y, L.spec_albedo.z, L.diffuse_color.x, L.diffuse_color.y, L.diffuse_color.z};
  memcpy(out, v, sizeof(v));
}
