// camera.hpp — pixels, primary rays, ray cones and the LDS staging of a world's scene
// (reference camera.py:54-191, :316-326; kernel_inputs.py:207-238).
#pragma once
#include "shade.hpp"

namespace mjpbr {

constexpr int kTileRows = 8;  // camera.py:47
constexpr int kTile = 16;     // a workgroup renders a 16 x 16 tile of one world, a wave 8 x 8

// Workgroup -> (world, camera, tile) map, built by the host.
struct TileMap {
  int nrender;
  int tiles_per_world;
  int nblocks;
  int xcd_chunk;  // > 0: blocks are remapped so each XCD gets contiguous chunks of this size
  int tile_base[MJPBR_MAX_CAMS + 1];
  int tiles_x[MJPBR_MAX_CAMS];
  int ray_offset[MJPBR_MAX_CAMS];
  int mesh_cache_nodes, mesh_cache_n;  // mesh-BVH top levels staged in LDS (render kernels)
};

struct Pixel {
  int camid, cam, pixelid, px, py, width, height, rayid;
  bool want_rgb, want_depth, want_seg;
  int projection;
  float fovy;
  double tan_half_fovy;  // tan(fovy / 2) in double (pinhole cameras without intrinsics)
  float sensor[2];
  float intr[4];
  float xmat[9];
  V3 xpos;
  const float* rorg;  // ray bundle of the camera (NULL: pinhole), from its first pixel
  const float* rdir;
};

// Reference ray id of pixel (px, py): the thread index of `camera.thread_pixel` under the render
// order (it seeds the pixel's random numbers).
MJ_DEV int ray_index(int order, int px, int py, int w, int h) {
  if (order == 1) {
    int band = py / kTileRows;
    int rows = min(kTileRows, h - band * kTileRows);
    return band * w * kTileRows + px * rows + (py - band * kTileRows);
  }
  return py * w + px;
}

// Block index after the XCD-aware remap: hardware dispatches workgroups round-robin over the 8
// XCDs, so consecutive logical tiles (same world, same L2) are given to one XCD.
MJ_DEV int logical_block(const TileMap& tm) {
  int b = blockIdx.x;
  if (tm.xcd_chunk <= 0) return b;
  int xcd = b & 7;
  int k = b >> 3;
  return xcd * tm.xcd_chunk + k;
}

MJ_DEV void camera_frame_ray(const Pixel& p, int subdiv, int sx, int sy, V3* dir, V3* off) {
  int W = p.width * subdiv, H = p.height * subdiv;
  float x = (float)(p.px * subdiv + sx), y = (float)(p.py * subdiv + sy);
  float u = (x + 0.5f) / (float)W - 0.5f;
  float v = (y + 0.5f) / (float)H - 0.5f;
  if (p.projection == 1) {
    float half_h = p.fovy * 0.5f;
    float half_w = half_h * (float)W / (float)H;
    *dir = V3{0.0f, 0.0f, -1.0f};
    *off = V3{2.0f * u * half_w, -2.0f * v * half_h, 0.0f};
    return;
  }
  V3 d;
  if (p.sensor[0] != 0.0f && p.sensor[1] != 0.0f) {
    float fx = p.intr[0] / p.sensor[0] * (float)W;
    float fy = p.intr[1] / p.sensor[1] * (float)H;
    float cx = p.intr[2] / p.sensor[0] * (float)W;
    float cy = p.intr[3] / p.sensor[1] * (float)H;
    d = V3{(x + 0.5f - 0.5f * (float)W - cx) / fx, -(y + 0.5f - 0.5f * (float)H - cy) / fy, -1.0f};
  } else {
    double h = p.tan_half_fovy;
    float ax = (float)(2.0 * h * (double)W / (double)H);
    float ay = (float)(2.0 * h);
    d = V3{u * ax, -v * ay, -1.0f};
  }
  *dir = normalize(d);
  *off = v3s(0.0f);
}

// ---------------------------------------------------------------- ray bundles (camera.py:194-347)
// Bundle rays (x0, x1) and weight t of sample i of n along an image axis of nb bundle rays: sample
// i sits at bundle coordinate (i + 0.5) nb / n - 0.5 = num / den, in integers so that t is exactly
// 0 on a bundle pixel centre.
MJ_DEV void bundle_axis(int i, int n, int nb, int* x0, int* x1, float* t) {
  const int64_t num = (int64_t)(2 * i + 1) * nb - n;
  const int64_t den = 2 * (int64_t)n;
  int a = (int)((num + den) / den) - 1;
  a = min(max(a, 0), max(nb - 2, 0));
  *x0 = a;
  *x1 = min(a + 1, nb - 1);
  *t = (float)(num - (int64_t)a * den) / (float)den;
}

MJ_DEV V3 lerp3(V3 a, V3 b, float t) { return a + (b - a) * t; }

// Camera-frame (direction, origin) through sub-pixel cell (sx, sy): bilinear between the four
// nearest bundle rays (on the image plane z = -1 when all point forward), the nearest ray next to
// a pixel without one. The direction is normalised, or zero.
MJ_DEV void bundle_ray(const Pixel& p, int subdiv, int sx, int sy, V3* dir, V3* org) {
  int x0, x1, y0, y1;
  float tx, ty;
  bundle_axis(p.px * subdiv + sx, p.width * subdiv, p.width, &x0, &x1, &tx);
  bundle_axis(p.py * subdiv + sy, p.height * subdiv, p.height, &y0, &y1, &ty);
  const int r0 = y0 * p.width, r1 = y1 * p.width;
  V3 d00 = ld3(p.rdir + 3 * (r0 + x0)), d10 = ld3(p.rdir + 3 * (r0 + x1));
  V3 d01 = ld3(p.rdir + 3 * (r1 + x0)), d11 = ld3(p.rdir + 3 * (r1 + x1));
  if (len_sq(d00) == 0.0f || len_sq(d10) == 0.0f || len_sq(d01) == 0.0f || len_sq(d11) == 0.0f) {
    const int near = (ty < 0.5f ? r0 : r1) + (tx < 0.5f ? x0 : x1);
    V3 d = ld3(p.rdir + 3 * near);
    *dir = len_sq(d) > 0.0f ? normalize(d) : d;
    *org = ld3(p.rorg + 3 * near);
    return;
  }
  V3 o00 = ld3(p.rorg + 3 * (r0 + x0)), o10 = ld3(p.rorg + 3 * (r0 + x1));
  V3 o01 = ld3(p.rorg + 3 * (r1 + x0)), o11 = ld3(p.rorg + 3 * (r1 + x1));
  if (d00.z < 0.0f && d10.z < 0.0f && d01.z < 0.0f && d11.z < 0.0f) {
    d00 = d00 * (-1.0f / d00.z);
    d10 = d10 * (-1.0f / d10.z);
    d01 = d01 * (-1.0f / d01.z);
    d11 = d11 * (-1.0f / d11.z);
  }
  V3 d = lerp3(lerp3(d00, d10, tx), lerp3(d01, d11, tx), ty);
  *org = lerp3(lerp3(o00, o10, tx), lerp3(o01, o11, tx), ty);
  *dir = len_sq(d) > 0.0f ? normalize(d) : d;
}

// (angle, distance) between bundle rays a and b, or between b and its linear extrapolation.
MJ_DEV void bundle_step(const Pixel& p, int a, int b, bool extrapolate, float* spread, float* width) {
  V3 da = ld3(p.rdir + 3 * a), db = ld3(p.rdir + 3 * b);
  *spread = 0.0f;
  *width = 0.0f;
  if (len_sq(da) == 0.0f || len_sq(db) == 0.0f) return;
  da = normalize(da);
  db = normalize(db);
  *width = len(ld3(p.rorg + 3 * b) - ld3(p.rorg + 3 * a));
  if (extrapolate) {
    if (da.z < 0.0f && db.z < 0.0f)
      *spread = len(normalize(db * (-2.0f / db.z) + da * (1.0f / da.z)) - db);
    else
      *spread = len(normalize(db * 2.0f - da) - db);
    return;
  }
  *spread = len(db - da);
}

MJ_DEV void bundle_cone(const Pixel& p, float* spread, float* width) {
  *spread = 0.0f;
  *width = 0.0f;
  if (p.width < 2) return;
  const int b = p.py * p.width + p.px;
  if (p.px + 1 < p.width)
    bundle_step(p, b, b + 1, false, spread, width);
  else
    bundle_step(p, b - 1, b, true, spread, width);
}

// Depth axis of bundle rays: depth is the distance along the ray.
constexpr V3 kAlongRay = V3{0.0f, 0.0f, -1.0f};

MJ_DEV void primary_ray(const Pixel& p, const MjpbrSettings& s, int subdiv, int sx, int sy, V3* o, V3* d,
                        float* max_dist, V3* dir_cam) {
  V3 dc, oc;
  if (p.rdir) {
    bundle_ray(p, subdiv, sx, sy, &dc, &oc);
    *max_dist = s.zfar > 0.0f ? s.zfar : kMaxVal;
    if (len_sq(dc) == 0.0f) {
      *max_dist = 0.0f;
      dc = kAlongRay;
    }
    *o = p.xpos + mat_mul(p.xmat, oc);
    *d = mat_mul(p.xmat, dc);
    *dir_cam = kAlongRay;
    return;
  }
  camera_frame_ray(p, subdiv, sx, sy, &dc, &oc);
  *o = p.xpos + mat_mul(p.xmat, oc);
  *d = mat_mul(p.xmat, dc);
  *dir_cam = dc;
  *max_dist = s.zfar > 0.0f ? s.zfar / fmaxf(-dc.z, 1.0e-6f) : kMaxVal;
}

MJ_DEV void pixel_cone(const Pixel& p, float* spread, float* width) {
  if (p.rdir) {
    bundle_cone(p, spread, width);
    return;
  }
  V3 d0, o0, d1, o1;
  camera_frame_ray(p, 1, 0, 0, &d0, &o0);
  camera_frame_ray(p, 1, 1, 0, &d1, &o1);
  *spread = len(d1 - d0);
  *width = len(o1 - o0);
}

MJ_DEV bool setup_pixel(const MjpbrCameras& cams, const MjpbrOutputs& out, const MjpbrSettings& s, int w, int c,
                        int px, int py, int ray_offset, Pixel* p) {
  int W = cams.cam_res[c][0], H = cams.cam_res[c][1];
  if (px >= W || py >= H) return false;
  p->camid = c;
  p->want_rgb = s.any_rgb && out.render_rgb[c];
  p->want_depth = out.render_depth[c] != 0;
  p->want_seg = out.render_seg[c] != 0;
  if (!p->want_rgb && !p->want_depth && !p->want_seg) return false;
  int cam = cams.cam_id_map[c];
  p->cam = cam;
  p->width = W;
  p->height = H;
  p->px = px;
  p->py = py;
  p->pixelid = py * W + px;
  p->rayid = ray_offset + ray_index(s.render_order, px, py, W, H);
  p->projection = cams.cam_projection[cam];
  p->fovy = cams.cam_fovy[(int64_t)(w % cams.cam_fovy_n) * cams.ncam + cam];
  p->sensor[0] = cams.cam_sensorsize[2 * cam];
  p->sensor[1] = cams.cam_sensorsize[2 * cam + 1];
  p->tan_half_fovy = p->projection != 1 && !(p->sensor[0] != 0.0f && p->sensor[1] != 0.0f)
                         ? tan((double)p->fovy * (3.14159265358979323846 / 360.0))
                         : 0.0;
  const float* in = cams.cam_intrinsic + ((int64_t)(w % cams.cam_intrinsic_n) * cams.ncam + cam) * 4;
#pragma unroll
  for (int k = 0; k < 4; ++k) p->intr[k] = in[k];
  const float* xm = cams.cam_xmat + ((int64_t)w * cams.ncam + cam) * 9;
#pragma unroll
  for (int k = 0; k < 9; ++k) p->xmat[k] = xm[k];
  p->xpos = ld3(cams.cam_xpos + ((int64_t)w * cams.ncam + cam) * 3);
  p->rorg = cams.ray_direction ? cams.ray_origin + 3 * (int64_t)ray_offset : nullptr;
  p->rdir = cams.ray_direction ? cams.ray_direction + 3 * (int64_t)ray_offset : nullptr;
  return true;
}

MJ_DEV void write_geometry(const MjpbrOutputs& out, int w, const Pixel& p, const Hit& hit, V3 dir_cam) {
  if (p.want_seg) {
    int2 seg = hit.geom_id != -1 ? make_int2(hit.geom_id, 5) : make_int2(-1, -1);
    reinterpret_cast<int2*>(out.seg)[(int64_t)w * out.seg_stride + out.seg_adr[p.camid] + p.pixelid] = seg;
  }
  if (p.want_depth) {
    float depth = hit.geom_id != -1 ? hit.dist * -dir_cam.z : 0.0f;
    out.depth[(int64_t)w * out.depth_stride + out.depth_adr[p.camid] + p.pixelid] = depth;
  }
}

}  // namespace mjpbr
