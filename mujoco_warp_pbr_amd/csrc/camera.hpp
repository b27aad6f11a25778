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
};

struct Pixel {
  int camid, cam, pixelid, px, py, width, height, rayid;
  bool want_rgb, want_depth, want_seg;
  int projection;
  float fovy;
  float sensor[2];
  float intr[4];
  float xmat[9];
  V3 xpos;
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
    double h = tan((double)p.fovy * (3.14159265358979323846 / 360.0));
    float ax = (float)(2.0 * h * (double)W / (double)H);
    float ay = (float)(2.0 * h);
    d = V3{u * ax, -v * ay, -1.0f};
  }
  *dir = normalize(d);
  *off = v3s(0.0f);
}

MJ_DEV void primary_ray(const Pixel& p, const MjpbrSettings& s, int subdiv, int sx, int sy, V3* o, V3* d,
                        float* max_dist, V3* dir_cam) {
  V3 dc, oc;
  camera_frame_ray(p, subdiv, sx, sy, &dc, &oc);
  *o = p.xpos + mat_mul(p.xmat, oc);
  *d = mat_mul(p.xmat, dc);
  *dir_cam = dc;
  *max_dist = s.zfar > 0.0f ? s.zfar / fmaxf(-dc.z, 1.0e-6f) : kMaxVal;
}

MJ_DEV void pixel_cone(const Pixel& p, float* spread, float* width) {
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
  const float* in = cams.cam_intrinsic + ((int64_t)(w % cams.cam_intrinsic_n) * cams.ncam + cam) * 4;
#pragma unroll
  for (int k = 0; k < 4; ++k) p->intr[k] = in[k];
  const float* xm = cams.cam_xmat + ((int64_t)w * cams.ncam + cam) * 9;
#pragma unroll
  for (int k = 0; k < 9; ++k) p->xmat[k] = xm[k];
  p->xpos = ld3(cams.cam_xpos + ((int64_t)w * cams.ncam + cam) * 3);
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
