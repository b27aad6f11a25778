// bvh.hip — per-world scene BVH build / refit (replaces the scene BVH of MuJoCo Warp's render
// context and `mujoco_warp.refit_bvh`, reference io.py:511-528, README.md:50-52).
//
// Layout: an implicit perfect binary tree of 2^depth leaves per world (see trace.hpp). Build
// assigns geoms to leaves by recursive median split of their box centres along the widest
// axis (balanced, so the depth is ceil(log2(nbvh))); refit keeps the assignment and recomputes
// the boxes bottom-up. One thread per world: scene BVHs hold tens of geoms.
#pragma clang fp contract(off)
#include "abi_util.hpp"

namespace mjpbr {

constexpr int kMaxBvhGeoms = 256;
constexpr float kHugeExtent = 1.0e8f;

// World-space box of geom gi, slightly enlarged so the slab test never culls a true hit.
MJ_DEV void geom_aabb(const MjpbrGeometry& g, int w, int gi, float* lo, float* hi) {
  int type = g.geom_type[gi];
  const float* sz = g.geom_size + (int64_t)(w % g.geom_size_n) * 3 * g.ngeom + 3 * gi;
  float framed[9];
  const float* R = traced_rotation(g, type, g.geom_dataid[(int64_t)(w % g.geom_dataid_n) * g.ngeom + gi],
                                   g.geom_xmat + ((int64_t)w * g.ngeom + gi) * 9, framed);
  V3 pos = ld3(g.geom_xpos + ((int64_t)w * g.ngeom + gi) * 3);
  V3 c = v3s(0.0f), e = v3s(0.0f);
  switch (type) {
    case MJPBR_GEOM_PLANE:
      e = V3{sz[0] > 0.0f ? sz[0] : kHugeExtent, sz[1] > 0.0f ? sz[1] : kHugeExtent, 0.0f};
      break;
    case MJPBR_GEOM_SPHERE:
      e = v3s(sz[0]);
      break;
    case MJPBR_GEOM_CAPSULE:
      e = V3{sz[0], sz[0], sz[1] + sz[0]};
      break;
    case MJPBR_GEOM_CYLINDER:
      e = V3{sz[0], sz[0], sz[1]};
      break;
    case MJPBR_GEOM_ELLIPSOID:
    case MJPBR_GEOM_BOX:
      e = V3{sz[0], sz[1], sz[2]};
      break;
    case MJPBR_GEOM_MESH:
    case MJPBR_GEOM_SDF:
    case MJPBR_GEOM_HFIELD: {
      int mid = g.geom_dataid[(int64_t)(w % g.geom_dataid_n) * g.ngeom + gi];
      if (mid >= 0 && type == MJPBR_GEOM_HFIELD) mid += g.nmesh;
      if (mid >= 0) {
        const float* a = g.mesh_aabb + 6 * mid;
        c = V3{0.5f * (a[0] + a[3]), 0.5f * (a[1] + a[4]), 0.5f * (a[2] + a[5])};
        e = V3{0.5f * (a[3] - a[0]), 0.5f * (a[4] - a[1]), 0.5f * (a[5] - a[2])};
      }
      break;
    }
    default:
      e = v3s(kHugeExtent);
  }
  V3 wc = pos + mat_mul(R, c);
  float we[3];
  for (int i = 0; i < 3; ++i)
    we[i] = fabsf(R[3 * i]) * e.x + fabsf(R[3 * i + 1]) * e.y + fabsf(R[3 * i + 2]) * e.z;
  float cc[3] = {wc.x, wc.y, wc.z};
  float m = fmaxf(we[0], fmaxf(we[1], we[2]));
  for (int i = 0; i < 3; ++i) {
    float pad = 1.0e-5f * m + 1.0e-6f * (fabsf(cc[i]) + we[i]) + 1.0e-6f;
    lo[i] = cc[i] - we[i] - pad;
    hi[i] = cc[i] + we[i] + pad;
  }
}

__global__ void scene_bvh_kernel(MjpbrGeometry g, int refit) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= g.nworld) return;
  const int D = g.tlas_depth, nleaf = 1 << D, ninner = nleaf - 1;
  float* boxes = g.tlas_box + (int64_t)w * ninner * 12;
  int* leaf = g.tlas_leaf + (int64_t)w * nleaf;
  if (!refit) {
    int perm[kMaxBvhGeoms];
    float key[kMaxBvhGeoms];
    for (int i = 0; i < g.nbvh; ++i) perm[i] = i;
    int st_node[32], st_level[32], st_begin[32], st_count[32];
    int sp = 0;
    st_node[0] = 0;
    st_level[0] = 0;
    st_begin[0] = 0;
    st_count[0] = g.nbvh;
    sp = 1;
    while (sp > 0) {
      --sp;
      int node = st_node[sp], level = st_level[sp], begin = st_begin[sp], count = st_count[sp];
      if (level == D) {
        leaf[node - ninner] = count > 0 ? g.bvh_geom[perm[begin]] : -1;
        continue;
      }
      if (count == 0) {
        int first = ((node + 1) << (D - level)) - 1;
        for (int k = 0; k < (1 << (D - level)); ++k) leaf[first - ninner + k] = -1;
        continue;
      }
      // widest axis of the box centres
      float cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int i = begin; i < begin + count; ++i) {
        float lo[3], hi[3];
        geom_aabb(g, w, g.bvh_geom[perm[i]], lo, hi);
        for (int a = 0; a < 3; ++a) {
          float cen = 0.5f * (fminf(fmaxf(lo[a], -kHugeExtent), kHugeExtent) + fminf(fmaxf(hi[a], -kHugeExtent), kHugeExtent));
          cl[a] = fminf(cl[a], cen);
          ch[a] = fmaxf(ch[a], cen);
        }
      }
      int axis = 0;
      if (ch[1] - cl[1] > ch[axis] - cl[axis]) axis = 1;
      if (ch[2] - cl[2] > ch[axis] - cl[axis]) axis = 2;
      for (int i = begin; i < begin + count; ++i) {
        float lo[3], hi[3];
        geom_aabb(g, w, g.bvh_geom[perm[i]], lo, hi);
        key[i] = 0.5f * (fminf(fmaxf(lo[axis], -kHugeExtent), kHugeExtent) + fminf(fmaxf(hi[axis], -kHugeExtent), kHugeExtent));
      }
      for (int i = begin + 1; i < begin + count; ++i) {  // insertion sort (stable)
        int p = perm[i];
        float k = key[i];
        int j = i - 1;
        while (j >= begin && key[j] > k) {
          perm[j + 1] = perm[j];
          key[j + 1] = key[j];
          --j;
        }
        perm[j + 1] = p;
        key[j + 1] = k;
      }
      int half = 1 << (D - level - 1);
      int nl = min((count + 1) / 2, half);
      st_node[sp] = 2 * node + 2;
      st_level[sp] = level + 1;
      st_begin[sp] = begin + nl;
      st_count[sp] = count - nl;
      ++sp;
      st_node[sp] = 2 * node + 1;
      st_level[sp] = level + 1;
      st_begin[sp] = begin;
      st_count[sp] = nl;
      ++sp;
    }
  }
  // boxes, bottom-up: inner node i holds its children's boxes
  for (int i = ninner - 1; i >= 0; --i) {
    for (int side = 0; side < 2; ++side) {
      int c = 2 * i + 1 + side;
      float lo[3], hi[3];
      if (c >= ninner) {
        int gi = leaf[c - ninner];
        if (gi >= 0) {
          geom_aabb(g, w, gi, lo, hi);
        } else {  // empty: a point far away, never entered before the cut
          for (int a = 0; a < 3; ++a) lo[a] = hi[a] = 1.0e30f;
        }
      } else {
        const float* cb = boxes + 12 * c;
        for (int a = 0; a < 3; ++a) {
          lo[a] = fminf(cb[a], cb[6 + a]);
          hi[a] = fmaxf(cb[3 + a], cb[9 + a]);
        }
      }
      float* dst = boxes + 12 * i + 6 * side;
      for (int a = 0; a < 3; ++a) {
        dst[a] = lo[a];
        dst[3 + a] = hi[a];
      }
    }
  }
}

}  // namespace mjpbr

using namespace mjpbr;

extern "C" int mjpbr_scene_bvh(const MjpbrGeometry* geo, int refit, void* stream) {
  if (geo->nbvh < 1 || geo->nbvh > kMaxBvhGeoms)
    return set_error(MJPBR_EINVAL, "scene BVH holds 1..%d geoms, got %d", kMaxBvhGeoms, geo->nbvh);
  if ((1 << geo->tlas_depth) < geo->nbvh || geo->tlas_depth < 1 || geo->tlas_depth > 12)
    return set_error(MJPBR_EINVAL, "tlas_depth %d too small for %d geoms", geo->tlas_depth, geo->nbvh);
  int threads = 64;
  int blocks = (geo->nworld + threads - 1) / threads;
  hipLaunchKernelGGL(scene_bvh_kernel, dim3(blocks), dim3(threads), 0, (hipStream_t)stream, *geo, refit);
  return check_launch("scene_bvh_kernel");
}
