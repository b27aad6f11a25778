// bvh.hip — per-world scene BVH build / refit (replaces the scene BVH of MuJoCo Warp's render
// context and `mujoco_warp.refit_bvh`, reference io.py:511-528, README.md:50-52).
//
// Layout: an implicit perfect binary tree of 2^depth leaves per world (see trace.hpp). Build
// assigns geoms to leaves by recursive median split of their box centres along the widest
// axis (balanced, so the depth is ceil(log2(nbvh))); refit keeps the assignment and recomputes
// the boxes bottom-up. One workgroup per world, up to 4096 geoms.
#pragma clang fp contract(off)
#include <algorithm>
#include "abi_util.hpp"

namespace mjpbr {

constexpr int kMaxBvhGeoms = 4096;  // 2^12 leaves
constexpr float kHugeExtent = 1.0e8f;
constexpr float kEmptyLeaf = 1.0e30f;

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

// (begin, count) of node k of `level` in the leaf-order permutation of n geoms. The split sizes
// depend on n and the depth only: a node of c geoms gives its left child min(ceil(c / 2), half its
// leaves) and its right child the rest, so every level of the tree is known before any geom moves.
MJ_DEV int2 scene_segment(int n, int D, int level, int k) {
  int begin = 0, count = n;
  for (int l = 0; l < level; ++l) {
    const int nl = min((count + 1) / 2, 1 << (D - l - 1));
    if ((k >> (level - l - 1)) & 1) {
      begin += nl;
      count -= nl;
    } else {
      count = nl;
    }
  }
  return make_int2(begin, count);
}

MJ_DEV float clamp_huge(float x) { return fminf(fmaxf(x, -kHugeExtent), kHugeExtent); }

// Float -> unsigned with the same order (-0 and +0 equal).
MJ_DEV uint32_t order_bits(float f) {
  const uint32_t u = __float_as_uint(f + 0.0f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// First widest axis of extents e.
MJ_DEV int widest_axis(float e0, float e1, float e2) {
  int axis = e1 > e0 ? 1 : 0;
  if (e2 > (axis ? e1 : e0)) axis = 2;
  return axis;
}

constexpr int kBuildThreads = 256;
constexpr int kSmallSegment = 64;  // segments up to this size are sorted by one thread each

// Sorts seg[0 .. c) (geom slots) by their centre on the widest axis of the segment's centres,
// ties in current order, with the block's threads: bitonic sort of (key, position) pairs.
MJ_DEV void block_sort_segment(const float* cen, int n, int* seg, int c, unsigned long long* sk, float (*red)[kBuildThreads]) {
  const int tid = threadIdx.x;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = tid; i < c; i += kBuildThreads)
    for (int a = 0; a < 3; ++a) {
      const float v = cen[a * n + seg[i]];
      lo[a] = fminf(lo[a], v);
      hi[a] = fmaxf(hi[a], v);
    }
  for (int a = 0; a < 3; ++a) {
    red[a][tid] = lo[a];
    red[3 + a][tid] = hi[a];
  }
  __syncthreads();
  for (int s = kBuildThreads / 2; s > 0; s >>= 1) {
    if (tid < s)
      for (int a = 0; a < 3; ++a) {
        red[a][tid] = fminf(red[a][tid], red[a][tid + s]);
        red[3 + a][tid] = fmaxf(red[3 + a][tid], red[3 + a][tid + s]);
      }
    __syncthreads();
  }
  const int axis = widest_axis(red[3][0] - red[0][0], red[4][0] - red[1][0], red[5][0] - red[2][0]);
  int p = 1;
  while (p < c) p <<= 1;
  for (int i = tid; i < p; i += kBuildThreads)
    sk[i] = i < c ? ((unsigned long long)order_bits(cen[axis * n + seg[i]]) << 32) | (unsigned)i : ~0ull;
  __syncthreads();  // also orders every read of red[][0] before the next segment's writes
  for (int k = 2; k <= p; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < p; i += kBuildThreads) {
        const int l = i ^ j;
        if (l > i) {
          const unsigned long long x = sk[i], y = sk[l];
          if ((y < x) == ((i & k) == 0)) {
            sk[i] = y;
            sk[l] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = tid; i < c; i += kBuildThreads) sk[i] = (unsigned long long)(unsigned)seg[(unsigned)sk[i]];
  __syncthreads();
  for (int i = tid; i < c; i += kBuildThreads) seg[i] = (int)(unsigned)sk[i];
  __syncthreads();
}

// The same order for a segment of at most kSmallSegment slots, by one thread (insertion sort).
MJ_DEV void thread_sort_segment(const float* cen, int n, int* seg, int c, float* key) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < c; ++i)
    for (int a = 0; a < 3; ++a) {
      const float v = cen[a * n + seg[i]];
      lo[a] = fminf(lo[a], v);
      hi[a] = fmaxf(hi[a], v);
    }
  const int axis = widest_axis(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
  for (int i = 0; i < c; ++i) key[i] = cen[axis * n + seg[i]];
  for (int i = 1; i < c; ++i) {
    const int v = seg[i];
    const float k = key[i];
    int j = i - 1;
    while (j >= 0 && key[j] > k) {
      seg[j + 1] = seg[j];
      key[j + 1] = key[j];
      --j;
    }
    seg[j + 1] = v;
    key[j + 1] = k;
  }
}

// Leaf assignment of world `blockIdx.x`: recursive median split of the geoms' box centres along
// the widest axis, level by level (one block a world). The world's tlas_box holds the centres
// meanwhile (3 n floats <= 12 (2^depth - 1)); scene_bvh_refit_kernel then writes the boxes.
__global__ void __launch_bounds__(kBuildThreads) scene_bvh_build_kernel(MjpbrGeometry g) {
  __shared__ int perm[kMaxBvhGeoms];
  __shared__ unsigned long long sk[kMaxBvhGeoms];
  __shared__ float red[6][kBuildThreads];
  const int w = blockIdx.x;
  const int n = g.nbvh, D = g.tlas_depth, nleaf = 1 << D;
  float* cen = g.tlas_box + (int64_t)w * (nleaf - 1) * 12;
  for (int i = threadIdx.x; i < n; i += kBuildThreads) {
    float lo[3], hi[3];
    geom_aabb(g, w, g.bvh_geom[i], lo, hi);
    for (int a = 0; a < 3; ++a) cen[a * n + i] = 0.5f * (clamp_huge(lo[a]) + clamp_huge(hi[a]));
    perm[i] = i;
  }
  __syncthreads();
  for (int level = 0; level < D; ++level) {
    const int nodes = 1 << level;
    // the left-most node of a level holds the most geoms
    if (scene_segment(n, D, level, 0).y > kSmallSegment) {
      for (int k = 0; k < nodes; ++k) {
        const int2 sg = scene_segment(n, D, level, k);
        if (sg.y > 1) block_sort_segment(cen, n, perm + sg.x, sg.y, sk, red);
      }
    } else {
      for (int k = threadIdx.x; k < nodes; k += kBuildThreads) {
        const int2 sg = scene_segment(n, D, level, k);
        if (sg.y > 1) thread_sort_segment(cen, n, perm + sg.x, sg.y, reinterpret_cast<float*>(sk) + sg.x);
      }
      __syncthreads();
    }
  }
  int* leaf = g.tlas_leaf + (int64_t)w * nleaf;
  for (int s = threadIdx.x; s < nleaf; s += kBuildThreads) {
    const int2 sg = scene_segment(n, D, D, s);
    leaf[s] = sg.y > 0 ? g.bvh_geom[perm[sg.x]] : -1;
  }
}

MJ_DEV void store_box(float* dst, const float* lo, const float* hi) {
  for (int a = 0; a < 3; ++a) {
    dst[a] = lo[a];
    dst[3 + a] = hi[a];
  }
}

// Boxes of world `blockIdx.x`, bottom-up a level at a time: inner node i stores its children's
// boxes. An empty leaf is the point (1e30, 1e30, 1e30), beyond every ray's cut; a child subtree's
// box is the union of its non-empty children's, so empty leaves never widen an ancestor.
__global__ void __launch_bounds__(256) scene_bvh_refit_kernel(MjpbrGeometry g) {
  const int w = blockIdx.x;
  const int D = g.tlas_depth, nleaf = 1 << D, ninner = nleaf - 1;
  float* boxes = g.tlas_box + (int64_t)w * ninner * 12;
  const int* leaf = g.tlas_leaf + (int64_t)w * nleaf;
  for (int i = (nleaf >> 1) - 1 + threadIdx.x; i < ninner; i += blockDim.x)
    for (int side = 0; side < 2; ++side) {
      const int gi = leaf[2 * i + 1 + side - ninner];
      float lo[3] = {kEmptyLeaf, kEmptyLeaf, kEmptyLeaf}, hi[3] = {kEmptyLeaf, kEmptyLeaf, kEmptyLeaf};
      if (gi >= 0) geom_aabb(g, w, gi, lo, hi);
      store_box(boxes + 12 * i + 6 * side, lo, hi);
    }
  for (int level = D - 2; level >= 0; --level) {
    __syncthreads();  // the level below is written
    for (int i = (1 << level) - 1 + threadIdx.x; i < (2 << level) - 1; i += blockDim.x)
      for (int side = 0; side < 2; ++side) {
        const float* cb = boxes + 12 * (2 * i + 1 + side);
        const bool e0 = cb[0] == kEmptyLeaf, e1 = cb[6] == kEmptyLeaf;
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
          lo[a] = e0 ? cb[6 + a] : (e1 ? cb[a] : fminf(cb[a], cb[6 + a]));
          hi[a] = e0 ? cb[9 + a] : (e1 ? cb[3 + a] : fmaxf(cb[3 + a], cb[9 + a]));
        }
        store_box(boxes + 12 * i + 6 * side, lo, hi);
      }
  }
}

}  // namespace mjpbr

using namespace mjpbr;

extern "C" int mjpbr_scene_bvh(const MjpbrGeometry* geo, int refit, void* stream) {
  if (geo->nbvh < 1 || geo->nbvh > kMaxBvhGeoms)
    return set_error(MJPBR_EINVAL, "scene BVH holds 1..%d geoms, got %d", kMaxBvhGeoms, geo->nbvh);
  if (geo->tlas_depth < 1 || geo->tlas_depth > 12 || (1 << geo->tlas_depth) < geo->nbvh)
    return set_error(MJPBR_EINVAL, "tlas_depth %d too small for %d geoms", geo->tlas_depth, geo->nbvh);
  if (geo->nworld < 1) return MJPBR_OK;
  hipStream_t st = (hipStream_t)stream;
  if (!refit) {
    hipLaunchKernelGGL(scene_bvh_build_kernel, dim3(geo->nworld), dim3(kBuildThreads), 0, st, *geo);
    if (int rc = check_launch("scene_bvh_build_kernel")) return rc;
  }
  // a thread per inner node of the lowest level, 64 .. 256 a world
  const int threads = std::min(256, std::max(64, 1 << (geo->tlas_depth - 1)));
  hipLaunchKernelGGL(scene_bvh_refit_kernel, dim3(geo->nworld), dim3(threads), 0, st, *geo);
  return check_launch("scene_bvh_refit_kernel");
}
