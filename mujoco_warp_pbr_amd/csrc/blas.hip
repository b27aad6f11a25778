// blas.hip — mesh BVH construction on the GPU (replaces the mesh BVHs MuJoCo Warp builds for the
// reference's render context, reference io.py:511-528, and its re-framed copies, io.py:138-211).
//
// The tree is the renderer's implicit perfect binary tree (include/mjpbr.h): triangles are
// assigned to 2^depth leaves of MJPBR_BLAS_LEAF slots by recursive median split of their
// centroids along the widest axis. Each level is one launch with a workgroup per node: the
// node's centroid bounds pick the axis, its segment of the triangle permutation is sorted by
// (centroid coordinate, position in the segment) - a stable order, so the result does not depend
// on the sort - and split so both halves fit their subtrees. Leaves then write the triangle
// slots {v0, face}, {v1 - v0}, {v2 - v0} and their boxes (inflated in double precision), and the
// inner boxes are merged bottom-up, empty subtrees not widening their parents.
#pragma clang fp contract(off)
#include "abi_util.hpp"

namespace mjpbr {

constexpr int kSplitThreads = 256;
constexpr int kLdsSortMax = 2048;  // segment entries sorted in LDS (larger ones in the workspace)
constexpr float kEmptyBox = 1.0e30f;

struct SortEntry {
  double key;
  int pos;   // position in the segment before sorting (tie-break: stable order)
  int face;  // payload
};

MJ_DEV bool entry_less(const SortEntry& a, const SortEntry& b) { return a.key < b.key || (a.key == b.key && a.pos < b.pos); }

// Centroids (v0 + v1 + v2) / 3 in double, summed in vertex order; the identity permutation and
// the root's segment.
__global__ void blas_centroid_kernel(const float* verts, const int32_t* faces, int nface, double* cen, int* perm, int2* root) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f == 0) *root = make_int2(0, nface);
  if (f >= nface) return;
  const int32_t* fv = faces + 3 * f;
  for (int k = 0; k < 3; ++k) {
    double s = (double)verts[3 * fv[0] + k] + (double)verts[3 * fv[1] + k];
    cen[3 * f + k] = (s + (double)verts[3 * fv[2] + k]) / 3.0;
  }
  perm[f] = f;
}

// Bitonic sort of n entries (padded to a power of two p with +inf keys) by the block's threads.
MJ_DEV void bitonic_sort(SortEntry* e, int p) {
  for (int k = 2; k <= p; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < p; i += blockDim.x) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          if (entry_less(e[l], e[i]) == up) {
            SortEntry t = e[i];
            e[i] = e[l];
            e[l] = t;
          }
        }
      }
      __syncthreads();
    }
  }
}

// One level of the split: node k of this level owns perm[seg[k].x .. + seg[k].y); writes the
// segments of its two children to next[2k], next[2k + 1]. `half_cap` is the number of triangle
// slots of one child subtree.
__global__ void __launch_bounds__(kSplitThreads) blas_split_kernel(const double* cen, int* perm, const int2* seg, int2* next,
                                                                  int half_cap, SortEntry* scratch) {
  __shared__ SortEntry lds_sort[kLdsSortMax];
  __shared__ double red[6][kSplitThreads];
  const int node = blockIdx.x;
  const int start = seg[node].x, n = seg[node].y;
  if (n == 0) {
    if (threadIdx.x == 0) next[2 * node] = next[2 * node + 1] = make_int2(start, 0);
    return;
  }
  // centroid bounds
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < n; i += kSplitThreads) {
    const double* c = cen + 3 * perm[start + i];
    for (int k = 0; k < 3; ++k) {
      lo[k] = fmin(lo[k], c[k]);
      hi[k] = fmax(hi[k], c[k]);
    }
  }
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = lo[k];
    red[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int s = kSplitThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  const double ex = red[3][0] - red[0][0], ey = red[4][0] - red[1][0], ez = red[5][0] - red[2][0];
  const int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);  // first widest axis
  // stable sort of the segment by the axis coordinate
  int p = 1;
  while (p < n) p <<= 1;
  SortEntry* e = p <= kLdsSortMax ? lds_sort : scratch + 2 * (int64_t)start;
  for (int i = threadIdx.x; i < p; i += kSplitThreads) {
    if (i < n) {
      const int f = perm[start + i];
      e[i] = SortEntry{cen[3 * f + axis], i, f};
    } else {
      e[i] = SortEntry{INFINITY, i, -1};
    }
  }
  __syncthreads();
  bitonic_sort(e, p);
  for (int i = threadIdx.x; i < n; i += kSplitThreads) perm[start + i] = e[i].face;
  if (threadIdx.x == 0) {
    int nl = min((n + 1) / 2, half_cap);
    nl = max(nl, n - half_cap);
    next[2 * node] = make_int2(start, nl);
    next[2 * node + 1] = make_int2(start + nl, n - nl);
  }
}

// Leaf k: its triangle slots and its box (double, inflated like the scene BVH's boxes).
__global__ void blas_leaf_kernel(const float* verts, const int32_t* faces, const int* perm, const int2* seg, int nleaf, float* tri,
                                 double* box) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nleaf) return;
  const int start = seg[k].x, n = seg[k].y;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int j = 0; j < MJPBR_BLAS_LEAF; ++j) {
    float* t = tri + 12 * ((int64_t)k * MJPBR_BLAS_LEAF + j);
    if (j >= n) {
      for (int q = 0; q < 12; ++q) t[q] = 0.0f;
      t[3] = __int_as_float(-1);
      continue;
    }
    const int f = perm[start + j];
    const float* a = verts + 3 * faces[3 * f];
    const float* b = verts + 3 * faces[3 * f + 1];
    const float* c = verts + 3 * faces[3 * f + 2];
    for (int q = 0; q < 3; ++q) {
      t[q] = a[q];
      t[4 + q] = b[q] - a[q];
      t[8 + q] = c[q] - a[q];
      lo[q] = fmin(lo[q], (double)fminf(fminf(a[q], b[q]), c[q]));
      hi[q] = fmax(hi[q], (double)fmaxf(fmaxf(a[q], b[q]), c[q]));
    }
    t[3] = __int_as_float(f);
    t[7] = t[11] = 0.0f;
  }
  double* out = box + 6 * (int64_t)k;
  if (n == 0) {
    for (int q = 0; q < 6; ++q) out[q] = (double)kEmptyBox;
    return;
  }
  const double ext = fmax(hi[0] - lo[0], fmax(hi[1] - lo[1], hi[2] - lo[2]));
  for (int q = 0; q < 3; ++q) {
    const double pad = 1.0e-5 * ext + 1.0e-6 * (fabs(lo[q]) + fabs(hi[q])) + 1.0e-6;
    out[q] = lo[q] - pad;
    out[3 + q] = hi[q] + pad;
  }
}

MJ_DEV bool box_empty(const double* b) { return b[0] == (double)kEmptyBox; }

// Inner nodes of one level: node i (of `count`, heap index first + i) stores its children's
// boxes and writes its own (the union of the non-empty children) for the level above.
__global__ void blas_inner_kernel(const double* child, int count, int first, float* inner, double* parent) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const double* c0 = child + 12 * (int64_t)i;
  const double* c1 = c0 + 6;
  float* o = inner + 12 * (int64_t)(first + i);
  for (int q = 0; q < 6; ++q) {
    o[q] = (float)c0[q];
    o[6 + q] = (float)c1[q];
  }
  double* p = parent + 6 * (int64_t)i;
  const bool e0 = box_empty(c0), e1 = box_empty(c1);
  for (int q = 0; q < 3; ++q) {
    p[q] = e0 ? c1[q] : (e1 ? c0[q] : fmin(c0[q], c1[q]));
    p[3 + q] = e0 ? c1[3 + q] : (e1 ? c0[3 + q] : fmax(c0[3 + q], c1[3 + q]));
  }
}

struct BlasWork {
  double* cen;
  int* perm;
  int2* seg_a;
  int2* seg_b;
  SortEntry* scratch;
  double* box_a;
  double* box_b;
};

int64_t blas_layout(int nface, int depth, BlasWork* w, char* base) {
  const int64_t nleaf = 1ll << depth;
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* p = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return p;
  };
  BlasWork t;
  t.cen = reinterpret_cast<double*>(take(24ll * nface));
  t.perm = reinterpret_cast<int*>(take(4ll * nface));
  t.seg_a = reinterpret_cast<int2*>(take(8 * nleaf));
  t.seg_b = reinterpret_cast<int2*>(take(8 * nleaf));
  t.scratch = reinterpret_cast<SortEntry*>(take(2ll * (int64_t)sizeof(SortEntry) * nface));
  t.box_a = reinterpret_cast<double*>(take(48 * nleaf));
  t.box_b = reinterpret_cast<double*>(take(48 * nleaf));
  if (w) *w = t;
  return off;
}

}  // namespace mjpbr

using namespace mjpbr;

extern "C" int64_t mjpbr_mesh_bvh_workspace(int nface, int depth) {
  if (nface < 0 || depth < 1 || depth > 28) return set_error(MJPBR_EINVAL, "mesh_bvh_workspace: bad nface %d or depth %d", nface, depth);
  return blas_layout(nface, depth, nullptr, nullptr);
}

extern "C" int mjpbr_mesh_bvh_build(const float* verts, const int32_t* faces, int nface, int depth, float* inner, float* tri, void* work,
                                    int64_t work_bytes, void* stream) {
  if (nface < 1 || depth < 1 || depth > 28) return set_error(MJPBR_EINVAL, "mesh_bvh_build: bad nface %d or depth %d", nface, depth);
  if ((int64_t)MJPBR_BLAS_LEAF << depth < nface)
    return set_error(MJPBR_EINVAL, "mesh_bvh_build: %d faces do not fit 2^%d leaves of %d", nface, depth, MJPBR_BLAS_LEAF);
  if (work_bytes < blas_layout(nface, depth, nullptr, nullptr)) return set_error(MJPBR_EINVAL, "mesh_bvh_build: workspace too small");
  BlasWork w;
  blas_layout(nface, depth, &w, static_cast<char*>(work));
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(blas_centroid_kernel, dim3((nface + 255) / 256), dim3(256), 0, st, verts, faces, nface, w.cen, w.perm, w.seg_a);
  if (int rc = check_launch("blas_centroid_kernel")) return rc;
  int2 *seg = w.seg_a, *next = w.seg_b;
  for (int level = 0; level < depth; ++level) {
    const int half_cap = MJPBR_BLAS_LEAF << (depth - level - 1);
    hipLaunchKernelGGL(blas_split_kernel, dim3(1u << level), dim3(kSplitThreads), 0, st, w.cen, w.perm, seg, next, half_cap, w.scratch);
    if (int rc = check_launch("blas_split_kernel")) return rc;
    int2* t = seg;
    seg = next;
    next = t;
  }
  const int nleaf = 1 << depth;
  hipLaunchKernelGGL(blas_leaf_kernel, dim3((nleaf + 127) / 128), dim3(128), 0, st, verts, faces, w.perm, seg, nleaf, tri, w.box_a);
  if (int rc = check_launch("blas_leaf_kernel")) return rc;
  double *child = w.box_a, *parent = w.box_b;
  for (int level = depth - 1; level >= 0; --level) {
    const int count = 1 << level;
    hipLaunchKernelGGL(blas_inner_kernel, dim3((count + 127) / 128), dim3(128), 0, st, child, count, count - 1, inner, parent);
    if (int rc = check_launch("blas_inner_kernel")) return rc;
    double* t = child;
    child = parent;
    parent = t;
  }
  return MJPBR_OK;
}
