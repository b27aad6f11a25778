// trace.hpp — closest-hit and any-hit ray queries over the per-world scene BVH (staged in LDS)
// and the shared mesh BVHs (HBM). Semantics follow the reference's ordered traversal,
// traverse.py:194-323: pruning cut dist*(1+tol)+tol, geoms whose rgba alpha is 0 skipped,
// back-face hits culled, ties within tol go to the higher geom id; any-hit returns at the first
// hit closer than max_dist.
//
// Both BVHs are implicit perfect binary trees (heap order) whose inner nodes hold their two
// children's boxes. Traversal is nearest-child-first with a one-register "trail": bit k set
// means the far child at level k is still pending; popping recomputes the pending node from
// the heap index, so no stack memory is used. A closest-hit query re-tests the popped box
// against the cut, which may have shrunk since the push; an any-hit query's cut is fixed, so
// the popped child is known to intersect and is entered without another box load.
#pragma once
#include "../../include/mjpbr.h"
#include "geom.hpp"

#include <type_traits>

namespace mjpbr {

// Staged geom record: 16 floats.
//   [0..2] xpos  [3] size.x  [4..12] xmat  [13] size.y  [14] size.z  [15] bits
// bits = type | alpha_zero << 8 | (mesh + 1) << 9. SDF geoms (traverse.py:185-196) and height
// fields (traverse.py:161-165) are staged as type MESH: mesh = dataid for an SDF geom, nmesh +
// dataid for a height field's triangle surface.
constexpr int kGeomRec = 16;
// Staged light record: 24 floats (see stage_light).
constexpr int kLightRec = 24;
// Staged surface record per geom: 16 floats (see stage_surface).
constexpr int kSurfRec = 16;
// Staged environment record per world: 4 floats (see stage_world_to).
constexpr int kEnvRec = 4;
// Leaf ids are padded to a multiple of 4 so the light records after them are 16-byte aligned.
MJ_HOST_DEV inline int staged_leaf_slots(int nleaf) { return (nleaf + 3) & ~3; }

struct Hit {
  int geom_id;  // -1: miss
  float dist;
  V3 normal;
  float bary_u, bary_v;  // weights of the triangle's first and second vertex (meshes)
  int face_id;
  int mesh_id;  // traced mesh; read only for mesh and SDF geoms (height fields shade as primitives)
};

struct SceneView {
  const float* tlas_box;  // [ninner][12] (LDS)
  const int* tlas_leaf;   // [nleaf] (LDS)
  const int* direct;      // [ndirect] geoms tested before the BVH (LDS)
  int ndirect;
  const float* geoms;     // [ngeom][16] (LDS)
  const float* lights;    // [nlight][kLightRec] (LDS)
  const float* surfs;     // [ngeom][kSurfRec] (LDS)
  const float* env;       // [kEnvRec]: cos(-yaw), sin(-yaw) of the world's sky rotation (LDS)
  int depth;
  // top levels of the mesh BVHs copied into LDS: mesh m < mesh_cache_n has its first
  // min(mesh_cache_nodes, 2^depth - 1) inner nodes at mesh_cache + 12 * mesh_cache_nodes * m
  const float* mesh_cache;
  int mesh_cache_nodes, mesh_cache_n;
  // mesh BVHs (global)
  const float* blas_box;
  const float* blas_tri;
  const int* blas_depth;
  const int* blas_node_adr;
  const int* blas_tri_adr;
};

MJ_DEV float cut_of(float dist) { return dist * (1.0f + kRayTolRel) + kRayTolAbs; }

typedef float NativeFloat4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const NativeFloat4 LdsFloat4;
typedef __attribute__((address_space(1))) const NativeFloat4 GlobalFloat4;

MJ_DEV float4 to_float4(NativeFloat4 v) { return make_float4(v.x, v.y, v.z, v.w); }

// Inner-node box storage of one BVH: load(i, q) fetches node i's 12 floats.
struct PlainBoxes {
  const float* p;
  MJ_DEV void load(int i, float4 (&q)[3]) const {
    const float4* b = reinterpret_cast<const float4*>(p) + 3 * i;
    q[0] = b[0];
    q[1] = b[1];
    q[2] = b[2];
  }
};
// A mesh BVH whose first `ncached` inner nodes (its top levels) are also in LDS: those are read
// with LDS loads, the others with global loads (explicit address spaces, so neither becomes a
// generic flat load).
struct CachedBoxes {
  const float* lds;
  const float* glob;
  int ncached;
  MJ_DEV void load(int i, float4 (&q)[3]) const {
    if (i < ncached) {
      LdsFloat4* b = (LdsFloat4*)(lds) + 3 * i;
      q[0] = to_float4(b[0]);
      q[1] = to_float4(b[1]);
      q[2] = to_float4(b[2]);
    } else {
      GlobalFloat4* b = (GlobalFloat4*)(glob) + 3 * i;
      q[0] = to_float4(b[0]);
      q[1] = to_float4(b[1]);
      q[2] = to_float4(b[2]);
    }
  }
};

// Visits the leaves of an implicit BVH nearest-first. leaf(slot) returns true to stop.
// kFmaSlabs: box tests as ray_aabb_fma (faster in the leanest real-time instance; slower where its
// six extra live registers spill).
template <bool kRetest, bool kFmaSlabs = false, class Boxes, class CutFn, class LeafFn>
MJ_DEV void traverse(Boxes boxes, int depth, V3 o, V3 rcp, CutFn cut, LeafFn leaf) {
  const int ninner = (1 << depth) - 1;
  const V3 rc = slab_rcp(rcp);
  const V3 orc = V3{o.x * rc.x, o.y * rc.y, o.z * rc.z};
  auto slab = [&](const float* lo, const float* hi, float max_t) {
    return kFmaSlabs ? ray_aabb_fma(orc, rc, lo, hi, max_t) : ray_aabb(o, rcp, lo, hi, max_t);
  };
  int node = 0, level = 0;
  uint32_t trail = 0;
  for (;;) {
    if (level < depth) {
      float4 q[3];
      boxes.load(node, q);
      const float* b = reinterpret_cast<const float*>(q);
      float c = cut();
      float ta = slab(b, b + 3, c);
      float tb = slab(b + 6, b + 9, c);
      if (ta >= 0.0f || tb >= 0.0f) {
        int nearside = (tb >= 0.0f && (ta < 0.0f || tb < ta)) ? 1 : 0;
        if (ta >= 0.0f && tb >= 0.0f) trail |= 1u << level;
        node = 2 * node + 1 + nearside;
        ++level;
        continue;
      }
    } else {
      if (leaf(node - ninner)) return;
    }
    for (;;) {
      if (trail == 0u) return;
      int L = 31 - __builtin_clz(trail);
      trail &= ~(1u << L);
      int anc1 = (node + 1) >> (level - L - 1);
      node = (anc1 ^ 1) - 1;
      level = L + 1;
      if (!kRetest) break;
      float4 q[3];
      boxes.load((node - 1) >> 1, q);
      const float* pb = reinterpret_cast<const float*>(q) + 6 * ((node - 1) & 1);
      if (slab(pb, pb + 3, cut()) >= 0.0f) break;
    }
  }
}

MJ_DEV V3 safe_rcp(V3 d) { return V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z}; }

// Closest triangle of mesh `mesh` (mesh-frame ray) with t < max_t; ties -> lower face id.
// Box storage of mesh `mesh`: its top levels from LDS when cached.
MJ_DEV CachedBoxes mesh_boxes(const SceneView& sv, int mesh) {
  const float* glob = sv.blas_box + 12 * (int64_t)sv.blas_node_adr[mesh];
  if (mesh < sv.mesh_cache_n) return CachedBoxes{sv.mesh_cache + 12 * sv.mesh_cache_nodes * mesh, glob, sv.mesh_cache_nodes};
  return CachedBoxes{glob, glob, 0};
}

MJ_DEV bool mesh_closest(const SceneView& sv, int mesh, V3 lo, V3 ld, float max_t, float* t_out,
                         int* face_out, float* b1_out, float* b2_out, V3* e1_out, V3* e2_out) {
  const CachedBoxes boxes = mesh_boxes(sv, mesh);
  const float* tris = sv.blas_tri + 12 * (int64_t)sv.blas_tri_adr[mesh];
  int depth = sv.blas_depth[mesh];
  float best = INFINITY;
  int best_face = -1;
  float bb1 = 0.0f, bb2 = 0.0f;
  V3 be1 = v3s(0.0f), be2 = v3s(0.0f);
  V3 rcp = safe_rcp(ld);
  traverse<true>(
      boxes, depth, lo, rcp, [&]() { return best == INFINITY ? max_t : fminf(max_t, cut_of(best)); },
      [&](int slot) {
        const float4* tr = reinterpret_cast<const float4*>(tris) + 3 * slot * MJPBR_BLAS_LEAF;
#pragma unroll
        for (int k = 0; k < MJPBR_BLAS_LEAF; ++k) {
          float4 a = tr[3 * k], b = tr[3 * k + 1], c = tr[3 * k + 2];
          int face = __float_as_int(a.w);
          if (face < 0) continue;
          V3 v0 = V3{a.x, a.y, a.z}, e1 = V3{b.x, b.y, b.z}, e2 = V3{c.x, c.y, c.z};
          float t, b1, b2;
          if (ray_triangle(lo, ld, v0, e1, e2, max_t, &t, &b1, &b2)) {
            if (t < best || (t == best && face < best_face)) {
              best = t;
              best_face = face;
              bb1 = b1;
              bb2 = b2;
              be1 = e1;
              be2 = e2;
            }
          }
        }
        return false;
      });
  if (best_face < 0) return false;
  *t_out = best;
  *face_out = best_face;
  *b1_out = bb1;
  *b2_out = bb2;
  *e1_out = be1;
  *e2_out = be2;
  return true;
}

MJ_DEV bool mesh_any(const SceneView& sv, int mesh, V3 lo, V3 ld, float max_t) {
  const CachedBoxes boxes = mesh_boxes(sv, mesh);
  const float* tris = sv.blas_tri + 12 * (int64_t)sv.blas_tri_adr[mesh];
  int depth = sv.blas_depth[mesh];
  bool hit = false;
  V3 rcp = safe_rcp(ld);
  traverse<false>(
      boxes, depth, lo, rcp, [&]() { return max_t; },
      [&](int slot) {
        const float4* tr = reinterpret_cast<const float4*>(tris) + 3 * slot * MJPBR_BLAS_LEAF;
#pragma unroll
        for (int k = 0; k < MJPBR_BLAS_LEAF; ++k) {
          float4 a = tr[3 * k], b = tr[3 * k + 1], c = tr[3 * k + 2];
          if (__float_as_int(a.w) < 0) continue;
          float t, b1, b2;
          if (ray_triangle(lo, ld, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, V3{c.x, c.y, c.z}, max_t, &t, &b1,
                           &b2)) {
            hit = true;
            return true;
          }
        }
        return false;
      });
  return hit;
}

// ---------------------------------------------------------------- mesh any-hit jobs of a wave
// Shadow rays reach mesh geoms in few lanes of a wave at a time, and each such lane walked the
// mesh BVH once per ray while the others waited (a tenth of the lanes active in those walks).
// trace_any_packet therefore defers the mesh geoms a packet reaches until its scene-BVH walk
// ends (primitives tested first may already block the rays), then the wave deals every pending
// (ray, mesh) pair of its lanes out to its lanes, one pair a lane a round. Any-hit answers do not
// depend on the order of the tests, so each ray's answer is unchanged.

// Pending mesh geoms of one lane's packet: geom id in the low kPendShift bits, the rays still to
// test against it above them (0 = free slot).
constexpr int kMeshPending = 4;
constexpr int kPendShift = 20;

MJ_DEV int lane_rank(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
MJ_DEV int pull_int(int lane, int v) { return __builtin_amdgcn_ds_bpermute(lane << 2, v); }
MJ_DEV float pull_float(int lane, float v) { return __int_as_float(__builtin_amdgcn_ds_bpermute(lane << 2, __float_as_int(v))); }

// Job table of each wave of a 256-thread workgroup (LDS): slot j holds the descriptor of the
// round's job j (source lane, ray, pending slot), and its answer in bit 31 after the round.
constexpr int kJobTableBytes = 4 * 64 * 4;

// Tests the pending (ray, mesh) pairs of every lane executing this call, spread over those lanes;
// returns this lane's rays that hit. The lanes executing the call must be the ones that reached
// it together (the exec mask), and every one of them runs the same number of rounds.
template <int N>
MJ_DEV uint32_t run_mesh_jobs(const SceneView& sv, V3 o, const V3 (&d)[N], const float (&tmax)[N],
                              uint32_t (&pend)[kMeshPending], uint32_t hit) {
  __shared__ uint32_t job_table[4][64];
  uint32_t* tab = job_table[(threadIdx.x >> 6) & 3];
  const uint64_t part = __ballot(1);
  const int me = lane_rank(~0ull);
  const int rank = lane_rank(part);
  const int nslot = __builtin_popcountll(part);
#pragma unroll
  for (int p = 0; p < kMeshPending; ++p) pend[p] &= ~(hit << kPendShift);
  for (;;) {
    int cnt = 0;
#pragma unroll
    for (int p = 0; p < kMeshPending; ++p) cnt += __builtin_popcount(pend[p] >> kPendShift);
    // wave-wide count and exclusive prefix of the lanes' job counts (cnt < 32: five bit planes)
    int total = 0, prefix = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
      const uint64_t plane = __ballot((cnt >> b) & 1);
      total += __builtin_popcountll(plane) << b;
      prefix += lane_rank(plane) << b;
    }
    if (total == 0) break;
    // sources: descriptors of the jobs that fall into this round's nslot slots
    uint32_t sent[kMeshPending];
    int j = prefix;
#pragma unroll
    for (int p = 0; p < kMeshPending; ++p) {
      sent[p] = 0u;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        if (!((pend[p] >> (kPendShift + k)) & 1u)) continue;
        if (j < nslot) {
          tab[j] = (uint32_t)me | ((uint32_t)k << 6) | ((uint32_t)p << 9);
          sent[p] |= 1u << k;
        }
        ++j;
      }
      pend[p] &= ~(sent[p] << kPendShift);
    }
    __builtin_amdgcn_wave_barrier();
    // destinations: the lane of rank r runs job r
    const bool has = rank < min(total, nslot);
    const uint32_t desc = has ? tab[rank] : 0u;
    const int src = desc & 63, k = (desc >> 6) & 7, p = (desc >> 9) & 3;
    int gslot = pull_int(src, (int)pend[0]);
#pragma unroll
    for (int q = 1; q < kMeshPending; ++q) {
      const int v = pull_int(src, (int)pend[q]);
      if (p == q) gslot = v;
    }
    // the slot's geom id survives in the low bits even once all its rays were sent
    const V3 so = V3{pull_float(src, o.x), pull_float(src, o.y), pull_float(src, o.z)};
    V3 sd = v3s(0.0f);
    float st = 0.0f;
#pragma unroll
    for (int q = 0; q < N; ++q) {
      const V3 v = V3{pull_float(src, d[q].x), pull_float(src, d[q].y), pull_float(src, d[q].z)};
      const float t = pull_float(src, tmax[q]);
      if (k == q) {
        sd = v;
        st = t;
      }
    }
    if (has) {
      const float* g = sv.geoms + kGeomRec * (gslot & ((1 << kPendShift) - 1));
      const int mid = (__float_as_int(g[15]) >> 9) - 1;
      const float* mat = g + 4;
      const bool h = mesh_any(sv, mid, mat_tmul(mat, so - V3{g[0], g[1], g[2]}), mat_tmul(mat, sd), st);
      tab[rank] = desc | (h ? 0x80000000u : 0u);
    }
    __builtin_amdgcn_wave_barrier();
    // sources: answers of the jobs they sent, in the order they were written
    j = prefix;
#pragma unroll
    for (int q = 0; q < kMeshPending; ++q) {
#pragma unroll
      for (int kk = 0; kk < N; ++kk) {
        if (!((sent[q] >> kk) & 1u)) continue;
        if (tab[j] >> 31) hit |= 1u << kk;
        ++j;
      }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int q = 0; q < kMeshPending; ++q) pend[q] &= ~(hit << kPendShift);
  }
  return hit;
}

// One geom against the ray (traverse.py:131-192). Returns the distance or -1. kMesh = false
// drops the mesh / SDF / height-field branch (for scenes without such geoms).
// kPlaneOnly: the geom is known to be a plane (the direct geoms).
template <bool kAnyHit, bool kMesh = true, bool kPlaneOnly = false>
MJ_DEV float intersect_geom(const SceneView& sv, const float* g, V3 o, V3 d, float max_t, V3* n, float* bu,
                            float* bv, int* face, int* mesh) {
  int bits = __float_as_int(g[15]);
  int type = bits & 0xff;
  V3 pos = V3{g[0], g[1], g[2]};
  const float* mat = g + 4;
  V3 size = V3{g[3], g[13], g[14]};
  if (kPlaneOnly) return type == MJPBR_GEOM_PLANE ? ray_plane(pos, mat, size, o, d, n) : -1.0f;
  switch (type) {
    case MJPBR_GEOM_PLANE:
      return ray_plane(pos, mat, size, o, d, n);
    case MJPBR_GEOM_SPHERE:
      return ray_sphere(pos, size.x * size.x, o, d, n);
    case MJPBR_GEOM_ELLIPSOID:
      return ray_ellipsoid(pos, mat, size, o, d, n);
    case MJPBR_GEOM_CAPSULE:
      return ray_capsule(pos, mat, size, o, d, n);
    case MJPBR_GEOM_CYLINDER:
      return ray_cylinder(pos, mat, size, o, d, n);
    case MJPBR_GEOM_BOX:
      return ray_box(pos, mat, size, o, d, n);
    case MJPBR_GEOM_MESH: {  // also SDF geoms and height fields, staged as meshes
      if (!kMesh) return -1.0f;
      int mid = (bits >> 9) - 1;
      if (mid < 0) return -1.0f;
      V3 lo = mat_tmul(mat, o - pos);
      V3 ld = mat_tmul(mat, d);
      if (kAnyHit) {
        *n = v3s(0.0f);
        return mesh_any(sv, mid, lo, ld, max_t) ? 0.0f : -1.0f;
      }
      float t, b1, b2;
      int f;
      V3 e1, e2;
      if (!mesh_closest(sv, mid, lo, ld, max_t, &t, &f, &b1, &b2, &e1, &e2)) return -1.0f;
      *n = normalize(mat_mul(mat, normalize(cross(e1, e2))));
      *bu = 1.0f - b1 - b2;
      *bv = b1;
      *face = f;
      *mesh = mid;
      return t;
    }
    default:
      return -1.0f;
  }
}

// kFmaSlabs: the scene BVH's box tests as ray_aabb_fma (instances pick it where it measured faster).
template <bool kMesh = true, bool kFmaSlabs = false>
MJ_DEV Hit trace_closest(const SceneView& sv, V3 o, V3 d, float max_dist, bool cull) {
  Hit h;
  h.geom_id = -1;
  h.dist = max_dist;
  h.normal = v3s(0.0f);
  h.bary_u = 0.0f;
  h.bary_v = 0.0f;
  h.face_id = -1;
  h.mesh_id = -1;
  V3 rcp = safe_rcp(d);
  // kPlane: a direct geom (a plane), tested without the other geom types' code
  auto visit = [&](int gi, auto kPlane) {
        if (gi < 0) return false;
        const float* g = sv.geoms + kGeomRec * gi;
        int bits = __float_as_int(g[15]);
        if (bits & 0x100) return false;  // alpha 0
        V3 n = v3s(0.0f);
        float bu = 0.0f, bv = 0.0f;
        int face = -1, mesh = -1;
        float dd = intersect_geom<false, kMesh, decltype(kPlane)::value>(sv, g, o, d, cut_of(h.dist), &n, &bu, &bv,
                                                                         &face, &mesh);
        if (cull && dd >= 0.0f && dot(d, n) > 0.0f) dd = -1.0f;
        float tol = kRayTolRel * fmaxf(h.dist, 1.0f);
        bool closer = (dd < h.dist - tol) || (fabsf(dd - h.dist) <= tol && (h.geom_id < 0 || gi > h.geom_id));
        if (dd >= 0.0f && closer) {
          h.dist = dd;
          h.normal = n;
          h.geom_id = gi;
          h.bary_u = bu;
          h.bary_v = bv;
          h.face_id = face;
          h.mesh_id = mesh;
        }
        return false;
      };
  // unbounded planes first (their hit, if any, becomes the cut of the BVH traversal)
  for (int k = 0; k < sv.ndirect; ++k) visit(sv.direct[k], std::true_type{});
  traverse<true, kFmaSlabs>(PlainBoxes{sv.tlas_box}, sv.depth, o, rcp, [&]() { return cut_of(h.dist); },
                 [&](int slot) { return visit(sv.tlas_leaf[slot], std::false_type{}); });
  return h;
}

template <bool kMesh = true, bool kFmaSlabs = false>
MJ_DEV bool trace_any(const SceneView& sv, V3 o, V3 d, float max_dist, bool cull) {
  V3 rcp = safe_rcp(d);
  bool hit = false;
  float cut = cut_of(max_dist);
  auto visit = [&](int gi, auto kPlane) {
        if (gi < 0) return false;
        const float* g = sv.geoms + kGeomRec * gi;
        int bits = __float_as_int(g[15]);
        if (bits & 0x100) return false;
        V3 n = v3s(0.0f);
        float bu, bv;
        int face, mesh;
        float dd = intersect_geom<true, kMesh, decltype(kPlane)::value>(sv, g, o, d, max_dist, &n, &bu, &bv, &face, &mesh);
        if (cull && dd >= 0.0f && dot(d, n) > 0.0f) dd = -1.0f;
        if (dd >= 0.0f && dd < max_dist) {
          hit = true;
          return true;
        }
        return false;
      };
  for (int k = 0; k < sv.ndirect; ++k)
    if (visit(sv.direct[k], std::true_type{})) return true;
  traverse<false, kFmaSlabs>(PlainBoxes{sv.tlas_box}, sv.depth, o, rcp, [&]() { return cut; },
                  [&](int slot) { return visit(sv.tlas_leaf[slot], std::false_type{}); });
  return hit;
}

// Any-hit queries of up to N rays from one origin (the soft-shadow rays towards one light), in
// one traversal of the scene BVH. Returns bit k set when ray k (k < n) hits something before
// tmax[k]. A node is entered when the hull of the rays' slab intervals meets its box: per axis,
// (bound - o) * r is monotonic in the reciprocal direction r, so the extreme rays' float products
// bound every ray's, and the hull test accepts every box a ray's own test accepts (an axis on
// which the rays' directions differ in sign or have a zero component constrains nothing). Each
// unresolved ray is tested against each reached leaf's geom exactly as in trace_any, so every
// ray's result is that of trace_any: a geom a ray hits lies in its padded box, which its own slab
// test, and so the hull test, accepts. With kMesh, the mesh geoms reached are tested after the
// walk (up to kMeshPending of them; more are tested where reached), shared out over the wave's
// lanes by run_mesh_jobs.
template <bool kMesh, int N>
MJ_DEV uint32_t trace_any_packet(const SceneView& sv, V3 o, const V3 (&d)[N], const float (&tmax)[N], int n, bool cull) {
  const uint32_t all = n >= 32 ? 0xffffffffu : ((1u << n) - 1u);
  uint32_t hit = 0u;
  // reciprocal ranges per axis; free[a]: the axis does not constrain the packet
  float rlo[3], rhi[3];
  bool free_axis[3];
  float cut = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    rlo[a] = INFINITY;
    rhi[a] = -INFINITY;
    free_axis[a] = false;
  }
  bool pos[3] = {false, false, false}, neg[3] = {false, false, false};
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (k >= n) break;
    const float c[3] = {d[k].x, d[k].y, d[k].z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float r = 1.0f / c[a];
      if (c[a] > 0.0f) pos[a] = true;
      else if (c[a] < 0.0f) neg[a] = true;
      else free_axis[a] = true;
      if (!(fabsf(r) < INFINITY)) free_axis[a] = true;
      rlo[a] = fminf(rlo[a], r);
      rhi[a] = fmaxf(rhi[a], r);
    }
    cut = fmaxf(cut, cut_of(tmax[k]));
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (pos[a] && neg[a]) free_axis[a] = true;
  const float oc[3] = {o.x, o.y, o.z};
  // entry distance of the packet's hull into box [lo, hi], or -1 when it misses before the cut
  auto hull_box = [&](const float* lo, const float* hi) {
    float tmin = 0.0f, tmx = cut;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (free_axis[a]) continue;
      const float t1 = lo[a] - oc[a], t2 = hi[a] - oc[a];
      const float p1 = t1 * rlo[a], p2 = t1 * rhi[a], p3 = t2 * rlo[a], p4 = t2 * rhi[a];
      tmin = fmaxf(tmin, fminf(fminf(p1, p2), fminf(p3, p4)));
      tmx = fminf(tmx, fmaxf(fmaxf(p1, p2), fmaxf(p3, p4)));
    }
    return (tmx >= tmin && tmin < cut) ? tmin : -1.0f;
  };
  uint32_t pend[kMeshPending] = {0u, 0u, 0u, 0u};
  // tests the unresolved rays against geom gi; true when every ray is resolved
  auto visit = [&](int gi, auto kPlane) {
    if (gi < 0) return false;
    const float* g = sv.geoms + kGeomRec * gi;
    const int bits = __float_as_int(g[15]);
    if (bits & 0x100) return false;
    if (kMesh && !decltype(kPlane)::value && (bits & 0xff) == MJPBR_GEOM_MESH) {
      // mesh, SDF and height-field geoms (staged as meshes): deferred to run_mesh_jobs
      if ((bits >> 9) - 1 < 0) return false;
      const uint32_t rec = (uint32_t)gi | ((all & ~hit) << kPendShift);
      bool queued = false;
#pragma unroll
      for (int p = 0; p < kMeshPending; ++p)
        if (!queued && pend[p] == 0u) {
          pend[p] = rec;
          queued = true;
        }
      if (queued) return false;
      // more pending meshes than slots: test this one here, ray by ray
#pragma unroll
      for (int k = 0; k < N; ++k) {
        if (k >= n || (hit >> k) & 1u) continue;
        V3 nrm;
        float bu, bv;
        int face, mesh;
        if (intersect_geom<true, true>(sv, g, o, d[k], tmax[k], &nrm, &bu, &bv, &face, &mesh) >= 0.0f) hit |= 1u << k;
      }
      return hit == all;
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if (k >= n || (hit >> k) & 1u) continue;
      V3 nrm = v3s(0.0f);
      float bu, bv;
      int face, mesh;
      float dd = intersect_geom<true, false, decltype(kPlane)::value>(sv, g, o, d[k], tmax[k], &nrm, &bu, &bv, &face, &mesh);
      if (cull && dd >= 0.0f && dot(d[k], nrm) > 0.0f) dd = -1.0f;
      if (dd >= 0.0f && dd < tmax[k]) hit |= 1u << k;
    }
    return hit == all;
  };
  for (int k = 0; k < sv.ndirect; ++k)
    if (visit(sv.direct[k], std::true_type{})) return hit;
  const int depth = sv.depth;
  const int ninner = (1 << depth) - 1;
  int node = 0, level = 0;
  uint32_t trail = 0u;
  for (;;) {
    if (level < depth) {
      const float* b = sv.tlas_box + 12 * node;
      const float ta = hull_box(b, b + 3), tb = hull_box(b + 6, b + 9);
      if (ta >= 0.0f || tb >= 0.0f) {
        const int nearside = (tb >= 0.0f && (ta < 0.0f || tb < ta)) ? 1 : 0;
        if (ta >= 0.0f && tb >= 0.0f) trail |= 1u << level;
        node = 2 * node + 1 + nearside;
        ++level;
        continue;
      }
    } else if (visit(sv.tlas_leaf[node - ninner], std::false_type{})) {
      return hit;
    }
    if (trail == 0u) break;
    const int L = 31 - __builtin_clz(trail);
    trail &= ~(1u << L);
    const int anc1 = (node + 1) >> (level - L - 1);
    node = (anc1 ^ 1) - 1;
    level = L + 1;
  }
  if (kMesh) hit = run_mesh_jobs<N>(sv, o, d, tmax, pend, hit);
  return hit;
}

}  // namespace mjpbr
