// stage.hpp — cooperative load of one world's scene BVH, geom frames and light records into
// LDS, so every closest-hit and shadow query of the workgroup traverses LDS instead of HBM, and
// the shading code reads each light and each hit geom's material constants as LDS vectors.
#pragma once
#include "camera.hpp"

namespace mjpbr {


// out = M R^T (row-major 3 x 3; sums in index order): the rotation a mesh geom is traced and
// shaded with when its mesh BVH is in frame R (MjpbrGeometry.mesh_frame).
MJ_HOST_DEV inline void framed_rotation(const float* M, const float* R, float* out) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] = M[3 * i] * R[3 * j] + M[3 * i + 1] * R[3 * j + 1] + M[3 * i + 2] * R[3 * j + 2];
}

// The rotation geom gi of world w is traced with: geom_xmat, or geom_xmat R^T for a mesh or SDF
// geom whose mesh is re-framed. Returns `xm` or `tmp`.
MJ_HOST_DEV inline const float* traced_rotation(const MjpbrGeometry& g, int type, int dataid, const float* xm, float* tmp) {
  if (!g.mesh_frame || dataid < 0 || (type != MJPBR_GEOM_MESH && type != MJPBR_GEOM_SDF)) return xm;
  framed_rotation(xm, g.mesh_frame + 9 * dataid, tmp);
  return tmp;
}

// Floats of one world's staged records: scene BVH boxes, geom records, leaf ids and direct geom
// ids (each padded to a multiple of 4), light records, surface records, the environment record
// (kEnvRec). A multiple of 4, so records copy as float4.
MJ_HOST_DEV inline int64_t stage_floats(const MjpbrScene& sc) {
  int nleaf = 1 << sc.geo.tlas_depth;
  return (int64_t)(nleaf - 1) * 12 + (int64_t)sc.geo.ngeom * kGeomRec + staged_leaf_slots(nleaf) +
         staged_leaf_slots(sc.geo.ndirect) + (int64_t)sc.lit.nlight * kLightRec + (int64_t)sc.geo.ngeom * kSurfRec +
         kEnvRec;
}

// Resolves world w's records into dst (LDS or global memory) with the block's threads: the
// per-world material, light and geom arrays (indexed w % n) are read here and nowhere else.
MJ_DEV void stage_world_to(const MjpbrScene& sc, int w, float* dst) {
  const MjpbrGeometry& g = sc.geo;
  int nleaf = 1 << g.tlas_depth;
  int ninner = nleaf - 1;
  float* box = dst;
  float* geoms = box + ninner * 12;
  int* leaf = reinterpret_cast<int*>(geoms + g.ngeom * kGeomRec);
  int* direct = leaf + staged_leaf_slots(nleaf);
  for (int i = threadIdx.x; i < staged_leaf_slots(g.ndirect); i += blockDim.x) direct[i] = i < g.ndirect ? g.direct_geom[i] : -1;
  float* lights = reinterpret_cast<float*>(direct + staged_leaf_slots(g.ndirect));
  for (int li = threadIdx.x; li < sc.lit.nlight; li += blockDim.x) stage_light(sc.lit, w, li, lights + li * kLightRec);
  float* surfs = lights + sc.lit.nlight * kLightRec;
  for (int gi = threadIdx.x; gi < g.ngeom; gi += blockDim.x) stage_surface(sc, w, gi, surfs + gi * kSurfRec);
  // environment record: cosine and sine of the negated sky rotation (the irradiance lookup
  // rotates the normal by -yaw), evaluated once per world
  if (threadIdx.x == 0) {
    float* env = surfs + g.ngeom * kSurfRec;
    const float yaw = sc.env.env_yaw[w % sc.env.env_yaw_n];
    env[0] = cosf(-yaw);
    env[1] = sinf(-yaw);
    env[2] = env[3] = 0.0f;
  }
  const float4* src = reinterpret_cast<const float4*>(g.tlas_box + (int64_t)w * ninner * 12);
  float4* bdst = reinterpret_cast<float4*>(box);
  for (int i = threadIdx.x; i < ninner * 3; i += blockDim.x) bdst[i] = src[i];
  const int* lsrc = g.tlas_leaf + (int64_t)w * nleaf;
  for (int i = threadIdx.x; i < staged_leaf_slots(nleaf); i += blockDim.x) leaf[i] = i < nleaf ? lsrc[i] : -1;
  for (int gi = threadIdx.x; gi < g.ngeom; gi += blockDim.x) {
    float* r = geoms + gi * kGeomRec;
    const float* xp = g.geom_xpos + ((int64_t)w * g.ngeom + gi) * 3;
    const float* sz = bat(g.geom_size, g.geom_size_n, w, 3 * g.ngeom) + 3 * gi;
    int type = g.geom_type[gi];
    int dataid = bat(g.geom_dataid, g.geom_dataid_n, w, g.ngeom)[gi];
    float framed[9];
    const float* xm = traced_rotation(g, type, dataid, g.geom_xmat + ((int64_t)w * g.ngeom + gi) * 9, framed);
    r[0] = xp[0];
    r[1] = xp[1];
    r[2] = xp[2];
    r[3] = sz[0];
#pragma unroll
    for (int k = 0; k < 9; ++k) r[4 + k] = xm[k];
    r[13] = sz[1];
    r[14] = sz[2];
    int mid;
    float4 rgba = geom_rgba(g, w, gi, &mid);
    if (type == MJPBR_GEOM_HFIELD && dataid >= 0) dataid += g.nmesh;
    int traced = (type == MJPBR_GEOM_SDF || type == MJPBR_GEOM_HFIELD) ? MJPBR_GEOM_MESH : type;
    int bits = traced | (rgba.w == 0.0f ? 0x100 : 0) | ((dataid + 1) << 9);
    r[15] = __int_as_float(bits);
  }
}

MJ_DEV SceneView scene_view(const MjpbrScene& sc, const float* lds) {
  const MjpbrGeometry& g = sc.geo;
  int nleaf = 1 << g.tlas_depth;
  SceneView sv;
  sv.tlas_box = lds;
  sv.geoms = lds + (nleaf - 1) * 12;
  sv.tlas_leaf = reinterpret_cast<const int*>(sv.geoms + g.ngeom * kGeomRec);
  sv.direct = sv.tlas_leaf + staged_leaf_slots(nleaf);
  sv.ndirect = g.ndirect;
  sv.lights = reinterpret_cast<const float*>(sv.direct + staged_leaf_slots(g.ndirect));
  sv.surfs = sv.lights + sc.lit.nlight * kLightRec;
  sv.env = sv.surfs + g.ngeom * kSurfRec;
  sv.depth = g.tlas_depth;
  sv.mesh_cache = nullptr;
  sv.mesh_cache_nodes = 0;
  sv.mesh_cache_n = 0;
  sv.blas_box = g.blas_box;
  sv.blas_tri = g.blas_tri;
  sv.blas_depth = g.blas_depth;
  sv.blas_node_adr = g.blas_node_adr;
  sv.blas_tri_adr = g.blas_tri_adr;
  return sv;
}

// Copies the first min(nodes, 2^depth - 1) inner nodes of meshes 0 .. n - 1 into dst (mesh m at
// dst + 12 * nodes * m), with the block's threads.
MJ_DEV void stage_mesh_cache(const MjpbrGeometry& g, int nodes, int n, float* dst) {
  const int per = 3 * nodes;  // float4 a mesh
  for (int i = threadIdx.x; i < per * n; i += blockDim.x) {
    const int m = i / per, k = i - m * per;
    const int inner = (1 << g.blas_depth[m]) - 1;
    if (k < 3 * inner)
      reinterpret_cast<float4*>(dst)[i] = reinterpret_cast<const float4*>(g.blas_box + 12 * (int64_t)g.blas_node_adr[m])[k];
  }
}

// World w's records into LDS: a float4 copy of the prepared records when geo.world_stage is set
// (stage_prep_kernel ran first on the stream), else resolved here. With cache_n > 0 the top
// levels (cache_nodes inner nodes) of the first cache_n mesh BVHs follow the records.
MJ_DEV SceneView stage_world(const MjpbrScene& sc, int w, float* lds, int cache_nodes = 0, int cache_n = 0) {
  float* cache = lds + stage_floats(sc);
  if (cache_n > 0) stage_mesh_cache(sc.geo, cache_nodes, cache_n, cache);
  if (sc.geo.world_stage) {
    const int64_t n4 = stage_floats(sc) / 4;
    const float4* src = reinterpret_cast<const float4*>(sc.geo.world_stage) + (int64_t)w * n4;
    float4* dst = reinterpret_cast<float4*>(lds);
    for (int64_t i = threadIdx.x; i < n4; i += blockDim.x) dst[i] = src[i];
  } else {
    stage_world_to(sc, w, lds);
  }
  __syncthreads();
  SceneView sv = scene_view(sc, lds);
  if (cache_n > 0) {
    sv.mesh_cache = cache;
    sv.mesh_cache_nodes = cache_nodes;
    sv.mesh_cache_n = cache_n;
  }
  return sv;
}

// World w's records read in place from geo.world_stage (stage_prep_kernel ran first on the
// stream), for worlds whose records do not fit the LDS; the mesh-BVH top levels (cache_n > 0)
// still go to LDS.
MJ_DEV SceneView global_world(const MjpbrScene& sc, int w, float* lds, int cache_nodes = 0, int cache_n = 0) {
  SceneView sv = scene_view(sc, sc.geo.world_stage + (int64_t)w * stage_floats(sc));
  if (cache_n > 0) {
    stage_mesh_cache(sc.geo, cache_nodes, cache_n, lds);
    __syncthreads();
    sv.mesh_cache = lds;
    sv.mesh_cache_nodes = cache_nodes;
    sv.mesh_cache_n = cache_n;
  }
  return sv;
}

// Mesh-BVH top levels an RT / PT workgroup caches in LDS: the largest 2^L - 1 inner nodes a mesh
// (L <= 8) whose copies for min(nmesh, kMaxCachedMeshes) meshes fit in `budget_bytes`.
constexpr int kMaxCachedMeshes = 16;
MJ_HOST_DEV inline void mesh_cache_shape(const MjpbrScene& sc, int budget_bytes, int* nodes, int* n) {
  *nodes = 0;
  *n = sc.geo.nmesh < kMaxCachedMeshes ? sc.geo.nmesh : kMaxCachedMeshes;
  if (*n <= 0) {
    *n = 0;
    return;
  }
  for (int L = 8; L >= 2; --L) {
    const int k = (1 << L) - 1;
    if ((int64_t)k * 48 * *n <= budget_bytes) {
      *nodes = k;
      return;
    }
  }
  *n = 0;
}

}  // namespace mjpbr
