// stage.hpp — cooperative load of one world's scene BVH, geom frames and light records into
// LDS, so every closest-hit and shadow query of the workgroup traverses LDS instead of HBM, and
// the shading code reads each light and each hit geom's material constants as LDS vectors.
#pragma once
#include "camera.hpp"

namespace mjpbr {


MJ_DEV SceneView stage_world(const MjpbrScene& sc, int w, float* lds) {
  const MjpbrGeometry& g = sc.geo;
  int nleaf = 1 << g.tlas_depth;
  int ninner = nleaf - 1;
  float* box = lds;
  float* geoms = box + ninner * 12;
  int* leaf = reinterpret_cast<int*>(geoms + g.ngeom * kGeomRec);
  float* lights = reinterpret_cast<float*>(leaf + staged_leaf_slots(nleaf));
  for (int li = threadIdx.x; li < sc.lit.nlight; li += blockDim.x) stage_light(sc.lit, w, li, lights + li * kLightRec);
  float* surfs = lights + sc.lit.nlight * kLightRec;
  for (int gi = threadIdx.x; gi < g.ngeom; gi += blockDim.x) stage_surface(sc, w, gi, surfs + gi * kSurfRec);
  const float4* src = reinterpret_cast<const float4*>(g.tlas_box + (int64_t)w * ninner * 12);
  float4* dst = reinterpret_cast<float4*>(box);
  for (int i = threadIdx.x; i < ninner * 3; i += blockDim.x) dst[i] = src[i];
  const int* lsrc = g.tlas_leaf + (int64_t)w * nleaf;
  for (int i = threadIdx.x; i < nleaf; i += blockDim.x) leaf[i] = lsrc[i];
  for (int gi = threadIdx.x; gi < g.ngeom; gi += blockDim.x) {
    float* r = geoms + gi * kGeomRec;
    const float* xp = g.geom_xpos + ((int64_t)w * g.ngeom + gi) * 3;
    const float* xm = g.geom_xmat + ((int64_t)w * g.ngeom + gi) * 9;
    const float* sz = bat(g.geom_size, g.geom_size_n, w, 3 * g.ngeom) + 3 * gi;
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
    int type = g.geom_type[gi];
    int dataid = bat(g.geom_dataid, g.geom_dataid_n, w, g.ngeom)[gi];
    if (type == MJPBR_GEOM_HFIELD && dataid >= 0) dataid += g.nmesh;
    int traced = (type == MJPBR_GEOM_SDF || type == MJPBR_GEOM_HFIELD) ? MJPBR_GEOM_MESH : type;
    int bits = traced | (rgba.w == 0.0f ? 0x100 : 0) | ((dataid + 1) << 9);
    r[15] = __int_as_float(bits);
  }
  __syncthreads();
  SceneView sv;
  sv.tlas_box = box;
  sv.tlas_leaf = leaf;
  sv.geoms = geoms;
  sv.depth = g.tlas_depth;
  sv.lights = lights;
  sv.surfs = surfs;
  sv.blas_box = g.blas_box;
  sv.blas_tri = g.blas_tri;
  sv.blas_depth = g.blas_depth;
  sv.blas_node_adr = g.blas_node_adr;
  sv.blas_tri_adr = g.blas_tri_adr;
  return sv;
}

}  // namespace mjpbr
