// atrous.hip — edge-avoiding a-trous denoiser of the path-traced tier
// (reference denoise.py:95-248).
//
// demodulate (hdr * exposure - emission) / albedo  ->  `iterations` 5x5 B3-spline passes with
// step 1, 2, 4, ... whose weights fall off with normal, depth, luminance (scaled by the local
// noise estimate) and albedo differences, variance propagated as sum w^2 var / (sum w)^2  ->
// remodulate, tone map, encode. The remodulation is fused into the last pass (one launch and one
// image round trip less than the reference's separate finish kernel). One thread per pixel; the
// per-world images are small enough that the 5x5 neighbourhood reads are served by L2. The
// demodulation packs each pixel's guides (normal and albedo as their f16 bits, depth) into one
// 16-byte record, so a tap reads its guides with one load beside its colour. The edge weight's
// power, exponential and albedo distance use the hardware v_log_f32 / v_exp_f32 / v_sqrt_f32
// (exp2(s log2 x), exp2(-d log2 e), sqrt without the correctly rounded fix-up); they differ from
// correctly rounded powf / expf / sqrtf by a few ulp of the weight, far inside the 1e-4 relative
// agreement the filtered image is held to.
#pragma clang fp contract(off)
#include "abi_util.hpp"

namespace mjpbr {

constexpr float kMinAlbedo = 1.0e-2f;  // denoise.py:40

struct Locate {
  int adr, w, h, pid;
};

MJ_DEV Locate locate(const MjpbrAtrousParams& p, int index) {
  for (int i = 0; i < p.nrender; ++i) {
    int adr = p.rgb_adr[i];
    int n = p.cam_res[i][0] * p.cam_res[i][1];
    if (adr >= 0 && index >= adr && index < adr + n) return Locate{adr, p.cam_res[i][0], p.cam_res[i][1], index - adr};
  }
  return Locate{-1, 0, 0, 0};
}

MJ_DEV V3 ld_half3(const uint16_t* p) { return V3{half_to_float(p[0]), half_to_float(p[1]), half_to_float(p[2])}; }

// Guide record: {n.x | n.y << 16, n.z | a.x << 16, a.y | a.z << 16, depth bits} (f16 bit patterns).
MJ_DEV uint4 pack_guides(const uint16_t* n, const uint16_t* a, float depth) {
  return make_uint4((uint32_t)n[0] | ((uint32_t)n[1] << 16), (uint32_t)n[2] | ((uint32_t)a[0] << 16),
                    (uint32_t)a[1] | ((uint32_t)a[2] << 16), __float_as_uint(depth));
}
MJ_DEV V3 guide_normal(uint4 g) {
  return V3{half_to_float((uint16_t)(g.x & 0xffffu)), half_to_float((uint16_t)(g.x >> 16)), half_to_float((uint16_t)(g.y & 0xffffu))};
}
MJ_DEV V3 guide_albedo(uint4 g) {
  return V3{half_to_float((uint16_t)(g.y >> 16)), half_to_float((uint16_t)(g.z & 0xffffu)), half_to_float((uint16_t)(g.z >> 16))};
}
MJ_DEV float guide_depth(uint4 g) { return __uint_as_float(g.w); }

MJ_DEV float b3(int i) { return i == 2 ? 0.375f : ((i == 1 || i == 3) ? 0.25f : 0.0625f); }

MJ_DEV V3 demodulated(V3 c, V3 albedo) {
  return V3{c.x / fmaxf(albedo.x, kMinAlbedo), c.y / fmaxf(albedo.y, kMinAlbedo), c.z / fmaxf(albedo.z, kMinAlbedo)};
}

// Remodulate, tone map and encode (denoise.py:181-201).
MJ_DEV void finish_pixel(const MjpbrAtrousParams& p, const MjpbrOutputs& out, int w, int index, float4 s) {
  int64_t a = (int64_t)w * out.aov_stride + index;
  V3 albedo = ld_half3(out.aov_albedo + 3 * a);
  V3 emission = ld3(out.aov_emission + 3 * a);
  V3 exposed = emission + V3{s.x * fmaxf(albedo.x, kMinAlbedo), s.y * fmaxf(albedo.y, kMinAlbedo), s.z * fmaxf(albedo.z, kMinAlbedo)};
  float exposure = p.exposure[w % p.exposure_n];
  if (p.render_hdr) st3(out.hdr + 3 * ((int64_t)w * out.hdr_stride + index), exposed / exposure);
  out.rgb[(int64_t)w * out.rgb_stride + index] = encode_display(display_transform(exposed, p.tone_mapper));
}

// denoise.py:95-102
__global__ void demodulate_kernel(MjpbrAtrousParams p, MjpbrOutputs out, float4* dst, uint4* guides, int finish) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)p.nworld * p.stride) return;
  int w = (int)(i / p.stride), index = (int)(i - (int64_t)w * p.stride);
  if (p.world_select && !p.world_select[w]) return;
  // the pixels of this launch's cameras only (a context's cameras may be filtered in several launches)
  if (locate(p, index).adr < 0) return;
  int64_t a = (int64_t)w * out.aov_stride + index;
  if (!finish) guides[i] = pack_guides(out.aov_normal + 3 * a, out.aov_albedo + 3 * a, out.aov_depth[a]);
  V3 albedo = ld_half3(out.aov_albedo + 3 * a);
  V3 c = ld3(out.hdr + 3 * ((int64_t)w * out.hdr_stride + index)) * p.exposure[w % p.exposure_n] - ld3(out.aov_emission + 3 * a);
  c = demodulated(V3{fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f)}, albedo);
  float scale = 1.0f / fmaxf(luminance(albedo), kMinAlbedo);
  float4 v = make_float4(c.x, c.y, c.z, out.aov_variance[a] * scale * scale);
  if (finish)
    finish_pixel(p, out, w, index, v);
  else
    dst[i] = v;
}

// denoise.py:105-178; `finish` fuses denoise.py:181-201 into the last pass.
// Workgroups are dispatched to the 8 XCDs round-robin; workgroup b covers the pixels of block
// (b % 8) * xcd_chunk + b / 8, so each XCD filters a contiguous run of rows and the dilated 5x5
// footprints of neighbouring rows are served by that XCD's L2 instead of being fetched by all 8.
__global__ void atrous_kernel(MjpbrAtrousParams p, MjpbrOutputs out, int step, const float4* src, const uint4* guides,
                              float4* dst, int finish, int xcd_chunk) {
  const int64_t blk = (int64_t)(blockIdx.x & 7) * xcd_chunk + (blockIdx.x >> 3);
  int64_t i = blk * blockDim.x + threadIdx.x;
  if (i >= (int64_t)p.nworld * p.stride) return;
  int w = (int)(i / p.stride), index = (int)(i - (int64_t)w * p.stride);
  if (p.world_select && !p.world_select[w]) return;
  Locate loc = locate(p, index);
  if (loc.adr < 0) return;
  const float4* srcw = src + (int64_t)w * p.stride;
  const uint4* gw = guides + (int64_t)w * p.stride;
  const uint4 g0 = gw[index];
  V3 n0 = guide_normal(g0);
  float4 s0 = srcw[index];
  if (len_sq(n0) == 0.0f) {
    if (finish)
      finish_pixel(p, out, w, index, s0);
    else
      dst[i] = s0;
    return;
  }
  int W = loc.w, H = loc.h;
  int px = loc.pid % W, py = loc.pid / W;
  float variance = 0.0f, vweight = 0.0f;
  for (int j = 0; j < 3; ++j) {
    int y = min(max(py + j - 1, 0), H - 1);
    for (int k = 0; k < 3; ++k) {
      int x = min(max(px + k - 1, 0), W - 1);
      float g = (k == 1 ? 0.5f : 0.25f) * (j == 1 ? 0.5f : 0.25f);
      variance += g * srcw[loc.adr + y * W + x].w;
      vweight += g;
    }
  }
  float inv_sigma_l = 1.0f / (p.sigma_color * sqrtf(fmaxf(variance / vweight, 0.0f)) + 1.0e-6f);
  float l0 = luminance(V3{s0.x, s0.y, s0.z});
  float z0 = guide_depth(g0);
  V3 a0 = guide_albedo(g0);
  float inv_sigma_a = 1.0f / p.sigma_albedo;
  float inv_sigma_z = 1.0f / fmaxf(p.sigma_depth * z0 * (float)step, 1.0e-6f);
  const float wn_zero = powf(0.0f, p.sigma_normal);  // the normal weight of a tap facing away
  V3 total = v3s(0.0f);
  float total_var = 0.0f, wsum = 0.0f;
  for (int j = 0; j < 5; ++j) {
    int y = py + (j - 2) * step;
    if (y < 0 || y >= H) continue;
    for (int k = 0; k < 5; ++k) {
      int x = px + (k - 2) * step;
      if (x < 0 || x >= W) continue;
      int q = loc.adr + y * W + x;
      float4 s = srcw[q];
      const uint4 gq = gw[q];
      V3 c = V3{s.x, s.y, s.z};
      const float nd = fmaxf(dot(n0, guide_normal(gq)), 0.0f);
      float wn = nd > 0.0f ? __builtin_amdgcn_exp2f(p.sigma_normal * __builtin_amdgcn_logf(nd)) : wn_zero;
      float dist = fabsf(guide_depth(gq) - z0) * inv_sigma_z + fabsf(luminance(c) - l0) * inv_sigma_l +
                   __builtin_amdgcn_sqrtf(len_sq(guide_albedo(gq) - a0)) * inv_sigma_a;
      float weight = b3(k) * b3(j) * wn * __expf(-dist);
      total += c * weight;
      total_var += weight * weight * s.w;
      wsum += weight;
    }
  }
  V3 c = total / wsum;
  float4 r = make_float4(c.x, c.y, c.z, total_var / (wsum * wsum));
  if (finish)
    finish_pixel(p, out, w, index, r);
  else
    dst[i] = r;
}

}  // namespace mjpbr

using namespace mjpbr;

extern "C" int mjpbr_atrous(const MjpbrAtrousParams* p, const MjpbrOutputs* out, float* work0, float* work1, float* guides,
                            void* stream) {
  if (p->nworld < 1 || p->stride < 1) return set_error(MJPBR_EINVAL, "atrous: empty image");
  if (p->nrender < 1 || p->nrender > MJPBR_MAX_CAMS) return set_error(MJPBR_EINVAL, "atrous: bad camera count");
  if (p->iterations < 0 || p->iterations > 16) return set_error(MJPBR_EINVAL, "atrous: iterations must be in [0, 16]");
  int64_t n = (int64_t)p->nworld * p->stride;
  int threads = 256;
  unsigned blocks = (unsigned)((n + threads - 1) / threads);
  const unsigned xcd_chunk = (blocks + 7) / 8;  // a-trous blocks each XCD filters (see atrous_kernel)
  hipStream_t st = (hipStream_t)stream;
  float4* a = reinterpret_cast<float4*>(work0);
  float4* b = reinterpret_cast<float4*>(work1);
  if (p->iterations > 0 && !guides) return set_error(MJPBR_EINVAL, "atrous: the guide buffer must not be NULL");
  uint4* g = reinterpret_cast<uint4*>(guides);
  hipLaunchKernelGGL(demodulate_kernel, dim3(blocks), dim3(threads), 0, st, *p, *out, a, g, p->iterations == 0 ? 1 : 0);
  int rc = check_launch("demodulate_kernel");
  if (rc) return rc;
  for (int it = 0; it < p->iterations; ++it) {
    int last = it == p->iterations - 1;
    hipLaunchKernelGGL(atrous_kernel, dim3(xcd_chunk * 8), dim3(threads), 0, st, *p, *out, 1 << it, a, g, b, last, (int)xcd_chunk);
    rc = check_launch("atrous_kernel");
    if (rc) return rc;
    float4* t = a;
    a = b;
    b = t;
  }
  return MJPBR_OK;
}
