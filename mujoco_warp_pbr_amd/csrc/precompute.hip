// precompute.hip — render-context tables built on the GPU: the split-sum DFG LUT
// (reference ibl.py:40-104), the GGX-prefiltered environment stack (envmap.py:171-221), and
// per-camera image extraction for get_hdr / get_albedo / get_normal (output.py:55-67).
#pragma clang fp contract(off)
#include "abi_util.hpp"

namespace mjpbr {

// Hammersley point i of n (ibl.py:51-58): (i / n, radical inverse of i).
MJ_DEV V2 hammersley(int i, float inv_n) {
  uint32_t bits = __builtin_bitreverse32((uint32_t)i);
  return V2{(float)i * inv_n, (float)bits * (1.0f / 4294967296.0f)};
}

// Half vector around +z distributed as D(h)(h.n) (ibl.py:61-69).
MJ_DEV V3 importance_sample_ggx(V2 u, float alpha) {
  float a2 = alpha * alpha;
  float sp, cp;
  sincos_turns(u.x, &sp, &cp);
  float cos2 = (1.0f - u.y) / (1.0f + (a2 - 1.0f) * u.y);
  float ct = sqrtf(cos2);
  float st = sqrtf(fmaxf(1.0f - cos2, 0.0f));
  return V3{st * cp, st * sp, ct};
}

__global__ void dfg_lut_kernel(float* lut, int n, int nsample) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;  // NoV column
  int j = blockIdx.y;                              // roughness row
  if (i >= n || j >= n) return;
  float nov = ((float)i + 0.5f) / (float)n;
  float pr = ((float)j + 0.5f) / (float)n;
  float alpha = pr * pr;
  V3 v = V3{sqrtf(1.0f - nov * nov), 0.0f, nov};
  float inv_n = 1.0f / (float)nsample;
  float x = 0.0f, y = 0.0f;
  for (int s = 0; s < nsample; ++s) {
    V3 h = importance_sample_ggx(hammersley(s, inv_n), alpha);
    float voh = dot(v, h);
    V3 l = h * (2.0f * voh) - v;
    float nol = l.z, noh = h.z;
    if (nol > 0.0f && voh > 0.0f) {
      float gv = v_smith_ggx_correlated(nov, nol, alpha) * 4.0f * voh * nol / noh;
      float t = 1.0f - voh;
      float fc = t * t * t * t * t;
      x += gv * fc;
      y += gv;
    }
  }
  lut[2 * (j * n + i)] = x * inv_n;
  lut[2 * (j * n + i) + 1] = y * inv_n;
}

__global__ void prefilter_kernel(const uint16_t* radiance, const int64_t* level_adr, const int32_t* desc,
                                 uint16_t* stack, int L, int H, int W, int nsample) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  int level = blockIdx.z;
  if (x >= W) return;
  V3 n = uv_to_direction(V2{((float)x + 0.5f) / (float)W, ((float)y + 0.5f) / (float)H});
  float pr = (float)level / (float)(L - 1);
  float alpha = fmaxf(pr * pr, 1.0e-4f);
  V3 up = fabsf(n.z) < 0.999f ? V3{0.0f, 0.0f, 1.0f} : V3{1.0f, 0.0f, 0.0f};
  V3 t = normalize(cross(up, n));
  V3 b = cross(n, t);
  float texel_sa = 4.0f * kPi / (float)(desc[0] * desc[1]);
  V3 color = v3s(0.0f);
  float weight = 0.0f;
  float inv_n = 1.0f / (float)nsample;
  for (int s = 0; s < nsample; ++s) {
    V3 hl = importance_sample_ggx(hammersley(s, inv_n), alpha);
    V3 h = t * hl.x + b * hl.y + n * hl.z;
    float noh = hl.z;
    V3 l = h * (2.0f * noh) - n;
    float nol = dot(n, l);
    if (nol > 0.0f) {
      float pdf = d_ggx(noh, alpha) * 0.25f;
      float sample_sa = inv_n / fmaxf(pdf, 1.0e-8f);
      float lod = level == 0 ? 0.0f : 0.5f * log2f(sample_sa / texel_sa) + 1.0f;
      float4 c = sample_mip(reinterpret_cast<const uint8_t*>(radiance), level_adr, desc, 2, direction_to_uv(l, 0.0f),
                            fmaxf(lod, 0.0f), false);
      color += V3{c.x, c.y, c.z} * nol;
      weight += nol;
    }
  }
  color = color / fmaxf(weight, 1.0e-8f);
  ushort4 o = make_ushort4(float_to_half(color.x), float_to_half(color.y), float_to_half(color.z), float_to_half(1.0f));
  reinterpret_cast<ushort4*>(stack)[((int64_t)level * H + y) * W + x] = o;
}

__global__ void extract_kernel(const float* src, int stride, int adr, int w, int h, int nout, float* dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)nout * w * h;
  if (i >= total) return;
  int64_t world = i / ((int64_t)w * h);
  int64_t p = i - world * w * h;
  const float* s = src + 3 * (world * stride + adr + p);
  dst[3 * i] = s[0];
  dst[3 * i + 1] = s[1];
  dst[3 * i + 2] = s[2];
}

}  // namespace mjpbr

using namespace mjpbr;

extern "C" int mjpbr_dfg_lut(float* lut, int size, int nsample, void* stream) {
  if (size < 1 || nsample < 1) return set_error(MJPBR_EINVAL, "dfg_lut: size and nsample must be positive");
  dim3 grid((size + 63) / 64, size);
  hipLaunchKernelGGL(dfg_lut_kernel, grid, dim3(64), 0, (hipStream_t)stream, lut, size, nsample);
  return check_launch("dfg_lut_kernel");
}

extern "C" int mjpbr_env_prefilter(const uint16_t* radiance, const int64_t* level_adr, const int32_t* desc,
                                   uint16_t* stack, int levels, int height, int width, int nsample, void* stream) {
  if (levels < 2 || height < 1 || width < 1 || nsample < 1)
    return set_error(MJPBR_EINVAL, "env_prefilter: bad stack shape %dx%dx%d", levels, height, width);
  dim3 grid((width + 63) / 64, height, levels);
  hipLaunchKernelGGL(prefilter_kernel, grid, dim3(64), 0, (hipStream_t)stream, radiance, level_adr, desc, stack, levels,
                     height, width, nsample);
  return check_launch("prefilter_kernel");
}

extern "C" int mjpbr_extract_rgb_image(const float* src, int src_stride, int adr, int width, int height, int nout,
                                       float* dst, void* stream) {
  int64_t total = (int64_t)nout * width * height;
  if (total <= 0) return MJPBR_OK;
  int threads = 256;
  int64_t blocks = (total + threads - 1) / threads;
  hipLaunchKernelGGL(extract_kernel, dim3((unsigned)blocks), dim3(threads), 0, (hipStream_t)stream, src, src_stride,
                     adr, width, height, nout, dst);
  return check_launch("extract_kernel");
}
