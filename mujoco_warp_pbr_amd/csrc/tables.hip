// tables.hip — render-context tables built on the GPU from uploaded images: texture and radiance
// mip chains (reference texture.py:63-101, envmap.py:366-399), the SH9 irradiance of an
// environment map (envmap.py:224-249) and its importance-sampling tables (envmap.py:402-430).
//
// Every table is a pure function of its image and comes out the same on every launch: reductions
// run in double precision in a fixed order (no atomics), and mip levels cascade in double from
// the unrounded level above, each level rounded once into its storage format.
#pragma clang fp contract(off)
#include <cstring>
#include "abi_util.hpp"

namespace mjpbr {

// IEEE binary16 of a double, round to nearest even (no intermediate float rounding).
MJ_HOST_DEV uint16_t double_to_half(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  const uint16_t sign = (uint16_t)((b >> 48) & 0x8000u);
  const uint64_t a = b & 0x7fffffffffffffffull;
  if (a >= 0x7ff0000000000000ull) return sign | (a > 0x7ff0000000000000ull ? 0x7e00u : 0x7c00u);
  const int e = (int)(a >> 52) - 1023;
  const uint64_t m = (a & 0xfffffffffffffull) | (1ull << 52);  // 53-bit significand
  if (a == 0) return sign;
  if (e >= -14) {  // normal range of binary16 (before rounding)
    uint64_t q = m >> 42, r = m & ((1ull << 42) - 1), halfway = 1ull << 41;
    if (r > halfway || (r == halfway && (q & 1))) ++q;
    int he = e + 15;
    if (q == 2048) {
      q = 1024;
      ++he;
    }
    if (he >= 31) return sign | 0x7c00u;
    return sign | (uint16_t)(he << 10) | (uint16_t)(q & 0x3ffu);
  }
  // subnormal: units of 2^-24
  const int shift = 28 - e;  // m * 2^(e - 52) / 2^-24 = m >> (28 - e)
  if (shift > 54) return sign;
  uint64_t q = m >> shift, r = m & ((1ull << shift) - 1), halfway = 1ull << (shift - 1);
  if (r > halfway || (r == halfway && (q & 1))) ++q;
  return sign | (uint16_t)q;  // q == 1024 encodes the smallest normal
}

// ---------------------------------------------------------------- mip chains
// Level 0 of a texture: RGBA8 texels stored as RGBA8, or decoded through the sRGB table (alpha
// / 255) into RGBA16F; `cur` receives the stored values in double for the cascade.
__global__ void texture_level0_kernel(const uint8_t* rgba8, int64_t n, int format, const uint16_t* srgb_table, uint8_t* dst,
                                      double* cur) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint8_t* t = rgba8 + 4 * i;
  if (format == MJPBR_TEX_RGBA16F) {
    uint16_t h[4] = {srgb_table[t[0]], srgb_table[t[1]], srgb_table[t[2]], double_to_half((double)t[3] / 255.0)};
    for (int k = 0; k < 4; ++k) {
      reinterpret_cast<uint16_t*>(dst)[4 * i + k] = h[k];
      cur[4 * i + k] = (double)half_to_float(h[k]);
    }
  } else {
    for (int k = 0; k < 4; ++k) {
      dst[4 * i + k] = t[k];
      cur[4 * i + k] = (double)t[k];
    }
  }
}

// Level 0 of a radiance map: RGB float32 -> RGBA16F with alpha 1.
__global__ void radiance_level0_kernel(const float* rgb, int64_t n, uint16_t* dst, double* cur) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < 4; ++k) {
    uint16_t h = k < 3 ? double_to_half((double)rgb[3 * i + k]) : double_to_half(1.0);
    dst[4 * i + k] = h;
    cur[4 * i + k] = (double)half_to_float(h);
  }
}

// One 2x2 box-filter step from the unrounded level above (w x h, RGBA doubles): columns are
// averaged first, then rows; an axis of one texel is kept as is. Writes the unrounded level
// (`next`) and its rounded copy in the level's format.
__global__ void mip_step_kernel(const double* cur, int w, int h, double* next, int w2, int h2, int format, uint8_t* dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)w2 * h2) return;
  const int x = (int)(i % w2), y = (int)(i / w2);
  const int xs = w > 1 ? 2 * x : x, ys = h > 1 ? 2 * y : y;
  for (int k = 0; k < 4; ++k) {
    auto col = [&](int row) {
      const double a = cur[4 * ((int64_t)row * w + xs) + k];
      return w > 1 ? 0.5 * (a + cur[4 * ((int64_t)row * w + xs + 1) + k]) : a;
    };
    const double v = h > 1 ? 0.5 * (col(ys) + col(ys + 1)) : col(ys);
    next[4 * i + k] = v;
    if (format == MJPBR_TEX_RGBA16F) {
      reinterpret_cast<uint16_t*>(dst)[4 * i + k] = double_to_half(v);
    } else {
      const double q = floor(v + 0.5);
      dst[4 * i + k] = (uint8_t)(q < 0.0 ? 0.0 : (q > 255.0 ? 255.0 : q));
    }
  }
}

int launch_mip_steps(uint8_t* pool, const int64_t* level_byte_adr, int nlevels, int width, int height, int format,
                     double* work, hipStream_t st) {
  double* cur = work;
  double* next = work + (int64_t)width * height * 4;
  int w = width, h = height;
  for (int l = 1; l < nlevels; ++l) {
    const int w2 = w > 1 ? w / 2 : 1, h2 = h > 1 ? h / 2 : 1;
    const int64_t n = (int64_t)w2 * h2;
    hipLaunchKernelGGL(mip_step_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cur, w, h, next, w2, h2, format,
                       pool + level_byte_adr[l]);
    if (int rc = check_launch("mip_step_kernel")) return rc;
    double* t = cur;
    cur = next;
    next = t;
    w = w2;
    h = h2;
  }
  return MJPBR_OK;
}

int check_chain(int width, int height, int nlevels, const int64_t* level_byte_adr, int64_t work_doubles) {
  if (width < 1 || height < 1) return set_error(MJPBR_EINVAL, "mip chain: bad size %dx%d", width, height);
  if (nlevels < 1 || nlevels > MJPBR_MAX_LEVELS) return set_error(MJPBR_EINVAL, "mip chain: nlevels must be in [1, %d]", MJPBR_MAX_LEVELS);
  if (!level_byte_adr) return set_error(MJPBR_EINVAL, "mip chain: level addresses must not be NULL");
  if (work_doubles < 8ll * width * height) return set_error(MJPBR_EINVAL, "mip chain: work needs %lld doubles", 8ll * width * height);
  return MJPBR_OK;
}

// ---------------------------------------------------------------- SH9 irradiance
constexpr int kShRowThreads = 256;

// Row y's contribution to the 9 x 3 projection of the radiance onto the real SH basis, weighted
// by each texel's solid angle (2 pi / w)(pi / h) sin(theta).
__global__ void __launch_bounds__(kShRowThreads) sh9_rows_kernel(const float* rgb, int w, int h, double* partial) {
  __shared__ double red[kShRowThreads];
  const int y = blockIdx.x;
  const double kPiD = 3.14159265358979323846;
  const double theta = ((double)y + 0.5) / h * kPiD;
  const double st = sin(theta), ct = cos(theta);
  const double sa = (2.0 * kPiD / w) * (kPiD / h) * st;
  double acc[27];
  for (int k = 0; k < 27; ++k) acc[k] = 0.0;
  for (int x = threadIdx.x; x < w; x += kShRowThreads) {
    const double phi = (((double)x + 0.5) / w - 0.5) * 2.0 * kPiD;
    const double dx = st * cos(phi), dy = st * sin(phi), dz = ct;
    const double basis[9] = {0.282095,
                             0.488603 * dy,
                             0.488603 * dz,
                             0.488603 * dx,
                             1.092548 * dx * dy,
                             1.092548 * dy * dz,
                             0.315392 * (3.0 * dz * dz - 1.0),
                             1.092548 * dx * dz,
                             0.546274 * (dx * dx - dy * dy)};
    const float* p = rgb + 3 * ((int64_t)y * w + x);
    for (int b = 0; b < 9; ++b)
      for (int c = 0; c < 3; ++c) acc[3 * b + c] += basis[b] * sa * (double)p[c];
  }
  for (int k = 0; k < 27; ++k) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = kShRowThreads / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)y * 27 + k] = red[0];
    __syncthreads();
  }
}

// Sum of the rows in order, times the Lambertian convolution of each band over pi (the shader
// evaluates irradiance / pi, i.e. the cosine-weighted mean radiance): pi, 2 pi / 3, pi / 4.
__global__ void sh9_finish_kernel(const double* partial, int h, float* out) {
  const int k = threadIdx.x;
  if (k >= 27) return;
  double s = 0.0;
  for (int y = 0; y < h; ++y) s += partial[(int64_t)y * 27 + k];
  const int band = k / 3 == 0 ? 0 : (k / 3 < 4 ? 1 : 2);
  const double conv = band == 0 ? 1.0 : (band == 1 ? 2.0 / 3.0 : 0.25);
  out[k] = (float)(s * conv);
}

// ---------------------------------------------------------------- importance-sampling tables
// One workgroup: the environment box-filtered to at most cdf_w texels wide (block means of the
// cropped image), weights lum * sin(theta) + 1e-6 * mean lum, the normalised probabilities, each
// row's conditional CDF (h x (w + 1)) and the marginal CDF over rows (h + 1), in double.
constexpr int kCdfThreads = 256;

__global__ void __launch_bounds__(kCdfThreads) cdf_kernel(const float* rgb, int W, int H, int halvings, double* cell, float* prob,
                                                          float* marginal, float* conditional) {
  __shared__ double red[kCdfThreads];
  const int bs = 1 << halvings;
  const int w = W >> halvings, h = H >> halvings;
  const double kPiD = 3.14159265358979323846;
  const int64_t n = (int64_t)w * h;
  double lsum = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kCdfThreads) {
    const int x = (int)(i % w), y = (int)(i / w);
    double c[3] = {0.0, 0.0, 0.0};
    for (int yy = 0; yy < bs; ++yy)
      for (int xx = 0; xx < bs; ++xx) {
        const float* p = rgb + 3 * ((int64_t)(y * bs + yy) * W + (x * bs + xx));
        for (int k = 0; k < 3; ++k) c[k] += (double)p[k];
      }
    const double inv = 1.0 / ((double)bs * bs);
    const double lum = 0.2126 * (c[0] * inv) + 0.7152 * (c[1] * inv) + 0.0722 * (c[2] * inv);
    cell[i] = lum;
    lsum += lum;
  }
  red[threadIdx.x] = lsum;
  __syncthreads();
  for (int s = kCdfThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double mean = red[0] / (double)n;
  const double floor_w = 1.0e-6 * (mean > 1.0e-12 ? mean : 1.0e-12);
  __syncthreads();
  double wsum = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kCdfThreads) {
    const int y = (int)(i / w);
    const double v = cell[i] * sin(((double)y + 0.5) / h * kPiD) + floor_w;
    cell[i] = v;
    wsum += v;
  }
  red[threadIdx.x] = wsum;
  __syncthreads();
  for (int s = kCdfThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double total = red[0];
  __syncthreads();
  // rows: probabilities, row mass, conditional CDF (sequential prefix in each row)
  for (int y = threadIdx.x; y < h; y += kCdfThreads) {
    double row = 0.0;
    for (int x = 0; x < w; ++x) {
      const double p = cell[(int64_t)y * w + x] / total;
      cell[(int64_t)y * w + x] = p;
      prob[(int64_t)y * w + x] = (float)p;
      row += p;
    }
    float* cond = conditional + (int64_t)y * (w + 1);
    cond[0] = 0.0f;
    double run = 0.0;
    const double inv_row = 1.0 / (row > 1.0e-30 ? row : 1.0e-30);
    for (int x = 0; x < w; ++x) {
      run += cell[(int64_t)y * w + x];
      cond[x + 1] = (float)(run * inv_row);
    }
    cond[w] = 1.0f;
    cell[n + y] = row;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = 0.0;
    marginal[0] = 0.0f;
    for (int y = 0; y < h; ++y) {
      run += cell[n + y];
      marginal[y + 1] = (float)run;
    }
    marginal[h] = 1.0f;
  }
}

}  // namespace mjpbr

using namespace mjpbr;

extern "C" int mjpbr_texture_mips(const uint8_t* rgba8, int width, int height, int format, const uint16_t* srgb_table,
                                  uint8_t* pool, const int64_t* level_byte_adr, int nlevels, double* work, int64_t work_doubles,
                                  void* stream) {
  if (int rc = check_chain(width, height, nlevels, level_byte_adr, work_doubles)) return rc;
  if (format != MJPBR_TEX_RGBA8 && format != MJPBR_TEX_RGBA16F) return set_error(MJPBR_EINVAL, "texture_mips: bad format %d", format);
  if (format == MJPBR_TEX_RGBA16F && !srgb_table) return set_error(MJPBR_EINVAL, "texture_mips: RGBA16F needs the sRGB table");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = (int64_t)width * height;
  hipLaunchKernelGGL(texture_level0_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rgba8, n, format, srgb_table,
                     pool + level_byte_adr[0], work);
  if (int rc = check_launch("texture_level0_kernel")) return rc;
  return launch_mip_steps(pool, level_byte_adr, nlevels, width, height, format, work, st);
}

extern "C" int mjpbr_radiance_mips(const float* rgb, int width, int height, uint16_t* pool, const int64_t* level_half_adr, int nlevels,
                                   double* work, int64_t work_doubles, void* stream) {
  if (int rc = check_chain(width, height, nlevels, level_half_adr, work_doubles)) return rc;
  int64_t bytes[MJPBR_MAX_LEVELS];
  for (int l = 0; l < nlevels; ++l) bytes[l] = 2 * level_half_adr[l];
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = (int64_t)width * height;
  hipLaunchKernelGGL(radiance_level0_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rgb, n, pool + level_half_adr[0],
                     work);
  if (int rc = check_launch("radiance_level0_kernel")) return rc;
  return launch_mip_steps(reinterpret_cast<uint8_t*>(pool), bytes, nlevels, width, height, MJPBR_TEX_RGBA16F, work, st);
}

extern "C" int mjpbr_env_sh9(const float* rgb, int width, int height, double* work, int64_t work_doubles, float* sh, void* stream) {
  if (width < 1 || height < 1) return set_error(MJPBR_EINVAL, "env_sh9: bad size %dx%d", width, height);
  if (work_doubles < 27ll * height) return set_error(MJPBR_EINVAL, "env_sh9: work needs %lld doubles", 27ll * height);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(sh9_rows_kernel, dim3(height), dim3(kShRowThreads), 0, st, rgb, width, height, work);
  if (int rc = check_launch("sh9_rows_kernel")) return rc;
  hipLaunchKernelGGL(sh9_finish_kernel, dim3(1), dim3(64), 0, st, work, height, sh);
  return check_launch("sh9_finish_kernel");
}

extern "C" int mjpbr_env_sampling_tables(const float* rgb, int width, int height, int cdf_width, double* work, int64_t work_doubles,
                                         float* prob, float* marginal, float* conditional, void* stream) {
  if (width < 1 || height < 1 || cdf_width < 1) return set_error(MJPBR_EINVAL, "env_sampling_tables: bad size %dx%d", width, height);
  int halvings = 0;
  while ((width >> halvings) > cdf_width) ++halvings;
  const int w = width >> halvings, h = height >> halvings;
  if (w < 1 || h < 1) return set_error(MJPBR_EINVAL, "env_sampling_tables: %dx%d has no texel left at width %d", width, height, cdf_width);
  if (work_doubles < (int64_t)w * h + h) return set_error(MJPBR_EINVAL, "env_sampling_tables: work needs %lld doubles", (long long)w * h + h);
  hipLaunchKernelGGL(cdf_kernel, dim3(1), dim3(kCdfThreads), 0, (hipStream_t)stream, rgb, width, height, halvings, work, prob, marginal,
                     conditional);
  return check_launch("cdf_kernel");
}
