// dmath.hpp — scalar/vector math of the gfx950 render kernels.
//
// Geometry and sampling code is compiled without FMA contraction (`#pragma clang fp
// contract(off)` in every render translation unit) so that ray/primitive arithmetic is the
// exact IEEE sequence written here; with correctly rounded division and square root (HIP's
// default) a hit distance is then a deterministic function of its inputs.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#define MJ_DEV __device__ __forceinline__
#define MJ_HOST_DEV __host__ __device__

namespace mjpbr {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kMaxVal = 1.0e10f;     // MuJoCo mjMAXVAL: closest-hit distance of a miss
constexpr float kMinVal = 1.0e-15f;    // MuJoCo mjMINVAL
constexpr float kRayOffset = 1.0e-4f;  // trace.py:47
constexpr float kShadowGap = 1.0e-3f;  // trace.py:49
constexpr float kUnboundedShadow = 1.0e8f;  // trace.py:51
constexpr float kRayTolRel = 1.0e-6f;  // tie-break tolerance of coincident hits
constexpr float kRayTolAbs = 1.0e-7f;

struct V3 {
  float x, y, z;
};
struct V2 {
  float x, y;
};

MJ_DEV V3 v3(float x, float y, float z) { return V3{x, y, z}; }
MJ_DEV V3 v3s(float s) { return V3{s, s, s}; }
MJ_DEV V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
MJ_DEV V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
MJ_DEV V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
MJ_DEV V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
MJ_DEV V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
MJ_DEV V3 cw(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
MJ_DEV V3& operator+=(V3& a, V3 b) {
  a = a + b;
  return a;
}
MJ_DEV float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MJ_DEV V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
MJ_DEV float len_sq(V3 a) { return dot(a, a); }
MJ_DEV float len(V3 a) { return sqrtf(dot(a, a)); }
MJ_DEV V3 normalize(V3 a) {
  float l = sqrtf(dot(a, a));
  return l > 0.0f ? a / l : a;
}
MJ_DEV V3 vmax(V3 a, float s) { return V3{fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s)}; }
MJ_DEV V3 vmin(V3 a, float s) { return V3{fminf(a.x, s), fminf(a.y, s), fminf(a.z, s)}; }
MJ_DEV float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
// log2 and x^y (x > 0) on the hardware v_log_f32 / v_exp_f32: a few ulp from the correctly
// rounded log2f / powf, whose full sequences cost some 40 instructions each. For shading and
// display values only (texture and sky LODs, specular occlusion, the sRGB curve), never for hits.
MJ_DEV float log2_hw(float x) { return __builtin_amdgcn_logf(x); }
MJ_DEV float pow_hw(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }

// Row-major 3x3: m[3 i + j].
MJ_DEV V3 mat_mul(const float* m, V3 v) {
  return V3{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
            m[6] * v.x + m[7] * v.y + m[8] * v.z};
}
MJ_DEV V3 mat_tmul(const float* m, V3 v) {
  return V3{m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
            m[2] * v.x + m[5] * v.y + m[8] * v.z};
}
MJ_DEV V3 ld3(const float* p) { return V3{p[0], p[1], p[2]}; }
MJ_DEV void st3(float* p, V3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}

MJ_DEV float half_to_float(uint16_t h) { return __half2float(__ushort_as_half(h)); }
MJ_DEV uint16_t float_to_half(float f) { return __half_as_ushort(__float2half_rn(f)); }

// sin and cos of 2*pi*t. The argument is reduced exactly (t*4 and its nearest integer are
// exact), then Taylor polynomials of degree 9/10 on [-pi/4, pi/4] (error < 4e-9).
MJ_DEV void sincos_turns(float t, float* s, float* c) {
  float q4 = t * 4.0f;
  float q = floorf(q4 + 0.5f);
  float f = q4 - q;
  float x = f * 1.57079632679489661923f;
  float x2 = x * x;
  float sp = x * (1.0f + x2 * (-1.6666666666666666e-1f +
                               x2 * (8.3333333333333333e-3f +
                                     x2 * (-1.9841269841269841e-4f + x2 * 2.7557319223985891e-6f))));
  float cp = 1.0f + x2 * (-0.5f + x2 * (4.1666666666666667e-2f +
                                        x2 * (-1.3888888888888889e-3f +
                                              x2 * (2.4801587301587302e-5f +
                                                    x2 * -2.7557319223985891e-7f))));
  int k = ((int)q) & 3;
  float ss = sp, cc = cp;
  if (k == 1) {
    ss = cp;
    cc = -sp;
  } else if (k == 2) {
    ss = -sp;
    cc = -cp;
  } else if (k == 3) {
    ss = -cp;
    cc = sp;
  }
  *s = ss;
  *c = cc;
}

// Counter-based PCG hash RNG (Jarzynski & Olano, "Hash Functions for GPU Rendering", JCGT
// 2020), the scheme of the reference's wp.rand_init(seed, offset) / wp.randf.
MJ_DEV uint32_t pcg_hash(uint32_t state) {
  uint32_t b = state * 747796405u + 2891336453u;
  uint32_t c = ((b >> ((b >> 28u) + 4u)) ^ b) * 277803737u;
  return (c >> 22u) ^ c;
}
MJ_DEV uint32_t rand_init(int32_t seed, int32_t offset) {
  return pcg_hash((uint32_t)seed + pcg_hash((uint32_t)offset));
}
MJ_DEV float randf(uint32_t state) {
  uint32_t s = pcg_hash(state);
  return (float)(s >> 8) * (1.0f / 16777216.0f);
}
MJ_DEV float rand_at(int32_t seed, int32_t counter) { return randf(rand_init(seed, counter)); }

// Branchless orthonormal basis (Duff et al. 2017), texture.py:104-110.
MJ_DEV void orthonormal_basis(V3 n, V3* t1, V3* t2) {
  float sign = n.z >= 0.0f ? 1.0f : -1.0f;
  float a = -1.0f / (sign + n.z);
  float b = n.x * n.y * a;
  *t1 = V3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
  *t2 = V3{b, sign + n.y * n.y * a, -n.y};
}

MJ_DEV float luminance(V3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

}  // namespace mjpbr
