/* oracle_common.h — shared scalar math of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is a plain-C restatement of the reference renderer
 * (julien-blanchon/mujoco_warp_pbr, mujoco_warp_pbr/_src/ files), used by tests/, bench.py's
 * cpu_baseline leg and __graft_entry__.smoke() as the checker. It is never part of the
 * product path. Built with gcc -O2 -ffp-contract=off (no FMA contraction, no fast math) so
 * that ray/primitive arithmetic is the exact IEEE sequence written.
 *
 * Parity unpinned (nothing here fixes these against the reference without importing its
 * dependencies; they are restated from MuJoCo's documented conventions and are identical in the
 * oracle and the HIP kernels): the float sequence inside MuJoCo's ray_* primitives and the
 * RAY_TOL_REL / RAY_TOL_ABS values; the orthographic and intrinsic branches of compute_ray; the
 * primitive UV conventions of _texture_*; Warp's texture filtering and mip weights (a bilinear /
 * trilinear sampler of this project's own); Warp's rand_init / randf (the PCG hash of Jarzynski
 * and Olano, JCGT 2020); the edge behaviour of Warp's watertight mesh query.
 */
#ifndef ORACLE_COMMON_H_
#define ORACLE_COMMON_H_

#include <math.h>
#include <stdint.h>
#include <string.h>
#include "../include/mjpbr.h"

#define O_PI 3.14159265358979323846f
#define O_INV_PI 0.318309886183790671538f
#define O_MAXVAL 1.0e10f          /* MuJoCo mjMAXVAL */
#define O_MINVAL 1.0e-15f         /* MuJoCo mjMINVAL */
#define O_RAY_OFFSET 1.0e-4f      /* trace.py:47 */
#define O_SHADOW_GAP 1.0e-3f      /* trace.py:49 */
#define O_UNBOUNDED 1.0e8f        /* trace.py:51 */
#define O_TOL_REL 1.0e-6f
#define O_TOL_ABS 1.0e-7f

typedef struct { float x, y, z; } v3;
typedef struct { float x, y; } v2;

static inline v3 mk3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 neg3(v3 a) { return mk3(-a.x, -a.y, -a.z); }
static inline v3 mul3(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline v3 div3(v3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
static inline v3 cwm(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float length3(v3 a) { return sqrtf(dot3(a, a)); }
static inline v3 norm3(v3 a) {
  float l = sqrtf(dot3(a, a));
  return l > 0.0f ? div3(a, l) : a;
}
static inline v3 load3(const float* p) { return mk3(p[0], p[1], p[2]); }
static inline float clampf_(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline v3 max3s(v3 a, float s) { return mk3(fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s)); }
static inline v3 min3s(v3 a, float s) { return mk3(fminf(a.x, s), fminf(a.y, s), fminf(a.z, s)); }
/* row-major 3x3 times vector, and its transpose */
static inline v3 mmul(const float* m, v3 v) {
  return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
             m[6] * v.x + m[7] * v.y + m[8] * v.z);
}
static inline v3 mtmul(const float* m, v3 v) {
  return mk3(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
             m[2] * v.x + m[5] * v.y + m[8] * v.z);
}

static inline float half2f(uint16_t h) {
  uint32_t sign = (uint32_t)(h >> 15) << 31;
  uint32_t exp = (h >> 10) & 0x1f;
  uint32_t man = h & 0x3ff;
  uint32_t bits;
  if (exp == 0) {
    if (man == 0) {
      bits = sign;
    } else { /* subnormal */
      int e = -1;
      do { man <<= 1; ++e; } while (!(man & 0x400));
      man &= 0x3ff;
      bits = sign | ((uint32_t)(127 - 15 - e) << 23) | (man << 13);
    }
  } else if (exp == 31) {
    bits = sign | 0x7f800000u | (man << 13);
  } else {
    bits = sign | ((exp + 112) << 23) | (man << 13);
  }
  float f;
  memcpy(&f, &bits, 4);
  return f;
}

static inline int f2i_bits(float f) { int i; memcpy(&i, &f, 4); return i; }

/* sin/cos of 2*pi*t by exact reduction to [-pi/4, pi/4] and Taylor polynomials (the same
 * evaluation the GPU uses, so sampled directions agree to the bit). */
static inline void sincos_turns_o(float t, float* s, float* c) {
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
                                              x2 * (2.4801587301587302e-5f + x2 * -2.7557319223985891e-7f))));
  int k = ((int)q) & 3;
  float ss = sp, cc = cp;
  if (k == 1) { ss = cp; cc = -sp; }
  else if (k == 2) { ss = -sp; cc = -cp; }
  else if (k == 3) { ss = -cp; cc = sp; }
  *s = ss;
  *c = cc;
}

/* PCG hash RNG (Jarzynski & Olano, JCGT 2020): warp.rand_init(seed, offset) / warp.randf. */
static inline uint32_t pcg_o(uint32_t state) {
  uint32_t b = state * 747796405u + 2891336453u;
  uint32_t c = ((b >> ((b >> 28u) + 4u)) ^ b) * 277803737u;
  return (c >> 22u) ^ c;
}
/* state = rand_init(seed, counter) = pcg(seed + pcg(counter)); randf advances it once more and
 * returns its top 24 bits as a float in [0, 1). */
static inline float rand_o(int32_t seed, int32_t counter) {
  uint32_t state = pcg_o((uint32_t)seed + pcg_o((uint32_t)counter));
  return (float)(pcg_o(state) >> 8) * (1.0f / 16777216.0f);
}

static inline float lum3(v3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

#endif
