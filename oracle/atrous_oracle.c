/* atrous_oracle.c — CPU oracle of the a-trous denoiser (reference denoise.py:63-248), three
 * stages as the reference: demodulate, `iterations` passes, finish. TEST INFRASTRUCTURE ONLY;
 * compiled inside oracle_all.c. */

static v3 half3_o(const uint16_t* p) { return mk3(half2f(p[0]), half2f(p[1]), half2f(p[2])); }

static float b3_o(int i) { return i == 2 ? 0.375f : ((i == 1 || i == 3) ? 0.25f : 0.0625f); }

/* The camera of RGB pixel `index` among the params' cameras: its first pixel (-1: none of them),
 * size and the pixel's index in it. A context's cameras may be filtered in several calls. */
static int locate_o(const MjpbrAtrousParams* p, int index, int* W, int* H, int* pid) {
  for (int c = 0; c < p->nrender; ++c) {
    int ca = p->rgb_adr[c], np_ = p->cam_res[c][0] * p->cam_res[c][1];
    if (ca >= 0 && index >= ca && index < ca + np_) {
      *W = p->cam_res[c][0];
      *H = p->cam_res[c][1];
      *pid = index - ca;
      return ca;
    }
  }
  return -1;
}

/* One edge-avoiding pass of footprint `step` (reference denoise.py:105-178): src -> dst, both
 * [nworld, stride, 4] (demodulated colour, variance). */
static void atrous_pass_o(const MjpbrAtrousParams* p, const MjpbrOutputs* out, const float* src, float* dst, int step) {
  int64_t n = (int64_t)p->nworld * p->stride;
  for (int64_t i = 0; i < n; ++i) {
    int w = (int)(i / p->stride), index = (int)(i - (int64_t)w * p->stride);
    int W = 0, H = 0, pid = 0;
    int adr = locate_o(p, index, &W, &H, &pid);
    if (adr < 0) continue;
    const float* sw = src + 4 * (int64_t)w * p->stride;
    int64_t fa = (int64_t)w * out->aov_stride;
    v3 n0 = half3_o(out->aov_normal + 3 * (fa + index));
    if (dot3(n0, n0) == 0.0f) {
      for (int k = 0; k < 4; ++k) dst[4 * i + k] = src[4 * i + k];
      continue;
    }
    int px = pid % W, py = pid / W;
    float variance = 0.0f, vw = 0.0f;
    for (int j = 0; j < 3; ++j) {
      int y = clampi(py + j - 1, 0, H - 1);
      for (int k = 0; k < 3; ++k) {
        int x = clampi(px + k - 1, 0, W - 1);
        float g = (k == 1 ? 0.5f : 0.25f) * (j == 1 ? 0.5f : 0.25f);
        variance += g * sw[4 * (adr + y * W + x) + 3];
        vw += g;
      }
    }
    float inv_l = 1.0f / (p->sigma_color * sqrtf(fmaxf(variance / vw, 0.0f)) + 1.0e-6f);
    const float* s0 = sw + 4 * index;
    float l0 = lum3(mk3(s0[0], s0[1], s0[2]));
    float z0 = out->aov_depth[fa + index];
    v3 a0 = half3_o(out->aov_albedo + 3 * (fa + index));
    float inv_a = 1.0f / p->sigma_albedo;
    float inv_z = 1.0f / fmaxf(p->sigma_depth * z0 * (float)step, 1.0e-6f);
    v3 total = mk3(0, 0, 0);
    float tv = 0.0f, ws = 0.0f;
    for (int j = 0; j < 5; ++j) {
      int y = py + (j - 2) * step;
      if (y < 0 || y >= H) continue;
      for (int k = 0; k < 5; ++k) {
        int x = px + (k - 2) * step;
        if (x < 0 || x >= W) continue;
        int q = adr + y * W + x;
        const float* s = sw + 4 * q;
        v3 c = mk3(s[0], s[1], s[2]);
        float wn = powf(fmaxf(dot3(n0, half3_o(out->aov_normal + 3 * (fa + q))), 0.0f), p->sigma_normal);
        float dist = fabsf(out->aov_depth[fa + q] - z0) * inv_z + fabsf(lum3(c) - l0) * inv_l +
                     length3(sub3(half3_o(out->aov_albedo + 3 * (fa + q)), a0)) * inv_a;
        float weight = b3_o(k) * b3_o(j) * wn * expf(-dist);
        total = add3(total, mul3(c, weight));
        tv += weight * weight * s[3];
        ws += weight;
      }
    }
    v3 c = div3(total, ws);
    dst[4 * i] = c.x; dst[4 * i + 1] = c.y; dst[4 * i + 2] = c.z;
    dst[4 * i + 3] = tv / (ws * ws);
  }
}

int oracle_atrous(const MjpbrAtrousParams* p, MjpbrOutputs* out, float* work0, float* work1) {
  int64_t n = (int64_t)p->nworld * p->stride;
  float* src = work0;
  float* dst = work1;
  int W, H, pid;
  /* _demodulate (the pixels of this call's cameras) */
  for (int64_t i = 0; i < n; ++i) {
    int w = (int)(i / p->stride), index = (int)(i - (int64_t)w * p->stride);
    if (locate_o(p, index, &W, &H, &pid) < 0) continue;
    int64_t a = (int64_t)w * out->aov_stride + index;
    v3 albedo = half3_o(out->aov_albedo + 3 * a);
    v3 c = sub3(mul3(load3(out->hdr + 3 * ((int64_t)w * out->hdr_stride + index)), p->exposure[w % p->exposure_n]),
                load3(out->aov_emission + 3 * a));
    c = mk3(fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f));
    c = mk3(c.x / fmaxf(albedo.x, 0.01f), c.y / fmaxf(albedo.y, 0.01f), c.z / fmaxf(albedo.z, 0.01f));
    float scale = 1.0f / fmaxf(lum3(albedo), 0.01f);
    src[4 * i] = c.x; src[4 * i + 1] = c.y; src[4 * i + 2] = c.z;
    src[4 * i + 3] = out->aov_variance[a] * scale * scale;
  }
  /* _atrous */
  for (int it = 0; it < p->iterations; ++it) {
    atrous_pass_o(p, out, src, dst, 1 << it);
    float* t = src;
    src = dst;
    dst = t;
  }
  /* _finish (the pixels of this call's cameras) */
  for (int64_t i = 0; i < n; ++i) {
    int w = (int)(i / p->stride), index = (int)(i - (int64_t)w * p->stride);
    if (locate_o(p, index, &W, &H, &pid) < 0) continue;
    int64_t a = (int64_t)w * out->aov_stride + index;
    v3 albedo = half3_o(out->aov_albedo + 3 * a);
    v3 exposed = add3(load3(out->aov_emission + 3 * a),
                      mk3(src[4 * i] * fmaxf(albedo.x, 0.01f), src[4 * i + 1] * fmaxf(albedo.y, 0.01f), src[4 * i + 2] * fmaxf(albedo.z, 0.01f)));
    float exposure = p->exposure[w % p->exposure_n];
    if (p->render_hdr) {
      float* h = out->hdr + 3 * ((int64_t)w * out->hdr_stride + index);
      v3 e = div3(exposed, exposure);
      h[0] = e.x; h[1] = e.y; h[2] = e.z;
    }
    out->rgb[(int64_t)w * out->rgb_stride + index] = encode_o(display_o(exposed, p->tone_mapper));
  }
  return 0;
}

/* One pass on its own, for the tests that pin it. */
int oracle_atrous_pass(const MjpbrAtrousParams* p, const MjpbrOutputs* out, const float* src, float* dst, int step) {
  atrous_pass_o(p, out, src, dst, step);
  return 0;
}
