/* pt_oracle.c — CPU oracle of the path-traced tier (reference mujoco_warp_pbr/_src/pathtrace.py,
 * brdf.py:116-191, envmap.py:433-490). TEST INFRASTRUCTURE ONLY; compiled inside oracle_all.c
 * after rt_oracle.c, whose brute-force ray queries and surface evaluation it reuses. */

#define PT_SUBPIXEL 15
#define PT_SLOTS 16
#define PT_DEPTHS 64
#define PT_JITTER_DEPTH 63

static float prand_o(int seed, int sample, int depth, int slot) {
  return rand_o(seed, (sample * PT_DEPTHS + depth) * PT_SLOTS + slot);
}

/* pathtrace.py:88-95 */
static v2 light_rand2_o(int seed, int sample, int depth, int nlight, int li) {
  v2 r;
  if (li < 4) {
    int slot = 4 + 2 * li;
    r.x = prand_o(seed, sample, depth, slot);
    r.y = prand_o(seed, sample, depth, slot + 1);
    return r;
  }
  int counter = (1 << 30) + ((sample * PT_DEPTHS + depth) * nlight + li) * 2;
  r.x = rand_o(seed, counter);
  r.y = rand_o(seed, counter + 1);
  return r;
}

static float power_heuristic_o(float a, float b) {
  float a2 = a * a;
  return a2 / fmaxf(a2 + b * b, 1.0e-30f);
}

/* brdf.py:116-191 */
static float smith_g1_o(float nov, float alpha) {
  float a2 = alpha * alpha;
  return 2.0f * nov / (nov + sqrtf(a2 + (1.0f - a2) * nov * nov));
}

static v3 vndf_o(v3 v, float alpha, v2 u) {
  v3 vh = norm3(mk3(alpha * v.x, alpha * v.y, v.z));
  float lensq = vh.x * vh.x + vh.y * vh.y;
  v3 t1 = mk3(1, 0, 0);
  if (lensq > 0.0f) t1 = div3(mk3(-vh.y, vh.x, 0.0f), sqrtf(lensq));
  v3 t2 = cross3(vh, t1);
  float r = sqrtf(u.x), sp, cp;
  sincos_turns_o(u.y, &sp, &cp);
  float p1 = r * cp, p2 = r * sp;
  float s = 0.5f * (1.0f + vh.z);
  p2 = (1.0f - s) * sqrtf(fmaxf(1.0f - p1 * p1, 0.0f)) + s * p2;
  v3 nh = add3(add3(mul3(t1, p1), mul3(t2, p2)), mul3(vh, sqrtf(fmaxf(1.0f - p1 * p1 - p2 * p2, 0.0f))));
  return norm3(mk3(alpha * nh.x, alpha * nh.y, fmaxf(nh.z, 0.0f)));
}

static v3 bsdf_cos_o(v3 diffuse, v3 f0, float alpha, v3 ec, v3 v, v3 l) {
  float nol = l.z, nov = fmaxf(v.z, 1.0e-4f);
  if (nol <= 0.0f) return mk3(0, 0, 0);
  v3 h = norm3(add3(l, v));
  v3 spec = cwm(spec_lobe_o(f0, alpha, fmaxf(h.z, 0.0f), nov, nol, fmaxf(dot3(l, h), 0.0f)), ec);
  return mul3(add3(mul3(diffuse, O_INV_PI), spec), nol);
}

static float pdf_bsdf_o(float alpha, float p_spec, v3 v, v3 l) {
  if (l.z <= 0.0f) return 0.0f;
  float nov = fmaxf(v.z, 1.0e-4f);
  v3 h = norm3(add3(l, v));
  float pdf_spec = smith_g1_o(nov, alpha) * d_ggx_o(fmaxf(h.z, 0.0f), alpha) / (4.0f * nov);
  return p_spec * pdf_spec + (1.0f - p_spec) * (l.z * O_INV_PI);
}

static float spec_prob_o(v3 diffuse, v3 spec) {
  float ws = lum3(spec), wd = lum3(diffuse);
  if (wd <= 0.0f) return 1.0f;
  return clampf_(ws / fmaxf(ws + wd, 1.0e-6f), 0.1f, 0.9f);
}

static v3 sample_bsdf_o(float alpha, float p_spec, v3 v, float u_lobe, v2 u) {
  if (u_lobe < p_spec) {
    v3 h = vndf_o(v, alpha, u);
    return sub3(mul3(h, 2.0f * dot3(v, h)), v);
  }
  float r = sqrtf(u.x), sp, cp;
  sincos_turns_o(u.y, &sp, &cp);
  return mk3(r * cp, r * sp, sqrtf(fmaxf(1.0f - u.x, 0.0f)));
}

/* envmap.py:433-490 */
static v3 rot_yaw_exact_o(v3 d, float yaw) {
  float s, c;
  sincos_turns_o(yaw * 0.159154943091895335768f, &s, &c);
  return mk3(c * d.x - s * d.y, s * d.x + c * d.y, d.z);
}

static int search_o(const float* cdf, int n, float u) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (cdf[mid] <= u) lo = mid;
    else hi = mid;
  }
  return lo;
}

static v3 env_sample_o(const MjpbrEnvironment* env, int e, float yaw, v2 u, float* pdf) {
  int h = env->cdf_h, w = env->cdf_w;
  const float* marginal = env->env_marginal + (int64_t)e * (h + 1);
  int r = search_o(marginal, h, u.x);
  float c0 = marginal[r];
  float dv = (u.x - c0) / fmaxf(marginal[r + 1] - c0, 1.0e-12f);
  const float* cond = env->env_conditional + ((int64_t)e * h + r) * (w + 1);
  int c = search_o(cond, w, u.y);
  float k0 = cond[c];
  float du = (u.y - k0) / fmaxf(cond[c + 1] - k0, 1.0e-12f);
  v2 uv = {((float)c + clampf_(du, 0.0f, 1.0f)) / (float)w, ((float)r + clampf_(dv, 0.0f, 1.0f)) / (float)h};
  v3 d = rot_yaw_exact_o(uv_to_dir_o(uv), yaw);
  float st, ct;
  sincos_turns_o(uv.y * 0.5f, &st, &ct);
  *pdf = env->env_prob[((int64_t)e * h + r) * w + c] * (float)(w * h) / (2.0f * O_PI * O_PI * fmaxf(st, 1.0e-6f));
  return d;
}

static float env_pdf_o(const MjpbrEnvironment* env, int e, float yaw, v3 dir) {
  int h = env->cdf_h, w = env->cdf_w;
  v2 uv = dir_to_uv_o(dir, yaw);
  int c = (int)(uv.x * (float)w);
  if (c > w - 1) c = w - 1;
  int r = (int)(uv.y * (float)h);
  if (r > h - 1) r = h - 1;
  float st, ct;
  sincos_turns_o(uv.y * 0.5f, &st, &ct);
  return env->env_prob[((int64_t)e * h + r) * w + c] * (float)(w * h) / (2.0f * O_PI * O_PI * fmaxf(st, 1.0e-6f));
}

/* pathtrace.py:128-162 */
static v3 light_sample_o(const Ctx* c, LightO L, const LobesO* lb, v3 point, v3 normal, v3 t1, v3 t2, v3 v, v2 u) {
  if (L.type != MJPBR_LIGHT_DIRECTIONAL && L.bulbradius > 0.0f) {
    v3 b1, b2;
    basis_o(norm3(sub3(point, L.pos)), &b1, &b2);
    float r = L.bulbradius * sqrtf(u.x), sp, cp;
    sincos_turns_o(u.y, &sp, &cp);
    L.pos = add3(L.pos, mul3(add3(mul3(b1, cp), mul3(b2, sp)), r));
  }
  v3 l, ds, ss;
  float dist;
  eval_light_o(&L, c->s->light_model, c->s->has_spot, point, &l, &dist, &ds, &ss);
  v3 ll = mk3(dot3(l, t1), dot3(l, t2), dot3(l, normal));
  if (ll.z <= 0.0f || (dot3(ds, ds) == 0.0f && dot3(ss, ss) == 0.0f)) return mk3(0, 0, 0);
  if (c->s->use_shadows && L.castshadow) {
    float max_t = L.type != MJPBR_LIGHT_DIRECTIONAL ? dist - O_SHADOW_GAP : O_UNBOUNDED;
    if (any_o(c, add3(point, mul3(normal, O_RAY_OFFSET)), l, max_t)) return mk3(0, 0, 0);
  }
  v3 zero = mk3(0, 0, 0);
  v3 diffuse = bsdf_cos_o(lb->diffuse_color, zero, lb->alpha, zero, v, ll);
  v3 specular = bsdf_cos_o(zero, lb->f0, lb->alpha, lb->energy_comp, v, ll);
  return add3(cwm(diffuse, ds), cwm(specular, ss));
}

static v3 direct_lights_o(const Ctx* c, const LobesO* lb, v3 point, v3 normal, v3 t1, v3 t2, v3 v, v3 head_dir, int seed,
                          int sample, int depth) {
  v3 radiance = mk3(0, 0, 0);
  const MjpbrLights* lt = &c->sc->lit;
  for (int li = 0; li < lt->nlight; ++li) {
    if (!lt->light_active[(c->w % lt->light_active_n) * lt->nlight + li]) continue;
    v2 u = light_rand2_o(seed, sample, depth, lt->nlight, li);
    radiance = add3(radiance, light_sample_o(c, fetch_light_o(lt, c->w, li), lb, point, normal, t1, t2, v, u));
  }
  if (c->s->headlight_active) {
    LightO H;
    memset(&H, 0, sizeof(H));
    H.type = MJPBR_LIGHT_DIRECTIONAL;
    H.dir = head_dir;
    H.diffuse = load3(c->s->headlight_diffuse);
    H.specular = load3(c->s->headlight_specular);
    H.attenuation = mk3(1, 0, 0);
    H.intensity = c->s->headlight_intensity / c->exposure;
    H.range = 1.0f;
    v2 u0 = {0.0f, 0.0f};
    radiance = add3(radiance, light_sample_o(c, H, lb, point, normal, t1, t2, v, u0));
  }
  return radiance;
}

static v3 env_radiance_o(const Ctx* c, v3 dir) {
  int e;
  float yaw, scale;
  env_lookup_o(&c->sc->env, c->w, &e, &yaw, &scale);
  return mul3(radiance_o(&c->sc->env, e, dir, yaw, 0.0f), scale);
}

typedef struct {
  HitO hit;
  v3 albedo, normal, emission, base_color;
} PrimaryO;

/* pathtrace.py:200-338 */
static v3 trace_path_o(const Ctx* c, v3 origin, v3 dir, float max_dist, v3 head_dir, int seed, int sample, PrimaryO* first) {
  const MjpbrSettings* s = c->s;
  const MjpbrEnvironment* env = &c->sc->env;
  v3 ambient = load3(env->ambient + 3 * (c->w % env->ambient_n));
  v3 radiance = mk3(0, 0, 0), beta = mk3(1, 1, 1);
  memset(first, 0, sizeof(*first));
  first->hit.geom_id = -1;
  first->albedo = mk3(1, 1, 1);
  int camera_ray = 1, delta_bounce = 1, depth = 0, passes = 0;
  float pdf_prev = 0.0f;
  while (depth <= s->max_depth) {
    HitO hit = closest_o(c, origin, dir, max_dist);
    if (camera_ray && passes == 0) first->hit = hit;
    if (hit.geom_id == -1) {
      if (camera_ray) {
        if (s->env_background) radiance = add3(radiance, cwm(beta, env_radiance_o(c, dir)));
        else radiance = add3(radiance, cwm(beta, div3(load3(env->background), c->exposure)));
      } else {
        v3 env_light = ambient;
        if (s->has_env) {
          float weight = 1.0f;
          if (!delta_bounce) {
            int e;
            float yaw, scale;
            env_lookup_o(env, c->w, &e, &yaw, &scale);
            weight = power_heuristic_o(pdf_prev, env_pdf_o(env, e, yaw, dir));
          }
          env_light = add3(env_light, mul3(env_radiance_o(c, dir), weight));
        }
        radiance = add3(radiance, cwm(beta, env_light));
      }
      break;
    }
    v3 point = add3(origin, mul3(dir, hit.dist));
    v3 normal;
    SurfO sf = eval_surface_o(c, &hit, point, dir, 0.0f, &normal);
    if (s->has_transparency) {
      if (prand_o(seed, sample, depth + passes, 0) >= sf.alpha && passes < 4 * s->max_layers) {
        origin = add3(point, mul3(dir, O_RAY_OFFSET));
        max_dist = max_dist - hit.dist;
        ++passes;
        continue;
      }
    }
    radiance = add3(radiance, cwm(beta, div3(sf.emissive, c->exposure)));
    if (depth == s->max_depth) break;
    max_dist = O_MAXVAL;
    if (camera_ray) {
      first->albedo = sf.base_color;
      first->normal = normal;
      first->base_color = sf.base_color;
      first->emission = div3(sf.emissive, c->exposure);
    }
    camera_ray = 0;
    if (s->has_reflection && prand_o(seed, sample, depth, 1) < mirror_o(c, hit.geom_id)) {
      dir = sub3(dir, mul3(normal, 2.0f * dot3(dir, normal)));
      origin = add3(point, mul3(normal, O_RAY_OFFSET));
      delta_bounce = 1;
      ++depth;
      continue;
    }
    v3 t1, t2;
    basis_o(normal, &t1, &t2);
    v3 view = neg3(dir);
    v3 v = mk3(dot3(view, t1), dot3(view, t2), dot3(view, normal));
    float nov = fmaxf(v.z, 1.0e-4f);
    LobesO lb = lobes_o(env, &sf, nov);
    float alpha = lb.alpha;
    v3 spec_weight = cwm(lb.spec_albedo, lb.energy_comp);
    float p_spec = spec_prob_o(lb.diffuse_color, spec_weight);
    if (depth == 0) first->albedo = add3(lb.diffuse_color, spec_weight);
    radiance = add3(radiance, cwm(beta, direct_lights_o(c, &lb, point, normal, t1, t2, v, head_dir, seed, sample, depth)));
    if (s->has_env) {
      int e;
      float yaw, scale, pdf_env;
      env_lookup_o(env, c->w, &e, &yaw, &scale);
      v2 u_env = {prand_o(seed, sample, depth, 2), prand_o(seed, sample, depth, 3)};
      v3 d_env = env_sample_o(env, e, yaw, u_env, &pdf_env);
      v3 l_env = mk3(dot3(d_env, t1), dot3(d_env, t2), dot3(d_env, normal));
      if (l_env.z > 0.0f && pdf_env > 0.0f) {
        v3 f = bsdf_cos_o(lb.diffuse_color, lb.f0, alpha, lb.energy_comp, v, l_env);
        if (!any_o(c, add3(point, mul3(normal, O_RAY_OFFSET)), d_env, O_UNBOUNDED)) {
          float weight = power_heuristic_o(pdf_env, pdf_bsdf_o(alpha, p_spec, v, l_env));
          radiance = add3(radiance, mul3(cwm(beta, cwm(f, env_radiance_o(c, d_env))), weight / pdf_env));
        }
      }
    }
    v2 u = {prand_o(seed, sample, depth, 13), prand_o(seed, sample, depth, 14)};
    v3 l = sample_bsdf_o(alpha, p_spec, v, prand_o(seed, sample, depth, 12), u);
    if (l.z <= 0.0f) break;
    float pdf = pdf_bsdf_o(alpha, p_spec, v, l);
    if (pdf <= 0.0f) break;
    beta = cwm(beta, div3(bsdf_cos_o(lb.diffuse_color, lb.f0, alpha, lb.energy_comp, v, l), pdf));
    dir = norm3(add3(add3(mul3(t1, l.x), mul3(t2, l.y)), mul3(normal, l.z)));
    origin = add3(point, mul3(normal, O_RAY_OFFSET));
    pdf_prev = pdf;
    delta_bounce = 0;
    ++depth;
    if (depth >= s->rr_depth) {
      float q = clampf_(fmaxf(beta.x, fmaxf(beta.y, beta.z)), 0.05f, 0.95f);
      if (prand_o(seed, sample, depth, 15) >= q) break;
      beta = div3(beta, q);
    }
  }
  return radiance;
}

static uint16_t f2half_o(float f) {
  /* round to nearest even float -> half (finite values in the renderer's range) */
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  int32_t exp = (int32_t)((x >> 23) & 0xff) - 127 + 15;
  uint32_t man = x & 0x7fffffu;
  if (((x >> 23) & 0xff) == 0xff) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u : 0u));
  if (exp >= 31) return (uint16_t)(sign | 0x7c00u);
  if (exp <= 0) {
    if (exp < -10) return (uint16_t)sign;
    man |= 0x800000u;
    int shift = 14 - exp;
    uint32_t half = man >> shift;
    uint32_t rem = man & ((1u << shift) - 1u), mid = 1u << (shift - 1);
    if (rem > mid || (rem == mid && (half & 1u))) ++half;
    return (uint16_t)(sign | half);
  }
  uint32_t half = ((uint32_t)exp << 10) | (man >> 13);
  uint32_t rem = man & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (half & 1u))) ++half;
  return (uint16_t)(sign | half);
}

/* Per-sample radiance of pixel (px, py) of camera c in world w, for parity tests. */
static void pt_pixel_o(const MjpbrScene* sc, const MjpbrSettings* s, MjpbrOutputs* out, int seed, int w, const PixO* pix,
                       float* samples) {
  Ctx c = {sc, s, w, 0, sc->env.exposure[w % sc->env.exposure_n]};
  const int centre = PT_SUBPIXEL / 2;
  if (!pix->want_rgb) {
    v3 o, d, dc;
    float md;
    primary_o(pix, s, PT_SUBPIXEL, centre, centre, &o, &d, &md, &dc);
    HitO hit = closest_o(&c, o, d, md);
    write_geometry_o(out, w, pix, &hit, dc);
    return;
  }
  v3 head_dir = neg3(mk3(pix->xmat[2], pix->xmat[5], pix->xmat[8]));
  int pseed = seed + w * sc->cams.total_rays + pix->rayid;
  v3 hdr = mk3(0, 0, 0), albedo_sum = mk3(0, 0, 0), normal_sum = mk3(0, 0, 0), emission_sum = mk3(0, 0, 0);
  float first_depth = 0.0f, lum_sq = 0.0f;
  int nsample = s->path_samples;
  for (int smp = 0; smp < nsample; ++smp) {
    int sx = centre, sy = centre;
    if (smp > 0) {
      sx = (int)(prand_o(pseed, smp, PT_JITTER_DEPTH, 0) * (float)PT_SUBPIXEL);
      sy = (int)(prand_o(pseed, smp, PT_JITTER_DEPTH, 1) * (float)PT_SUBPIXEL);
    }
    v3 o, d, dc;
    float md;
    primary_o(pix, s, PT_SUBPIXEL, sx, sy, &o, &d, &md, &dc);
    PrimaryO first;
    v3 radiance = trace_path_o(&c, o, d, md, head_dir, pseed, smp, &first);
    if (samples) {
      samples[3 * smp] = radiance.x;
      samples[3 * smp + 1] = radiance.y;
      samples[3 * smp + 2] = radiance.z;
    }
    if (smp == 0) {
      write_geometry_o(out, w, pix, &first.hit, dc);
      if (s->render_guides) {
        int64_t a = (int64_t)w * out->guide_stride + out->rgb_adr[pix->camid] + pix->pixelid;
        float* al = out->albedo + 3 * a;
        float* nr = out->normal + 3 * a;
        al[0] = first.base_color.x; al[1] = first.base_color.y; al[2] = first.base_color.z;
        nr[0] = first.normal.x; nr[1] = first.normal.y; nr[2] = first.normal.z;
      }
      if (first.hit.geom_id != -1) first_depth = first.hit.dist * -dc.z;
    }
    hdr = add3(hdr, radiance);
    albedo_sum = add3(albedo_sum, first.albedo);
    normal_sum = add3(normal_sum, first.normal);
    emission_sum = add3(emission_sum, first.emission);
    if (s->denoise) {
      float lum = lum3(radiance) * c.exposure;
      lum_sq += lum * lum;
    }
  }
  float fn = (float)nsample;
  hdr = div3(hdr, fn);
  int64_t adr = out->rgb_adr[pix->camid] + pix->pixelid;
  if (s->denoise) {
    int64_t a = (int64_t)w * out->aov_stride + adr;
    float* h = out->hdr + 3 * ((int64_t)w * out->hdr_stride + adr);
    h[0] = hdr.x; h[1] = hdr.y; h[2] = hdr.z;
    v3 alb = div3(albedo_sum, fn);
    out->aov_albedo[3 * a] = f2half_o(alb.x);
    out->aov_albedo[3 * a + 1] = f2half_o(alb.y);
    out->aov_albedo[3 * a + 2] = f2half_o(alb.z);
    v3 em = mul3(emission_sum, c.exposure / fn);
    out->aov_emission[3 * a] = em.x; out->aov_emission[3 * a + 1] = em.y; out->aov_emission[3 * a + 2] = em.z;
    if (dot3(normal_sum, normal_sum) > 0.0f) normal_sum = norm3(normal_sum);
    out->aov_normal[3 * a] = f2half_o(normal_sum.x);
    out->aov_normal[3 * a + 1] = f2half_o(normal_sum.y);
    out->aov_normal[3 * a + 2] = f2half_o(normal_sum.z);
    out->aov_depth[a] = first_depth;
    float mean_lum = lum3(hdr) * c.exposure;
    out->aov_variance[a] = fmaxf(lum_sq / fn - mean_lum * mean_lum, 0.0f) / fn;
    return;
  }
  if (s->render_hdr) {
    float* h = out->hdr + 3 * ((int64_t)w * out->hdr_stride + adr);
    h[0] = hdr.x; h[1] = hdr.y; h[2] = hdr.z;
  }
  out->rgb[(int64_t)w * out->rgb_stride + adr] = encode_o(display_o(mul3(hdr, c.exposure), s->tone_mapper));
}

/* Path-traces worlds [w0, w1). samples (optional): per-sample radiance
 * [w1 - w0][total_rays][path_samples][3] in camera-concatenated row-major pixel order. */
int oracle_pt_render(const MjpbrScene* sc, const MjpbrSettings* s, MjpbrOutputs* out, int32_t seed, int w0, int w1,
                     float* samples) {
  for (int w = w0; w < w1; ++w) {
    int ray_offset = sc->cams.ray_base;  /* a launch of the context's later cameras */
    for (int c = 0; c < sc->cams.nrender; ++c) {
      int W = sc->cams.cam_res[c][0], H = sc->cams.cam_res[c][1];
      for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
          PixO pix;
          float* smp = samples ? samples + 3 * (int64_t)s->path_samples * ((int64_t)(w - w0) * sc->cams.total_rays + ray_offset + py * W + px) : 0;
          if (setup_pixel_o(sc, out, s, w, c, px, py, ray_offset, &pix)) pt_pixel_o(sc, s, out, seed, w, &pix, smp);
        }
      ray_offset += W * H;
    }
  }
  return 0;
}
