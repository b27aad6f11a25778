// wavefront.hip — persistent path tracer over a batch of world groups with mixed camera
// resolutions, in one launch (no counterpart in the reference, whose path tracer is the per-pixel
// megakernel of pathtrace.py:200-438; the per-path arithmetic here is that kernel's, path.hpp).
//
// Work items are runs of 16x16 pixel tiles of one camera of one world of one group (every group
// has its own cameras, world count and outputs; four tiles a run at 1 sample a pixel, so that a
// run holds more paths than a workgroup has lanes, else one), handed out by a global atomic
// counter to a persistent grid. A
// workgroup stages the item's world (scene BVH + geom frames) in LDS (a world whose records exceed
// the LDS is staged into its group's world_stage by a launch before, and read from there); its
// four waves then draw the item's tiles x 256 x spp paths from an LDS counter, pixel by pixel
// (path id pixel * spp + sample: the paths a wave claims together share few pixels, so their camera
// segments trace nearly the same rays; sample-major ids measured 7-9 % slower). Paths end after
// different numbers of segments (misses, Russian roulette, depth limit): after each segment a
// wave ballots its finished lanes and, once at least `refill` lanes are free, claims that many
// new paths with one LDS atomic and hands them to the free lanes by their prefix rank (mbcnt).
// Lanes stay busy with live paths instead of idling until the wave's longest path ends. A path's radiance (and with the denoiser
// its first-hit features) goes to its sample slot in the workgroup's own region; once every path
// of the item has ended, the workgroup averages each pixel's samples in sample order and writes
// the display colour (or the noisy HDR and features), so the results are bit-identical to the
// megakernel's.
//
// Every launch is asynchronous and allocation-free on the caller's stream (the work counter is
// cleared by a kernel, not a memset), so the launches (counter reset, the staging of oversized
// worlds when a group has them, the persistent kernel) can be captured in a graph. A batch holds
// up to MJPBR_WF_MAX_GROUPS groups; an item finds its group by a binary search over the groups'
// first items.
#pragma clang fp contract(off)
#include "abi_util.hpp"
#include "path.hpp"

namespace mjpbr {

constexpr int kWfMaxGroups = MJPBR_WF_MAX_GROUPS;
constexpr int kTilePixels = kTile * kTile;
static_assert(kTilePixels == 256, "a path's slot keeps its sample above bit 8 and its tile pixel below");

struct WfGroup {
  MjpbrScene sc;
  MjpbrOutputs out;
  TileMap tm;
  int64_t item_base;   // first work item of the group
  int64_t stage_base;  // global groups: first index of their worlds in wf_stage_kernel's grid
  int32_t global;      // 1: the world's records are read in place from sc.geo.world_stage
  int32_t items_per_world;                    // every camera's tiles in runs of tiles_per_item
  int32_t item_cam_base[MJPBR_MAX_CAMS + 1];  // first item of each camera within a world
};

struct WfHeader {
  int32_t ngroup;
  int32_t spp;
  int64_t nitems;
  int64_t npixels;
  WfGroup groups[kWfMaxGroups];
};

// The records of every world of the global groups into their world_stage (as stage_prep_kernel
// does for one scene), one workgroup a world; runs before wf_kernel on the stream.
__global__ void __launch_bounds__(256) wf_stage_kernel(const WfHeader* __restrict__ hdr) {
  const int64_t b = blockIdx.x;
  for (int g = 0; g < hdr->ngroup; ++g) {
    const WfGroup& G = hdr->groups[g];
    if (!G.global || b < G.stage_base || b >= G.stage_base + G.sc.geo.nworld) continue;
    const int w = (int)(b - G.stage_base);
    stage_world_to(G.sc, w, G.sc.geo.world_stage + (int64_t)w * stage_floats(G.sc));
    return;
  }
}

constexpr size_t kHeaderBytes = (sizeof(WfHeader) + 255) / 256 * 256;
constexpr size_t kCounterBytes = 256;

MJ_DEV int lanes_below(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// The last group whose first work item is at most `item`: a binary search over the groups'
// non-decreasing item bases.
MJ_DEV int find_group(const WfHeader* h, int64_t item) {
  int lo = 0, hi = h->ngroup - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (item >= h->groups[mid].item_base) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void wf_reset_kernel(int* counter) {
  if (threadIdx.x == 0) counter[0] = 0;
}

// Floats a path sample keeps until its tile is resolved: the radiance, and with the denoiser also
// the first hit's albedo, normal and emission (the per-pixel kernel's per-sample sums).
MJ_HOST_DEV inline int slot_floats(const MjpbrSettings& s) { return s.denoise ? 12 : 3; }

// Tiles of one camera a work item holds: four at 1 sample a pixel, so that the lanes have more
// paths than lanes to refill from (one tile would be one path a lane), else one. An item's pixels
// q = tile * 256 + tile pixel fit `item_bits(spp)` bits.
MJ_HOST_DEV inline int tiles_per_item(int spp) { return spp == 1 ? 4 : 1; }
MJ_HOST_DEV inline int item_bits(int spp) { return spp == 1 ? 10 : 8; }

MJ_HOST_DEV inline int64_t region_floats(const MjpbrSettings& s) {
  return (int64_t)tiles_per_item(s.path_samples) * kTilePixels * s.path_samples * slot_floats(s);
}

// Work items of one world of a group (each camera's tiles in runs of tiles_per_item), and the
// first item of every camera.
MJ_HOST_DEV inline int item_layout(const TileMap& tm, int spp, int32_t* cam_base) {
  const int t = tiles_per_item(spp);
  int items = 0;
  for (int c = 0; c < tm.nrender; ++c) {
    if (cam_base) cam_base[c] = items;
    items += (tm.tile_base[c + 1] - tm.tile_base[c] + t - 1) / t;
  }
  if (cam_base) cam_base[tm.nrender] = items;
  return items;
}

// Pixel of item pixel q (tile q >> 8 of the item's run, from tile `first_tile` of its camera).
MJ_DEV void item_pixel(int first_tile, int tiles_x, int q, int* px, int* py) {
  const int tl = first_tile + (q >> 8), lp = q & 255;
  const int tx = tl % tiles_x, ty = tl / tiles_x;
  const int blk = lp >> 6, ln = lp & 63;
  *px = tx * kTile + (blk & 1) * 8 + (ln & 7);
  *py = ty * kTile + (blk >> 1) * 8 + (ln >> 3);
}

// A pixel's outputs from its sample sums (in sample order) over `fn` samples: the per-pixel
// kernel's arithmetic after its sample loop (pathtrace.py:340-438, pt.hip); the display colour and
// HDR, or with the features the noisy HDR and the features.
template <bool kFeatures>
MJ_DEV void write_pixel(const MjpbrOutputs& out, const MjpbrSettings& s, int w, int64_t adr, float exposure, float fn, V3 hdr,
                        V3 albedo_sum, V3 normal_sum, V3 emission_sum, float lum_sq) {
  hdr = hdr / fn;
  if (!kFeatures) {
    if (s.render_hdr) st3(out.hdr + 3 * ((int64_t)w * out.hdr_stride + adr), hdr);
    out.rgb[(int64_t)w * out.rgb_stride + adr] = encode_display(display_transform(hdr * exposure, s.tone_mapper));
    return;
  }
  const int64_t a = (int64_t)w * out.aov_stride + adr;
  st3(out.hdr + 3 * ((int64_t)w * out.hdr_stride + adr), hdr);
  const V3 alb = albedo_sum / fn;
  out.aov_albedo[3 * a] = float_to_half(alb.x);
  out.aov_albedo[3 * a + 1] = float_to_half(alb.y);
  out.aov_albedo[3 * a + 2] = float_to_half(alb.z);
  st3(out.aov_emission + 3 * a, emission_sum * (exposure / fn));
  if (len_sq(normal_sum) > 0.0f) normal_sum = normalize(normal_sum);
  out.aov_normal[3 * a] = float_to_half(normal_sum.x);
  out.aov_normal[3 * a + 1] = float_to_half(normal_sum.y);
  out.aov_normal[3 * a + 2] = float_to_half(normal_sum.z);
  const float mean_lum = luminance(hdr) * exposure;
  out.aov_variance[a] = fmaxf(lum_sq / fn - mean_lum * mean_lum, 0.0f) / fn;
}

// Pixel q of a finished item from the workgroup's sample region, where sample k of item pixel q
// sits in slot (k << item_bits) | q (a wave's lanes read adjacent slots): the per-pixel kernel's
// arithmetic after its sample loop (pathtrace.py:340-438, pt.hip), sums in sample order; the
// display colour and HDR, or with the features (s.denoise) the noisy HDR and the features.
template <bool kFeatures>
MJ_DEV void resolve_pixel(const WfGroup& G, const MjpbrSettings& s, int w, int cidx, int first_tile, int q, const float* region) {
  int px, py;
  item_pixel(first_tile, G.tm.tiles_x[cidx], q, &px, &py);
  Pixel pix;
  if (!setup_pixel(G.sc.cams, G.out, s, w, cidx, px, py, G.tm.ray_offset[cidx], &pix) || !pix.want_rgb) return;
  const int spp = s.path_samples, sf = kFeatures ? 12 : 3;
  const float exposure = G.sc.env.exposure[w % G.sc.env.exposure_n];
  const MjpbrOutputs& out = G.out;
  const int64_t adr = out.rgb_adr[cidx] + pix.pixelid;
  const float* sl = region + (int64_t)q * sf;
  const int step = (1 << item_bits(spp)) * sf;
  V3 hdr = v3s(0.0f);
  const float fn = (float)spp;
  if (!kFeatures) {
    // eight samples' loads in flight at a time, summed in sample order
    int k = 0;
    for (; k + 8 <= spp; k += 8, sl += 8 * step) {
      V3 r[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) r[j] = ld3(sl + j * step);
#pragma unroll
      for (int j = 0; j < 8; ++j) hdr += r[j];
    }
    for (; k < spp; ++k, sl += step) hdr += ld3(sl);
    write_pixel<false>(out, s, w, adr, exposure, fn, hdr, hdr, hdr, hdr, 0.0f);
    return;
  }
  V3 albedo_sum = v3s(0.0f), normal_sum = v3s(0.0f), emission_sum = v3s(0.0f);
  float lum_sq = 0.0f;
  auto add = [&](const float* v) {
    const V3 radiance = ld3(v);
    hdr += radiance;
    albedo_sum += ld3(v + 3);
    normal_sum += ld3(v + 6);
    emission_sum += ld3(v + 9);
    const float lum = luminance(radiance) * exposure;
    lum_sq += lum * lum;
  };
  // four samples' loads in flight at a time, summed in sample order
  int k = 0;
  for (; k + 4 <= spp; k += 4, sl += 4 * step) {
    float v[4][12];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int q = 0; q < 12; ++q) v[j][q] = sl[j * step + q];
#pragma unroll
    for (int j = 0; j < 4; ++j) add(v[j]);
  }
  for (; k < spp; ++k, sl += step) add(sl);
  write_pixel<true>(out, s, w, adr, exposure, fn, hdr, albedo_sum, normal_sum, emission_sum, lum_sq);
}

// kCount: adds the rays traced to ray_counts (the same image). kGlobal: some group reads its
// worlds' records in place from world_stage (a per-item branch; 3 % slower, so only batches with
// such a group take this instance). kFeatures (= s.denoise): the first-hit features are kept per
// sample; as a runtime flag they stayed live, and spilled, in the path loop of every launch
// (material_grid at 16 spp 74.3 against 78.4 ms without them). Workgroup b keeps its item's
// samples in region b of `slots` and resolves the item at the top of the next item iteration,
// after the barrier that ends it.
//
// kTiles: tiles a work item (tiles_per_item, 4 at 1 sample a pixel): its own instance, since the
// single-tile instance keeps the tile's corner uniform where the general one looks each path's
// tile up (1-3 % slower at 16 samples, 5 % with meshes at 8 waves).
template <int kWavesPerEu, bool kCount, bool kGlobal, bool kFeatures, int kTiles>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWavesPerEu, kWavesPerEu)))
wf_kernel(const WfHeader* __restrict__ hdr, int* counter, float* slots, MjpbrSettings s, int seed, int refill, int64_t item_limit,
          uint64_t* ray_counts) {
  extern __shared__ float lds[];
  __shared__ int s_item;
  __shared__ int s_pool;
  __shared__ int2 s_corner[kTiles];
  const int lane = threadIdx.x & 63;
  const int64_t nitems = item_limit > 0 && item_limit < hdr->nitems ? item_limit : hdr->nitems;
  const int spp = s.path_samples;
  constexpr int bits = kTiles == 4 ? 10 : 8, per_item = kTiles;
  const int sf = kFeatures ? 12 : 3;
  float* region = slots + (int64_t)blockIdx.x * region_floats(s);
  RayCount rays{0u, 0u, 0u};
  int done_g = -1, done_w = 0, done_cidx = 0, done_first = 0;  // the item to resolve
  for (;;) {
    __syncthreads();  // every wave is done with the previous item's LDS, pool and sample slots
    // (items of four tiles hold one sample a pixel, written as their paths end)
    if (kTiles == 1 && done_g >= 0) resolve_pixel<kFeatures>(hdr->groups[done_g], s, done_w, done_cidx, done_first, threadIdx.x, region);
    if (threadIdx.x == 0) {
      s_item = atomicAdd(counter, 1);
      s_pool = 0;
    }
    __syncthreads();  // (also: the resolve has read every slot before new paths overwrite them)
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    if (item >= nitems) break;
    const int g = __builtin_amdgcn_readfirstlane(find_group(hdr, item));
    const WfGroup& G = hdr->groups[g];
    const MjpbrScene& sc = G.sc;
    const int64_t local = item - G.item_base;
    const int w = (int)(local / G.items_per_world);
    const int r = (int)(local - (int64_t)w * G.items_per_world);
    int cidx = 0;
    while (cidx + 1 < G.tm.nrender && r >= G.item_cam_base[cidx + 1]) ++cidx;
    // the item: tiles first_tile .. + ntiles of camera cidx, its paths pixel by pixel
    const int first_tile = (r - G.item_cam_base[cidx]) * per_item;
    const int ntiles = min(per_item, G.tm.tile_base[cidx + 1] - G.tm.tile_base[cidx] - first_tile);
    const int pool = ntiles * kTilePixels * spp;
    // the corner of each tile of the item: one uniform value, or with several tiles in LDS (read
    // where a path starts and ends: registers held through the path loop would spill)
    const int2 corner0 = make_int2((first_tile % G.tm.tiles_x[cidx]) * kTile, (first_tile / G.tm.tiles_x[cidx]) * kTile);
    if (kTiles > 1) {
      if (threadIdx.x < kTiles) {
        const int tl = first_tile + min((int)threadIdx.x, ntiles - 1);
        s_corner[threadIdx.x] = make_int2((tl % G.tm.tiles_x[cidx]) * kTile, (tl / G.tm.tiles_x[cidx]) * kTile);
      }
      __syncthreads();  // the corners before any path reads them (global_world need not wait)
    }
    auto pixel_of = [&](int q, int* px, int* py) {
      const int lp = q & 255, blk = lp >> 6, ln = lp & 63;
      const int2 c = kTiles == 1 ? corner0 : s_corner[q >> 8];
      *px = c.x + (blk & 1) * 8 + (ln & 7);
      *py = c.y + (blk >> 1) * 8 + (ln >> 3);
    };
    done_g = g;
    done_w = w;
    done_cidx = cidx;
    done_first = first_tile;
    // (G is the same for the whole workgroup, so both branches keep their barriers uniform)
    SceneView sv = kGlobal && G.global ? global_world(sc, w, lds, G.tm.mesh_cache_nodes, G.tm.mesh_cache_n)
                                       : stage_world(sc, w, lds, G.tm.mesh_cache_nodes, G.tm.mesh_cache_n);
    const float exposure = sc.env.exposure[w % sc.env.exposure_n];
    PtCtx c{sc, s, sv, w, exposure, &rays};
    const int total_rays = sc.cams.total_rays;
    const int ray_offset = G.tm.ray_offset[cidx];
    const float* xm = sc.cams.cam_xmat + ((int64_t)w * sc.cams.ncam + sc.cams.cam_id_map[cidx]) * 9;
    const V3 head_dir = -V3{xm[2], xm[5], xm[8]};
    const bool want_depth = G.out.render_depth[cidx] != 0, want_seg = G.out.render_seg[cidx] != 0;

    PathState st;
    Primary first;
    bool alive = false, pool_done = false;
    int pseed = 0, slot = 0;  // slot: (sample << bits) | item pixel (the path's sample and pixel)
    float dcz = 0.0f;
    for (;;) {
      const uint64_t dead = __ballot(!alive);
      const int ndead = __popcll(dead);
      if (!pool_done && ndead >= refill) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&s_pool, ndead);
        base = __shfl(base, 0);
        if (base >= pool) {
          pool_done = true;
        } else if (!alive) {
          const int pid = base + lanes_below(dead);
          if (pid < pool) {
            const int q = pid / spp, sample = pid - q * spp;  // a pixel's samples on adjacent lanes
            int px, py;
            pixel_of(q, &px, &py);
            Pixel pix;
            if (setup_pixel(sc.cams, G.out, s, w, cidx, px, py, ray_offset, &pix)) {
              V3 o, d, dc;
              float md;
              if (!pix.want_rgb) {
                if (sample == 0) {
                  const int centre = kSubpixel / 2;
                  primary_ray(pix, s, kSubpixel, centre, centre, &o, &d, &md, &dc);
                  ++rays.segments;
                  Hit hit = trace_closest(sv, o, d, md, s.cull_backfaces != 0);
                  write_geometry(G.out, w, pix, hit, dc);
                }
              } else {
                pseed = seed + w * total_rays + pix.rayid;
                path_primary(pix, s, pseed, sample, &o, &d, &md, &dc);
                path_begin(&st, &first, o, d, md);
                slot = (sample << bits) | q;
                dcz = dc.z;
                alive = true;
              }
            }
          }
        }
      }
      if (ndead == 64 && pool_done) break;
      if (alive && !path_step(c, &st, head_dir, pseed, slot >> bits, &first)) {
        alive = false;
        if (kTiles == 1) {
          float* sl = region + (int64_t)slot * sf;
          st3(sl, st.radiance);
          if (kFeatures) {
            st3(sl + 3, first.albedo);
            st3(sl + 6, first.normal);
            st3(sl + 9, first.emission);
          }
        }
        if ((slot >> bits) == 0) {  // sample 0
          int px, py;
          pixel_of(slot, &px, &py);
          const int pixelid = py * sc.cams.cam_res[cidx][0] + px;
          const float depth = first.hit.geom_id != -1 ? first.hit.dist * -dcz : 0.0f;
          if (want_seg) {
            int2 seg = first.hit.geom_id != -1 ? make_int2(first.hit.geom_id, 5) : make_int2(-1, -1);
            reinterpret_cast<int2*>(G.out.seg)[(int64_t)w * G.out.seg_stride + G.out.seg_adr[cidx] + pixelid] = seg;
          }
          if (want_depth) G.out.depth[(int64_t)w * G.out.depth_stride + G.out.depth_adr[cidx] + pixelid] = depth;
          if (s.render_guides) {
            int64_t a = (int64_t)w * G.out.guide_stride + G.out.rgb_adr[cidx] + pixelid;
            st3(G.out.albedo + 3 * a, first.base_color);
            st3(G.out.normal + 3 * a, first.normal);
          }
          if (kFeatures) G.out.aov_depth[(int64_t)w * G.out.aov_stride + G.out.rgb_adr[cidx] + pixelid] = depth;
          if (kTiles > 1) {  // one sample a pixel: its outputs now (the resolve's sums of one sample)
            V3 hdr = v3s(0.0f), albedo_sum = v3s(0.0f), normal_sum = v3s(0.0f), emission_sum = v3s(0.0f);
            float lum_sq = 0.0f;
            hdr += st.radiance;
            if (kFeatures) {
              albedo_sum += first.albedo;
              normal_sum += first.normal;
              emission_sum += first.emission;
              const float lum = luminance(st.radiance) * exposure;
              lum_sq += lum * lum;
            }
            write_pixel<kFeatures>(G.out, s, w, G.out.rgb_adr[cidx] + pixelid, exposure, 1.0f, hdr, albedo_sum, normal_sum, emission_sum,
                                   lum_sq);
          }
        }
      }
    }
  }
  if (kCount) flush_ray_counts(ray_counts, rays);  // the item loop's exit is uniform: every lane is here
}

int device_cus() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cached[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 256;
    cached[dev] = n;
  }
  return cached[dev];
}

constexpr int kMaxGroupsPerCu = 16;

// Work items (16x16 tiles) of the batch, or a negative code when a group's cameras are invalid.
int64_t count_items(const MjpbrScene* scenes, int ngroup, int spp) {
  int64_t items = 0;
  for (int g = 0; g < ngroup; ++g) {
    TileMap tm;
    int nlaunch;
    if (make_tile_map(&scenes[g], &tm, &nlaunch)) return -1;
    items += (int64_t)item_layout(tm, spp, nullptr) * scenes[g].geo.nworld;
  }
  return items;
}

}  // namespace mjpbr

using namespace mjpbr;

extern "C" int64_t mjpbr_pt_wavefront_workspace(const MjpbrScene* scenes, int ngroup, const MjpbrSettings* s, int groups_per_cu) {
  if (ngroup < 1 || ngroup > kWfMaxGroups || s->path_samples < 1 || groups_per_cu < 1 || groups_per_cu > kMaxGroupsPerCu) return -1;
  const int64_t items = count_items(scenes, ngroup, s->path_samples);
  if (items < 1) return -1;
  int64_t regions = (int64_t)device_cus() * groups_per_cu;
  if (regions > items) regions = items;
  return (int64_t)(kHeaderBytes + kCounterBytes) + regions * region_floats(*s) * (int64_t)sizeof(float);
}

extern "C" int mjpbr_pt_wavefront_prepare(const MjpbrScene* scenes, const MjpbrOutputs* outs, int ngroup, const MjpbrSettings* s,
                                          void* workspace, int64_t workspace_bytes, MjpbrWavefrontPlan* plan, void* stream) {
  if (ngroup < 1 || ngroup > kWfMaxGroups) return set_error(MJPBR_EINVAL, "wavefront: 1 to %d world groups, got %d", kWfMaxGroups, ngroup);
  if (s->path_samples < 1) return set_error(MJPBR_EINVAL, "path_samples must be at least 1");
  if (s->max_depth < 0 || s->max_depth + 4 * s->max_layers >= kJitterDepth)
    return set_error(MJPBR_EINVAL, "max_depth + 4 * max_layers must be below %d", kJitterDepth);
  const int64_t region_bytes = region_floats(*s) * (int64_t)sizeof(float);
  const int64_t fixed = (int64_t)(kHeaderBytes + kCounterBytes);
  if (!workspace || workspace_bytes < fixed + region_bytes)
    return set_error(MJPBR_EINVAL, "wavefront: workspace of %lld bytes, need at least %lld (one sample region)", (long long)workspace_bytes,
                     (long long)(fixed + region_bytes));
  WfHeader* h = new WfHeader();
  h->ngroup = ngroup;
  h->spp = s->path_samples;
  int64_t items = 0, pixels = 0, staged = 0;
  size_t lds = 0;
  for (int g = 0; g < ngroup; ++g) {
    WfGroup& G = h->groups[g];
    G.sc = scenes[g];
    G.out = outs[g];
    int nlaunch;
    int rc = make_tile_map(&scenes[g], &G.tm, &nlaunch);
    bool global = false;
    if (!rc) rc = global_records(&scenes[g], &global);
    if (!rc && global && reinterpret_cast<uintptr_t>(scenes[g].geo.world_stage) % 16)
      rc = set_error(MJPBR_EINVAL, "wavefront: group %d world_stage must be 16-byte aligned", g);
    if (rc) {
      delete h;
      return rc;
    }
    // groups whose records fit the LDS stage their world directly (no per-render prep); the others
    // are staged into their world_stage by wf_stage_kernel and read from there
    G.global = global ? 1 : 0;
    G.stage_base = staged;
    if (global) staged += scenes[g].geo.nworld;
    else G.sc.geo.world_stage = nullptr;
    size_t b = global ? 0 : host_stage_bytes(&scenes[g]);
    b += plan_mesh_cache(&scenes[g], s->has_meshes != 0, b, &G.tm);
    lds = b > lds ? b : lds;
    G.item_base = items;
    G.items_per_world = item_layout(G.tm, s->path_samples, G.item_cam_base);
    items += (int64_t)G.items_per_world * scenes[g].geo.nworld;
    pixels += (int64_t)scenes[g].geo.nworld * scenes[g].cams.total_rays;
  }
  if (items >= (1ll << 31)) {
    delete h;
    return set_error(MJPBR_EINVAL, "wavefront: too many tiles (%lld)", (long long)items);
  }
  h->nitems = items;
  h->npixels = pixels;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(workspace, h, sizeof(WfHeader), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  delete h;
  if (e != hipSuccess) return set_error(MJPBR_ELAUNCH, "wavefront prepare: %s", hipGetErrorString(e));
  plan->nitems = items;
  plan->npixels = pixels;
  plan->nstaged = staged;
  plan->nregions = (workspace_bytes - fixed) / region_bytes;
  plan->lds_bytes = (int32_t)lds;
  plan->ngroup = ngroup;
  plan->settings = *s;
  return MJPBR_OK;
}

extern "C" int mjpbr_pt_wavefront_launch(void* workspace, const MjpbrWavefrontPlan* plan, const MjpbrWavefrontConfig* cfg, int32_t seed,
                                         void* stream) {
  if (!workspace || plan->nitems < 1 || plan->nregions < 1) return set_error(MJPBR_EINVAL, "wavefront: not prepared");
  int refill = cfg->refill_threshold;
  if (refill < 1 || refill > 64) return set_error(MJPBR_EINVAL, "wavefront: refill_threshold must be in [1, 64], got %d", refill);
  if (cfg->groups_per_cu < 1 || cfg->groups_per_cu > kMaxGroupsPerCu)
    return set_error(MJPBR_EINVAL, "wavefront: groups_per_cu must be in [1, %d], got %d", kMaxGroupsPerCu, cfg->groups_per_cu);
  const WfHeader* hdr = reinterpret_cast<const WfHeader*>(workspace);
  int* counter = reinterpret_cast<int*>(reinterpret_cast<char*>(workspace) + kHeaderBytes);
  float* slots = reinterpret_cast<float*>(reinterpret_cast<char*>(workspace) + kHeaderBytes + kCounterBytes);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(wf_reset_kernel, dim3(1), dim3(64), 0, st, counter);
  int rc = check_launch("wf_reset_kernel");
  if (rc) return rc;
  if (plan->nstaged > 0) {
    if (plan->nstaged >= (1ll << 31)) return set_error(MJPBR_EINVAL, "wavefront: too many staged worlds (%lld)", (long long)plan->nstaged);
    hipLaunchKernelGGL(wf_stage_kernel, dim3((unsigned)plan->nstaged), dim3(256), 0, st, hdr);
    rc = check_launch("wf_stage_kernel");
    if (rc) return rc;
  }
  int64_t items = cfg->item_limit > 0 && cfg->item_limit < plan->nitems ? cfg->item_limit : plan->nitems;
  // workgroup b owns sample region b: the grid never exceeds the regions the workspace holds
  int64_t grid = (int64_t)device_cus() * cfg->groups_per_cu;
  if (grid > items) grid = items;
  if (grid > plan->nregions) grid = plan->nregions;
  const MjpbrSettings& s = plan->settings;
  using WfFn = void (*)(const WfHeader*, int*, float*, MjpbrSettings, int, int, int64_t, uint64_t*);
  const bool count = cfg->ray_counts != nullptr;
  const bool global = plan->nstaged > 0;
  const bool features = s.denoise != 0;
  if (cfg->waves_per_eu != 4 && cfg->waves_per_eu != 8)
    return set_error(MJPBR_EINVAL, "wavefront: waves_per_eu must be 4 or 8, got %d", cfg->waves_per_eu);
  // [waves 8][count][global][features]
  static const WfFn kFns[2][2][2][2] = {
      {{{wf_kernel<4, false, false, false, 1>, wf_kernel<4, false, false, true, 1>}, {wf_kernel<4, false, true, false, 1>, wf_kernel<4, false, true, true, 1>}},
       {{wf_kernel<4, true, false, false, 1>, wf_kernel<4, true, false, true, 1>}, {wf_kernel<4, true, true, false, 1>, wf_kernel<4, true, true, true, 1>}}},
      {{{wf_kernel<8, false, false, false, 1>, wf_kernel<8, false, false, true, 1>}, {wf_kernel<8, false, true, false, 1>, wf_kernel<8, false, true, true, 1>}},
       {{wf_kernel<8, true, false, false, 1>, wf_kernel<8, true, false, true, 1>}, {wf_kernel<8, true, true, false, 1>, wf_kernel<8, true, true, true, 1>}}}};
  // items of four tiles (1 sample a pixel): 4 waves a SIMD whatever waves_per_eu asks (the
  // images do not depend on it)
  static const WfFn kFns4[2][2][2] = {
      {{wf_kernel<4, false, false, false, 4>, wf_kernel<4, false, false, true, 4>}, {wf_kernel<4, false, true, false, 4>, wf_kernel<4, false, true, true, 4>}},
      {{wf_kernel<4, true, false, false, 4>, wf_kernel<4, true, false, true, 4>}, {wf_kernel<4, true, true, false, 4>, wf_kernel<4, true, true, true, 4>}}};
  const WfFn fn = tiles_per_item(s.path_samples) == 4 ? kFns4[count][global][features] : kFns[cfg->waves_per_eu == 8][count][global][features];
  hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(256), plan->lds_bytes, st, hdr, counter, slots, s, seed, refill, (int64_t)cfg->item_limit,
                     cfg->ray_counts);
  return check_launch("wf_kernel");
}
