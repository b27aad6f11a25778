/* mjpbr.h — C ABI of libmjpbr_hip.so, the MI355X (gfx950) renderer behind the
 * mujoco_warp_pbr_amd Python package.
 *
 * Every pointer in the structs below is a DEVICE pointer (HBM) when the struct is passed to an
 * mjpbr_* entry point of libmjpbr_hip.so. The CPU oracle (oracle/, test infrastructure only)
 * reads the same structs with host pointers.
 *
 * Conventions (mirroring the reference's Python API, mujoco_warp_pbr/_src/types.py:126-132):
 *   - "batched" model arrays carry a leading size `*_n` (1 or nworld) and are indexed
 *     `worldid % *_n`; per-frame data arrays (geom_xpos, cam_xmat, light_xpos, ...) are indexed
 *     by worldid directly.
 *   - matrices are row-major 3x3 (MuJoCo xmat: columns are the frame axes in world).
 *   - every entry point returns 0 on success or a negative MJPBR_E* code; mjpbr_last_error()
 *     returns the message of the last failure on the calling thread. Nothing allocates, frees
 *     or synchronises; work is enqueued on `stream` (a hipStream_t, NULL = default stream).
 */
#ifndef MJPBR_H_
#define MJPBR_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MJPBR_ABI_VERSION 1

#define MJPBR_OK 0
#define MJPBR_EINVAL (-1)  /* bad argument (shape, size, enum) */
#define MJPBR_ELAUNCH (-2) /* HIP launch or runtime error */

/* Rendering cameras of one launch at most; their layout is passed by value (kernel arguments of
 * the render kernels stay under 4 KB). A context with more cameras is rendered in launches of at
 * most this many consecutive cameras (MjpbrCameras.ray_base). */
#define MJPBR_MAX_CAMS 32
/* Texture roles (MuJoCo mjtTextureRole). */
#define MJPBR_NTEXROLE 10
/* Mip levels per texture / environment map at most. */
#define MJPBR_MAX_LEVELS 16
/* World groups of one wavefront batch at most (mjpbr_pt_wavefront_prepare). */
#define MJPBR_WF_MAX_GROUPS 64
/* Triangles per mesh-BVH leaf (padding slots have face id -1). */
#define MJPBR_BLAS_LEAF 4

/* MuJoCo geom types (mjtGeom). */
enum { MJPBR_GEOM_PLANE = 0, MJPBR_GEOM_HFIELD = 1, MJPBR_GEOM_SPHERE = 2, MJPBR_GEOM_CAPSULE = 3,
       MJPBR_GEOM_ELLIPSOID = 4, MJPBR_GEOM_CYLINDER = 5, MJPBR_GEOM_BOX = 6, MJPBR_GEOM_MESH = 7,
       MJPBR_GEOM_SDF = 8 };
/* MuJoCo light types (mjtLightType). */
enum { MJPBR_LIGHT_SPOT = 0, MJPBR_LIGHT_DIRECTIONAL = 1, MJPBR_LIGHT_POINT = 2, MJPBR_LIGHT_IMAGE = 3 };
/* Texture storage formats. */
enum { MJPBR_TEX_RGBA8 = 0, MJPBR_TEX_RGBA16F = 1 };

/* Scene geometry and acceleration structures (reference: trace.py:77-113 `Geometry`). */
typedef struct MjpbrGeometry {
  int32_t nworld;
  int32_t ngeom;
  int32_t nmesh;  /* model meshes; height field h is traced as mesh nmesh + h */
  int32_t pad0;
  const int32_t* geom_type;    /* [ngeom] */
  const int32_t* geom_dataid;  /* [n, ngeom] mesh id of mesh geoms */
  const int32_t* geom_matid;   /* [n, ngeom] */
  const float* geom_size;      /* [n, ngeom, 3] */
  const float* geom_rgba;      /* [n, ngeom, 4] */
  const float* mat_rgba;       /* [n, nmat, 4] */
  const float* geom_xpos;      /* [nworld, ngeom, 3] */
  const float* geom_xmat;      /* [nworld, ngeom, 9] */
  int32_t geom_dataid_n, geom_matid_n, geom_size_n, geom_rgba_n, mat_rgba_n, nmat;
  /* Per-world scene BVH: an implicit perfect binary tree of 2^tlas_depth leaves (heap order,
   * children of inner node i are 2i+1 and 2i+2). Inner node i stores the boxes of its two
   * children: tlas_box[w][i] = {lo0.xyz, hi0.xyz, lo1.xyz, hi1.xyz}. Leaves hold one geom id
   * (or -1). Build or refit with mjpbr_scene_bvh. */
  int32_t tlas_depth;
  int32_t nbvh;                /* number of geoms in the BVH (enabled geom groups) */
  float* tlas_box;             /* [nworld, 2^depth - 1, 12] */
  int32_t* tlas_leaf;          /* [nworld, 2^depth] */
  const int32_t* bvh_geom;     /* [nbvh] enabled geom ids in the BVH */
  /* Enabled geoms every ray query tests directly, before the BVH: ground planes, whose boxes
   * would widen every ancestor node. Planes only (the direct test is the plane routine; other
   * types listed here would never be hit). Ascending ids, disjoint from bvh_geom; the enabled
   * set is bvh_geom and direct_geom together. */
  const int32_t* direct_geom;  /* [ndirect] */
  int32_t ndirect;
  int32_t pad1;
  /* Mesh BVHs (shared by all worlds), same implicit layout, MJPBR_BLAS_LEAF triangles a leaf.
   * blas_tri[slot] = {v0.xyz, face id (int bits)}, {e1 = v1 - v0}, {e2 = v2 - v0} in the mesh
   * frame. */
  const int32_t* blas_depth;   /* [nmesh + nhfield] */
  const int32_t* blas_node_adr;/* [nmesh + nhfield] first inner node */
  const int32_t* blas_tri_adr; /* [nmesh + nhfield] first triangle slot */
  const float* blas_box;       /* [sum inner nodes, 12] */
  const float* blas_tri;       /* [slots, 12] */
  const float* mesh_aabb;      /* [nmesh + nhfield, 6] mesh-frame bounds {lo.xyz, hi.xyz} */
  /* Optional [nworld, mjpbr_world_stage_floats] buffer: mjpbr_rt_render / mjpbr_pt_render first
   * resolve every world's scene BVH, geom, light and surface records into it (one small launch),
   * then each workgroup copies its world's block into LDS. NULL: each workgroup resolves its
   * world itself (same results). The wavefront tracer ignores it. */
  float* world_stage;
  /* Optional [nmesh, 9] row-major rotations R of meshes whose BVH (blas_*, mesh_aabb) and vertex
   * arrays (MjpbrMaterials.mesh_vert / mesh_normal) are in the frame R (vertices R v): ray queries
   * and shading of their mesh and SDF geoms use geom_xmat R^T (reference trace.py:20-30,
   * io.py:138-211). NULL: every mesh in its MuJoCo frame. Height fields are never re-framed. */
  const float* mesh_frame;
} MjpbrGeometry;

/* Materials, mesh attributes and textures (reference: material.py:224-249 `Materials`). */
typedef struct MjpbrMaterials {
  const float* geom_metallic;    /* [n, ngeom] */
  const float* geom_roughness;   /* [n, ngeom] */
  const float* geom_reflectance; /* [n, ngeom] dielectric reflectance, F0 = 0.16 r^2 */
  const float* mat_reflectance;  /* [n, nmat] MuJoCo mirror reflectance */
  const float* mat_emission;     /* [n, nmat] */
  const int32_t* mat_texid;      /* [n, nmat, 10] */
  const int32_t* mat_texuniform; /* [n, nmat] */
  const float* mat_texrepeat;    /* [n, nmat, 2] */
  int32_t geom_metallic_n, geom_roughness_n, geom_reflectance_n, mat_reflectance_n;
  int32_t mat_emission_n, mat_texid_n, mat_texuniform_n, mat_texrepeat_n;
  /* Mesh attributes, mesh frame; height fields follow the meshes (flat, no uv). */
  const int32_t* mesh_vertadr;      /* [nmesh + nhfield] */
  const int32_t* mesh_faceadr;      /* [nmesh + nhfield] */
  const int32_t* mesh_facenum;      /* [nmesh + nhfield] */
  const int32_t* mesh_normaladr;    /* [nmesh + nhfield] */
  const int32_t* mesh_texcoordadr;  /* [nmesh + nhfield], -1: no texture coordinates */
  const float* mesh_vert;           /* [nvert, 3] */
  const float* mesh_normal;         /* [nnormal, 3] */
  const int32_t* mesh_face;         /* [nface, 3] */
  const int32_t* mesh_facenormal;   /* [nface, 3] */
  const int32_t* mesh_facetexcoord; /* [nface, 3] */
  const float* mesh_texcoord;       /* [ntexcoord, 2] */
  /* Textures: mip chains in one byte pool. Level l of texture t starts at byte
   * tex_level_adr[t * MJPBR_MAX_LEVELS + l]; tex_desc[t] = {width, height, nlevels, format}. */
  int32_t ntex;
  int32_t pad1;
  const int64_t* tex_level_adr;
  const int32_t* tex_desc;
  const uint8_t* tex_data;
} MjpbrMaterials;

/* Lights (reference: lights.py:40-60 `Lights`). */
typedef struct MjpbrLights {
  int32_t nlight;
  int32_t light_type_n, light_castshadow_n, light_active_n, light_attenuation_n, light_cutoff_n;
  int32_t light_exponent_n, light_diffuse_n, light_specular_n, light_intensity_n, light_range_n;
  int32_t light_bulbradius_n;
  const int32_t* light_type;       /* [n, nlight] */
  const int32_t* light_castshadow; /* [n, nlight] */
  const int32_t* light_active;     /* [n, nlight] */
  const float* light_attenuation;  /* [n, nlight, 3] */
  const float* light_cutoff;       /* [n, nlight] degrees */
  const float* light_exponent;     /* [n, nlight] */
  const float* light_diffuse;      /* [n, nlight, 3] */
  const float* light_specular;     /* [n, nlight, 3] */
  const float* light_xpos;         /* [nworld, nlight, 3] */
  const float* light_xdir;         /* [nworld, nlight, 3] */
  const float* light_intensity;    /* [n, nlight] lux / candela */
  const float* light_range;        /* [n, nlight] */
  const float* light_bulbradius;   /* [n, nlight] */
} MjpbrLights;

/* Cameras (reference: kernel_inputs.py:151-169 `Cameras`). Rendering camera i (of nrender)
 * is model camera cam_id_map[i] with resolution cam_res[i]. */
typedef struct MjpbrCameras {
  int32_t nrender;
  int32_t ncam;
  int32_t cam_fovy_n;
  int32_t cam_intrinsic_n;
  int32_t total_rays;               /* sum of every rendering camera's pixels, this launch's or not */
  int32_t ray_base;                 /* ray id of this launch's first camera's first pixel: the pixels of
                                       the context's cameras before it (0 when one launch renders them
                                       all); ray ids seed the per-pixel random numbers */
  int32_t cam_res[MJPBR_MAX_CAMS][2];
  int32_t cam_id_map[MJPBR_MAX_CAMS];
  const int32_t* cam_projection;    /* [ncam] 0 perspective, 1 orthographic */
  const float* cam_fovy;            /* [n, ncam] degrees (orthographic: extent) */
  const float* cam_sensorsize;      /* [ncam, 2] */
  const float* cam_intrinsic;       /* [n, ncam, 4] fx, fy, cx, cy (length units) */
  const float* cam_xpos;            /* [nworld, ncam, 3] */
  const float* cam_xmat;            /* [nworld, ncam, 9] */
  /* Ray bundles (reference rays.py:84-143, camera.py:194-347): NULL renders the pinhole cameras;
   * else one camera-frame ray a pixel, [total_rays, 3] each, rendering camera i's rays at its
   * pixel offset (the sum of the earlier cameras' pixels), shared by all worlds. A zero direction
   * marks a pixel without a ray. */
  const float* ray_origin;
  const float* ray_direction;
} MjpbrCameras;

/* Environment lighting (reference: kernel_inputs.py:172-187 `Environment`). */
typedef struct MjpbrEnvironment {
  const float* exposure;         /* [n] */
  const float* ambient;          /* [n, 3] constant ambient radiance */
  int32_t exposure_n, ambient_n;
  float background[3];           /* linear background colour (before 1 / exposure) */
  int32_t dfg_size;
  const float* dfg_lut;          /* [dfg_size (roughness), dfg_size (NoV), 2] */
  int32_t nenv;
  int32_t env_id_n, env_intensity_n, env_yaw_n;
  /* Radiance maps: RGBA16F mip chains, level l of map e at half offset
   * env_level_adr[e * MJPBR_MAX_LEVELS + l]; env_desc[e] = {width, height, nlevels, 0}. */
  const int64_t* env_level_adr;
  const int32_t* env_desc;
  const uint16_t* env_radiance;
  /* GGX-prefiltered stacks: [nenv, pf_levels, pf_height, pf_width, 4] f16. */
  int32_t pf_levels, pf_height, pf_width, pad2;
  const uint16_t* env_prefiltered;
  const float* env_sh;           /* [nenv, 9, 3] */
  const int32_t* env_id;         /* [n] */
  const float* env_intensity;    /* [n] */
  const float* env_yaw;          /* [n] radians */
  /* Importance sampling tables (path tracer): prob [nenv, h, w], marginal [nenv, h + 1],
   * conditional [nenv, h, w + 1]. */
  int32_t cdf_h, cdf_w;
  const float* env_prob;
  const float* env_marginal;
  const float* env_conditional;
} MjpbrEnvironment;

/* Render settings (reference: kernel_inputs.py:55-116 `Settings`). */
typedef struct MjpbrSettings {
  int32_t render_order;     /* 0 ROW, 1 TILED, 2 WORLD */
  int32_t samples_per_axis; /* real-time tier N (N x N samples a pixel) */
  float znear, zfar;
  int32_t tone_mapper;      /* 0 PBR neutral, 1 ACES, 2 clamp */
  int32_t render_hdr;
  int32_t any_rgb;
  int32_t light_model;      /* 0 physical, 1 MuJoCo */
  int32_t has_spot;
  int32_t use_shadows;
  int32_t shadow_samples;
  float shadow_light_fraction;
  int32_t headlight_active;
  float headlight_diffuse[3];
  float headlight_specular[3];
  float headlight_intensity;
  int32_t has_env;
  int32_t env_background;
  float inv_texel_angle;
  int32_t textures, base_color_map, normal_map, orm_map, emissive_map, opacity_map;
  int32_t specular_aa;
  int32_t vertex_normals;
  int32_t has_reflection;
  int32_t has_transparency;
  int32_t max_layers;
  int32_t denoise;
  int32_t cull_backfaces;
  int32_t render_guides;
  /* path tracer */
  int32_t path_samples, max_depth, rr_depth;
  /* 1 when an enabled geom is a mesh, SDF or height field (reference io.py:317 `has_meshes`);
   * 0 selects kernel instances without mesh traversal, which then skip such geoms. */
  int32_t has_meshes;
  /* Triangles of the rendered mesh geoms, counted once a geom (instances add up): a scheduling
   * hint only (the path tracer's occupancy), the images do not depend on it. */
  int32_t mesh_triangles;
} MjpbrSettings;

/* Output buffers (reference: kernel_inputs.py:190-204 `Outputs`). Per-world flat images:
 * camera c's pixels start at *_adr[c] within a world row of *_stride elements. */
typedef struct MjpbrOutputs {
  int32_t render_rgb[MJPBR_MAX_CAMS];
  int32_t render_depth[MJPBR_MAX_CAMS];
  int32_t render_seg[MJPBR_MAX_CAMS];
  int32_t rgb_adr[MJPBR_MAX_CAMS];
  int32_t depth_adr[MJPBR_MAX_CAMS];
  int32_t seg_adr[MJPBR_MAX_CAMS];
  int32_t rgb_stride, depth_stride, seg_stride, hdr_stride, guide_stride, aov_stride;
  uint32_t* rgb;     /* [nworld, rgb_stride] packed sRGB8 0xAARRGGBB: B | G<<8 | R<<16 | 255<<24 */
  float* depth;      /* [nworld, depth_stride] */
  int32_t* seg;      /* [nworld, seg_stride, 2] (id, object type) */
  float* hdr;        /* [nworld, hdr_stride, 3] linear radiance (optional) */
  float* albedo;     /* [nworld, guide_stride, 3] first-hit base colour (optional) */
  float* normal;     /* [nworld, guide_stride, 3] first-hit world normal (optional) */
  /* path tracer denoiser features, [nworld, aov_stride] (reference denoise.py:43-51) */
  uint16_t* aov_albedo;   /* f16 x3 */
  uint16_t* aov_normal;   /* f16 x3 */
  float* aov_depth;
  float* aov_emission;    /* x3 */
  float* aov_variance;
  /* Optional [3]: mjpbr_pt_render adds the rays it traced - path segments (closest hit), light
   * shadow rays, environment shadow rays - through a counting instance of its kernel (same
   * image, not for timing). NULL: not counted. */
  uint64_t* ray_counts;
} MjpbrOutputs;

/* One frame's inputs. */
typedef struct MjpbrScene {
  MjpbrGeometry geo;
  MjpbrMaterials mats;
  MjpbrLights lit;
  MjpbrCameras cams;
  MjpbrEnvironment env;
} MjpbrScene;

/* À-trous parameters (reference: denoise.py:204-248, types.py:303-306). */
typedef struct MjpbrAtrousParams {
  int32_t nworld;
  int32_t nrender;
  int32_t cam_res[MJPBR_MAX_CAMS][2];
  int32_t rgb_adr[MJPBR_MAX_CAMS];
  int32_t stride;          /* pixels a world row */
  int32_t iterations;
  float sigma_color, sigma_normal, sigma_depth, sigma_albedo;
  int32_t tone_mapper;
  int32_t render_hdr;
  const float* exposure;
  int32_t exposure_n;
  int32_t pad;
  /* NULL: filter every world; else only worlds w with world_select[w] != 0 (the others are left
   * untouched: the learned denoiser writes them). No counterpart in the reference. */
  const int32_t* world_select;
} MjpbrAtrousParams;

/* Persistent wavefront path tracer (no counterpart in the reference): tuning knobs of a launch. */
typedef struct MjpbrWavefrontConfig {
  int32_t waves_per_eu;      /* 4 or 8: register budget / occupancy variant of the kernel (at 1
                                sample a pixel, whose work items hold four tiles, always 4) */
  int32_t groups_per_cu;     /* persistent workgroups per compute unit (1..16) */
  int32_t refill_threshold;  /* a wave refills its lanes once this many paths ended (1..64) */
  int32_t item_limit;        /* > 0: only the first item_limit tiles (autotuner samples) */
  int32_t pad0;
  /* Optional [3]: rays the launch traced, added as MjpbrOutputs.ray_counts (a counting instance
   * of the kernel). NULL: not counted. */
  uint64_t* ray_counts;
} MjpbrWavefrontConfig;

/* What mjpbr_pt_wavefront_prepare recorded for the launches (host memory, filled by prepare). */
typedef struct MjpbrWavefrontPlan {
  int64_t nitems;   /* 16x16 tiles over every world of every group */
  int64_t npixels;  /* pixels over every world of every group */
  int64_t nstaged;  /* worlds of the groups whose records exceed the LDS (staged into their
                       geo.world_stage by each launch and read from there) */
  int64_t nregions; /* per-workgroup sample regions the workspace holds: the launch's grid at most */
  int32_t lds_bytes;
  int32_t ngroup;
  MjpbrSettings settings;
} MjpbrWavefrontPlan;

/* ---------------------------------------------------------------------------------------- */
/* Library */
int mjpbr_abi_version(void);
/* First 16 hex digits of the SHA-256 of this header as the library was built against it (the
 * bindings compare it with the header they read and refuse a stale library). */
const char* mjpbr_header_hash(void);
const char* mjpbr_last_error(void);
/* sizeof of every ABI struct, in declaration order, so bindings can check their mirrors. */
int mjpbr_struct_sizes(int64_t* sizes, int n);

/* Floats a world takes in MjpbrGeometry.world_stage (a multiple of 4), or a negative code. */
int64_t mjpbr_world_stage_floats(const MjpbrScene* scene);

/* Per-world scene BVH over the enabled geoms: build (leaf assignment by median split + boxes)
 * or refit (boxes only, leaf assignment kept). Replaces `mujoco_warp.refit_bvh` and the BVH of
 * `mujoco_warp.create_render_context` (reference io.py:511-528, README.md:50-52). */
int mjpbr_scene_bvh(const MjpbrGeometry* geo, int refit, void* stream);

/* Mesh BVH of one triangle mesh on the GPU (replaces the mesh BVHs of MuJoCo Warp's render
 * context, reference io.py:511-528; re-framed meshes pass their re-framed vertices, io.py:138-211):
 * verts [nvert][3], faces [nface][3] (device). Fills inner [2^depth - 1][12] (the layout of
 * MjpbrGeometry.blas_box) and tri [2^depth * MJPBR_BLAS_LEAF][12] (blas_tri). Triangles go to leaves
 * by recursive median split of their centroids along the widest axis (a stable sort a level), so
 * the tree is a pure function of the mesh. work: mjpbr_mesh_bvh_workspace bytes. */
int64_t mjpbr_mesh_bvh_workspace(int nface, int depth);
int mjpbr_mesh_bvh_build(const float* verts, const int32_t* faces, int nface, int depth, float* inner, float* tri, void* work,
                         int64_t work_bytes, void* stream);

/* Real-time tier megakernel: one thread per (world, pixel), all samples of a pixel.
 * Replaces `render` with tier REALTIME (reference render.py:430-460, megakernel :400-427). */
int mjpbr_rt_render(const MjpbrScene* scene, const MjpbrSettings* s, MjpbrOutputs* out, int32_t seed,
                    void* stream);

/* mjpbr_rt_render without a caller-provided world_stage and without any allocation: the
 * per-world staging lives in a static device array of MJPBR_NOALLOC_STAGE_BYTES the library holds
 * on each GPU it runs on (scene->geo.world_stage is ignored). No counterpart in the reference; it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops). Returns MJPBR_EINVAL when nworld x mjpbr_world_stage_floats x 4 exceeds the maximum.
 * Calls share one array per GPU: calls that overlap on one GPU (different streams) race. */
#define MJPBR_NOALLOC_STAGE_BYTES 33554432 /* 32 MiB */
int mjpbr_rt_render_noalloc(const MjpbrScene* scene, const MjpbrSettings* s, MjpbrOutputs* out, int32_t seed, void* stream);

/* Feature mask of the real-time kernel instance mjpbr_rt_render launches for these settings
 * (bits: 1 meshes, 2 supersampling, 4 soft shadows, 8 mirror, 16 layers, 32 textures): the
 * reference's kernel specialisation key (kernel_inputs.py:321-337) as compiled instances; bit 128
 * marks the instances of 2x2 supersampling that run one lane a sample (same images). */
int mjpbr_rt_kernel_features(const MjpbrSettings* s);

/* Path-traced tier megakernel: one thread per (world, pixel), `path_samples` paths each
 * (reference pathtrace.py:340-438). With s->denoise it writes the noisy HDR and the first-hit
 * features (out->aov_*) instead of RGB. */
int mjpbr_pt_render(const MjpbrScene* scene, const MjpbrSettings* s, MjpbrOutputs* out, int32_t seed,
                    void* stream);

/* Persistent wavefront path tracer over up to MJPBR_WF_MAX_GROUPS world groups, each with its own cameras, world
 * count and outputs (mixed resolutions in one launch); images, and with s->denoise the noisy HDR
 * and first-hit features, bit-identical to mjpbr_pt_render of each group with the same seed (no
 * counterpart in the reference, whose path tracer is the per-pixel megakernel of
 * pathtrace.py:340-438). Workspace: a descriptor header, the work counter and one sample region
 * per persistent workgroup (a 16x16 tile's path samples: 3 floats each, 12 with s->denoise);
 * mjpbr_pt_wavefront_workspace gives the bytes for launches of up to `groups_per_cu` (1..16)
 * workgroups per compute unit, and prepare accepts any workspace holding at least one region (a
 * launch's grid is capped at the regions the workspace holds). prepare validates, writes the descriptors (pointers of scenes[g] and outs[g] as they
 * are now) into the workspace and synchronises `stream`; launch enqueues the counter reset, the
 * staging of oversized worlds and the persistent kernel (which resolves each tile itself) on
 * `stream` without allocating or synchronising (graph-capturable). */
int64_t mjpbr_pt_wavefront_workspace(const MjpbrScene* scenes, int ngroup, const MjpbrSettings* s, int groups_per_cu);
int mjpbr_pt_wavefront_prepare(const MjpbrScene* scenes, const MjpbrOutputs* outs, int ngroup, const MjpbrSettings* s,
                               void* workspace, int64_t workspace_bytes, MjpbrWavefrontPlan* plan, void* stream);
int mjpbr_pt_wavefront_launch(void* workspace, const MjpbrWavefrontPlan* plan, const MjpbrWavefrontConfig* cfg, int32_t seed,
                              void* stream);

/* À-trous denoiser over the path tracer's noisy HDR and first-hit features (out->hdr, aov_*):
 * demodulate, `iterations` edge-avoiding passes, remodulate + tone map + encode into out->rgb
 * (and out->hdr with render_hdr) (reference denoise.py:214-248). work0 / work1: two
 * [nworld, stride, 4] f32 buffers; guides: one more [nworld, stride, 4] 32-bit buffer that the
 * demodulation fills with each pixel's normal, albedo (f16, as in aov_normal / aov_albedo) and
 * depth packed into 16 bytes, so every filter tap reads its guides in one load. */
int mjpbr_atrous(const MjpbrAtrousParams* p, const MjpbrOutputs* out, float* work0, float* work1, float* guides,
                 void* stream);

/* Context precompute (reference ibl.py:62-104, envmap.py:171-221). */
int mjpbr_dfg_lut(float* lut, int size, int nsample, void* stream);
int mjpbr_env_prefilter(const uint16_t* radiance, const int64_t* level_adr, const int32_t* desc,
                        uint16_t* stack, int levels, int height, int width, int nsample, void* stream);

/* Context tables built on the GPU from uploaded images. Every table is a pure function of its
 * image: reductions run in double precision in a fixed order, mip levels cascade in double from
 * the unrounded level above (columns averaged first, then rows; an axis of one texel is kept) and
 * each level is rounded once into its format (RGBA8: floor(v + 0.5); RGBA16F: round to nearest
 * even). `work` is caller-provided device memory of `work_doubles` doubles.
 *
 * Texture mip chain (reference texture.py:63-101): level 0 from RGBA8 texels, kept RGBA8, or
 * (format RGBA16F) colour decoded through srgb_table (256 f16 values) with alpha / 255; levels are
 * written at pool + level_byte_adr[l] (host array of nlevels byte offsets). work: 8 w h. */
int mjpbr_texture_mips(const uint8_t* rgba8, int width, int height, int format, const uint16_t* srgb_table, uint8_t* pool,
                       const int64_t* level_byte_adr, int nlevels, double* work, int64_t work_doubles, void* stream);
/* Radiance mip chain of an environment map (reference envmap.py:366-399): RGB float32 (height x
 * width x 3) -> RGBA16F levels (alpha 1) at pool + level_half_adr[l] halves (host array). work: 8 w h. */
int mjpbr_radiance_mips(const float* rgb, int width, int height, uint16_t* pool, const int64_t* level_half_adr, int nlevels,
                        double* work, int64_t work_doubles, void* stream);
/* sh[9][3]: the radiance projected on the real SH basis (solid-angle weighted), each band times
 * its Lambertian convolution over pi (1, 2/3, 1/4), so that evaluating it gives the cosine-weighted
 * mean radiance around a normal (reference envmap.py:224-249, Ramamoorthi & Hanrahan 2001).
 * work: 27 height. */
int mjpbr_env_sh9(const float* rgb, int width, int height, double* work, int64_t work_doubles, float* sh, void* stream);
/* Importance-sampling tables of an environment map (reference envmap.py:402-430): the map box
 * filtered to at most cdf_width texels wide (w x h), p proportional to luminance sin(theta) +
 * 1e-6 mean luminance; prob [h][w], marginal [h + 1] and conditional [h][w + 1] CDFs (last entries
 * exactly 1). work: w h + h. */
int mjpbr_env_sampling_tables(const float* rgb, int width, int height, int cdf_width, double* work, int64_t work_doubles, float* prob,
                              float* marginal, float* conditional, void* stream);

/* get_hdr / get_albedo / get_normal extraction of one camera into (n, h, w, 3)
 * (reference output.py:55-67). */
int mjpbr_extract_rgb_image(const float* src, int src_stride, int adr, int width, int height, int nout,
                            float* dst, void* stream);

/* ---------------------------------------------------------------- learned denoiser (U-Net)
 * No counterpart in the reference, whose only denoiser is the a-trous filter
 * (mujoco_warp_pbr/denoise.py:63-201). These are the building blocks of a 3-level U-Net trained
 * on the renderer's own low/high-sample pairs. Activations are NHWC BF16 (uint16 bit patterns)
 * with channel counts that are multiples of 32 (inputs of a convolution) or 16 (outputs);
 * accumulation is FP32 on matrix cores. Every call is asynchronous on `stream`. */

/* 3x3 convolution, zero padding, stride 1: y = act(bias + conv(x * (mask > 0), w)).
 * w is [cout][9][cin] BF16 (forward) or the mirrored transposed layout [cin][9][cout] for a data
 * gradient (then x is the output gradient and mask the layer's output activation: ReLU backward
 * fused into the load). bias may be NULL; y32, if not NULL, receives the first out_f32 channels
 * in FP32 ([pixel][out_f32]), and then y may be NULL (no BF16 output). h and w must be multiples
 * of 16, cin of 32, cout of 16. */
int mjpbr_unet_conv3x3(const uint16_t* x, const uint16_t* mask, const uint16_t* w, const float* bias, uint16_t* y, float* y32,
                       int out_f32, int n, int h, int wd, int cin, int cout, int relu, void* stream);
/* mjpbr_unet_conv3x3 (no mask, no FP32 output) that also writes the 2x2 average pooling of y,
 * pool [n][h/2][w/2][cout], equal bit for bit to mjpbr_unet_avgpool2 of y: the U-Net encoder's
 * last convolution of a level, its pooling fused into the epilogue. */
int mjpbr_unet_conv3x3_pool(const uint16_t* x, const uint16_t* w, const float* bias, uint16_t* y, uint16_t* pool, int n, int h, int wd,
                            int cin, int cout, int relu, void* stream);
/* Two consecutive 3x3 convolutions whose intermediate stays on chip (the U-Net's last two layers
 * at inference): m = relu(conv(x, w1) + b1) with x [n][h][w][32], w1 [32][9][32]; then
 * y32 [pixel][out_f32] = the first out_f32 (1..4) channels of conv(m, w2) + b2, w2 [>= 16][9][32]
 * (its first 16 output channels computed). Equal bit for bit to mjpbr_unet_conv3x3 run twice. */
int mjpbr_unet_conv3x3_pair(const uint16_t* x, const uint16_t* w1, const float* b1, const uint16_t* w2, const float* b2, float* y32,
                            int out_f32, int n, int h, int wd, void* stream);
/* The same convolution over concat(upsample2x(low), skip) without materialising it: low is
 * [n][h/2][w/2][c_split] (nearest 2x upsampling), skip [n][h][w][cin - c_split]; c_split a
 * multiple of 32. The U-Net decoder's input. */
int mjpbr_unet_conv3x3_upcat(const uint16_t* low, const uint16_t* skip, int c_split, const uint16_t* w, const float* bias,
                             uint16_t* y, int n, int h, int wd, int cin, int cout, int relu, void* stream);
/* Floats of the `part` workspace of mjpbr_unet_conv3x3_wgrad with nchunk pixel chunks. */
int64_t mjpbr_unet_wgrad_workspace(int n, int h, int wd, int cin, int cout, int nchunk);
/* Weight gradient grad[cout][9][cin] (FP32, overwritten) of a 3x3 convolution with input x and
 * output gradient g * (mask > 0); and, if bias_grad is not NULL, the bias gradient (bias_part
 * holds mjpbr_unet_bias_workspace floats). h and w multiples of 8, cin of 16, cout of 16. */
int mjpbr_unet_conv3x3_wgrad(const uint16_t* x, const uint16_t* g, const uint16_t* mask, float* part, int nchunk, float* grad,
                             float* bias_part, float* bias_grad, int n, int h, int wd, int cin, int cout, void* stream);
int64_t mjpbr_unet_bias_workspace(int n, int h, int wd, int cout);
/* 2x2 average pooling (grad = 0: x [n][h][w][c] -> y [n][h/2][w/2][c]); grad = 1: the pooling's
 * backward (x at h/2 x w/2 -> y at h x w). */
int mjpbr_unet_avgpool2(const uint16_t* x, uint16_t* y, int n, int h, int wd, int c, int grad, void* stream);
/* out [n][h][w][cl + cs] = concat(nearest 2x upsampling of low [n][h/2][w/2][cl], skip). */
int mjpbr_unet_upcat(const uint16_t* low, const uint16_t* skip, uint16_t* out, int n, int h, int wd, int cl, int cs,
                     void* stream);
/* low_grad [n][h/2][w/2][cl] = 2x2 sums of the first cl channels of g [n][h][w][cl + cs]. */
int mjpbr_unet_upcat_low_grad(const uint16_t* g, uint16_t* low_grad, int n, int h, int wd, int cl, int cs, void* stream);
/* skip_grad [n][h][w][cs] = pooling backward of pool_grad [n][h/2][w/2][cs] + the last cs
 * channels of cat_grad [n][h][w][cl + cs]. */
int mjpbr_unet_skip_grad(const uint16_t* pool_grad, const uint16_t* cat_grad, uint16_t* skip_grad, int n, int h, int wd, int cs,
                         int cl, void* stream);
/* L1 loss of pred [npix][3] + the first 3 channels of base [npix][cb] (BF16; NULL for none)
 * against target [npix][3]: loss_part[ceil(npix / 256)] partial sums of |d| and the BF16
 * gradient sign(d) / (3 npix) into grad [npix][cp] (channels 3.. zero). */
int mjpbr_unet_l1(const float* pred, const uint16_t* base, int cb, const float* target, int64_t npix, int cp, uint16_t* grad,
                  float* loss_part, void* stream);
/* Adam step `step` (1-based) on FP32 w/m/v with gradient g; if wf is not NULL, also refreshes
 * the BF16 forward (wf) and data-gradient (wd) layouts of a [cout][9][cin] weight tensor. */
int mjpbr_unet_adam(float* w, float* m, float* v, const float* g, int64_t count, float lr, float b1, float b2, float eps, int step,
                    uint16_t* wf, uint16_t* wd, int cin, int cout, void* stream);
/* BF16 forward and data-gradient layouts of FP32 weights [cout][9][cin]. */
int mjpbr_unet_weight_layouts(const float* w, int cin, int cout, uint16_t* wf, uint16_t* wd, void* stream);
/* Network input [n][h][w][cp] BF16 of one camera: log1p(max(hdr * exposure, 0)), albedo,
 * normal, depth / (1 + depth), zeros. hdr is the per-world flat HDR buffer (vec3, stride
 * hdr_stride pixels a world), albedo/normal FP16 vec3 and depth FP32 per-world AOVs (stride
 * aov_stride), adr the camera's pixel offset; exposure per world (index world % exposure_n). */
int mjpbr_unet_pack_input(const float* hdr, int hdr_stride, const uint16_t* albedo, const uint16_t* normal, const float* depth,
                          int aov_stride, int adr, const float* exposure, int exposure_n, int n, int h, int wd, int cp,
                          uint16_t* out, void* stream);
/* Denoised HDR (expm1(max(pred + base, 0)) / exposure) and display RGB of one camera from the
 * network output; hdr or rgb may be NULL; images with skip[img] != 0 (skip may be NULL) are not
 * written. */
int mjpbr_unet_unpack_output(const float* pred, const uint16_t* base, int cb, int n, int h, int wd, const float* exposure,
                             int exposure_n, int tone_mapper, float* hdr, int hdr_stride, uint32_t* rgb, int rgb_stride, int adr,
                             const int32_t* skip, void* stream);
/* Per-image guide statistics of a packed input x [n][h][w][cp]: stats [n][4] = {mean log
 * radiance, background fraction, mean depth feature over hits, mean albedo over hits}. With
 * range = {lo[4], hi[4]} (device memory, may be NULL), fallback[img] = 1 when a statistic lies
 * outside the range the network was trained on (ORed into fallback when accumulate != 0): the
 * caller then filters that image with the a-trous denoiser instead. */
int mjpbr_unet_guide_stats(const uint16_t* x, int n, int h, int wd, int cp, const float* range, float* stats, int32_t* fallback,
                           int accumulate, void* stream);
/* D [16][16] FP32 = A [16][32] . B [32][16] (BF16) through the convolutions' matrix-core operand
 * layout: a self-test of the lane mapping. */
int mjpbr_unet_mfma_probe(const uint16_t* A, const uint16_t* B, float* D, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MJPBR_H_ */
