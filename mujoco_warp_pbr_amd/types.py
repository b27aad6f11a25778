"""Enums, defaults and the render context (reference mujoco_warp_pbr/_src/types.py:31-322)."""

from __future__ import annotations

import dataclasses
import enum
from typing import Any


class LightModel(enum.IntEnum):
  """How MuJoCo light parameters are interpreted (types.py:31-42)."""

  PHYSICAL = 0
  MUJOCO = 1


class Tier(enum.IntEnum):
  """Rendering algorithm (types.py:45-56)."""

  REALTIME = 0
  PATH = 1


class ToneMapper(enum.IntEnum):
  """Display transform before sRGB encoding (types.py:59-70)."""

  PBR_NEUTRAL = 0
  ACES = 1
  CLAMP = 2


class RenderOrder(enum.IntEnum):
  """Pixel order of the reference's threads (types.py:73-88).

  On this renderer a workgroup always renders a 16x16 tile of one world; the order only fixes
  each pixel's ray id, which seeds its random numbers, so images match the reference's order.
  """

  ROW = 0
  TILED = 1
  WORLD = 2


class UnsetMaterial(enum.IntEnum):
  """Resolution of material fields MuJoCo leaves unset (types.py:91-100)."""

  GLTF = 0
  HEURISTIC = 1


# Bake groups of the reference (types.py:107-119). This renderer compiles every setting as a
# runtime value of the kernels' settings struct; the field is kept for API compatibility.
DEFAULT_BAKED = frozenset(
  {"samples_per_axis", "transparency", "has_reflection", "textures", "nlight", "has_env", "light_model", "denoise", "render_order"}
)


@dataclasses.dataclass
class RenderBase:
  """Geometry, cameras and output buffers (the MuJoCo Warp RenderContext part of the reference).

  Attributes:
    nworld: number of worlds.
    nrender: number of rendering cameras.
    cam_id_map: model camera of each rendering camera                 (nrender,)
    cam_res: (width, height) of each rendering camera                 (nrender, 2)
    render_rgb / render_depth / render_seg: per rendering camera      (nrender,)
    rgb_adr / depth_adr / seg_adr: first pixel of each camera in the flat per-world buffers
      (-1 when the camera does not render that output)                (nrender,)
    total_rays: pixels per world over all rendering cameras.
    rgb_data: packed sRGB8 0xAARRGGBB (B | G << 8 | R << 16 | 255 << 24) (nworld, nrgb) uint32 bits in int32
    depth_data: planar depth                                          (nworld, ndepth)
    seg_data: (geom id, object type), (-1, -1) for the background     (nworld, nseg, 2)
    znear / zfar: clip distances (zfar <= 0: unbounded).
    enabled_geom_ids: rendered geoms (enabled groups)                (nenabled,)
    tlas_depth: depth of the per-world scene BVH.
    tlas_box / tlas_leaf: per-world scene BVH (see include/mjpbr.h)
    blas_*: mesh BVHs shared by all worlds.
  """

  nworld: int
  nrender: int
  cam_id_map: Any
  cam_res: Any
  render_rgb: Any
  render_depth: Any
  render_seg: Any
  rgb_adr: Any
  depth_adr: Any
  seg_adr: Any
  total_rays: int
  rgb_data: Any
  depth_data: Any
  seg_data: Any
  znear: float
  zfar: float
  enabled_geom_ids: Any
  tlas_depth: int
  tlas_box: Any
  tlas_leaf: Any
  blas_depth: Any
  blas_node_adr: Any
  blas_tri_adr: Any
  blas_box: Any
  blas_tri: Any
  mesh_aabb: Any
  enable_backface_culling: bool = True
  enable_vertex_normals: bool = True
  use_fast_math: bool = True
  headlight_active: bool = False
  headlight_diffuse: tuple = (0.4, 0.4, 0.4)
  headlight_specular: tuple = (0.5, 0.5, 0.5)
  has_spot_lights: bool = False
  # an enabled geom is a mesh, SDF or height field (reference io.py:317 `has_meshes`)
  has_meshes: bool = True
  # triangles of the rendered mesh geoms, counted once a geom (the path tracer's kernel choice)
  mesh_triangles: int = 0
  # [nworld, floats] per-world records the render kernels resolve once a render and copy into
  # LDS (MjpbrGeometry.world_stage); None: each workgroup resolves its world itself
  world_stage: Any = None
  # enabled geoms split into the scene BVH's (bounded) and the directly tested ones (unbounded
  # planes, MjpbrGeometry.direct_geom); None: every enabled geom is in the BVH
  bvh_geom_ids: Any = None
  # mesh BVHs built in another frame (mesh_bvh_frame="auto"): per-mesh rotations R, and the traced
  # vertex / normal arrays (those meshes rotated); None: every mesh in its MuJoCo frame
  mesh_frame: Any = None
  trace_mesh_vert: Any = None
  trace_mesh_normal: Any = None
  # host description of the mesh BVHs (vertices, faces) for host-side harnesses
  mesh_bvhs: Any = None
  direct_geom_ids: Any = None


@dataclasses.dataclass
class PbrRenderContext:
  """Context for physically based rendering of nworld worlds (reference types.py:122-322).

  Batched (leading size 1 or nworld) per-world tables are indexed `worldid % shape[0]`: replace
  one with a tensor of leading size nworld to vary the value per world.
  """

  base: RenderBase
  light_model: int
  tone_mapper: int
  use_shadows: bool
  shadow_light_fraction: float
  geom_metallic: Any
  geom_roughness: Any
  geom_reflectance: Any
  mat_reflectance: Any
  light_intensity: Any
  light_range: Any
  light_bulbradius: Any
  headlight_intensity: float
  exposure: Any
  ambient_radiance: Any
  dfg_lut: Any
  background_linear: tuple
  has_env: bool
  env_background: bool
  env: Any  # envmap.Environments
  env_texel_angle: float
  env_id: Any
  env_intensity: Any
  env_yaw: Any
  textures: Any  # texture.TexturePool
  features: Any
  samples_per_pixel: int
  shadow_samples: int
  has_transparency: bool
  max_layers: int
  has_reflection: bool
  render_order: int
  block_dim: int | None
  max_registers: int
  ordered_traversal: bool
  tier: int
  path_samples: int
  max_depth: int
  rr_depth: int
  denoise: bool
  denoise_iterations: int
  aov_albedo: Any
  aov_normal: Any
  aov_depth: Any
  aov_emission: Any
  aov_variance: Any
  denoise_buffers: tuple
  render_hdr: bool
  hdr_data: Any
  render_guides: bool
  guide_albedo: Any
  guide_normal: Any
  traversal: str = "software"
  ray_origin: Any = None  # ray bundle (set_camera_rays): camera-frame origins [total_rays, 3]
  ray_direction: Any = None  # and directions; None renders the pinhole cameras
  denoiser: str = "atrous"  # "atrous" (the reference's filter) or "learned" (U-Net, a-trous fallback)
  denoiser_weights: str | None = None  # learned denoiser weights (None: the packaged ones)
  path_kernel: str = "auto"  # path tracer: "wavefront" (persistent), "megakernel" (per pixel) or "auto"
  denoise_sigma_color: float = 4.0
  denoise_sigma_normal: float = 128.0
  denoise_sigma_depth: float = 0.02
  denoise_sigma_albedo: float = 0.05
  baked: frozenset = DEFAULT_BAKED
  _host_cache: dict = dataclasses.field(default_factory=dict)
