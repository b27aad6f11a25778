"""create_render_context (reference io.py:331-634)."""

from __future__ import annotations

import dataclasses
import math
import os
import warnings
from collections.abc import Sequence

import numpy as np
import torch

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd import bvh as bvh_lib
from mujoco_warp_pbr_amd import envmap as envmap_lib
from mujoco_warp_pbr_amd import material
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd.mjmodel import as_host_model
from mujoco_warp_pbr_amd import texture
from mujoco_warp_pbr_amd.types import LightModel
from mujoco_warp_pbr_amd.types import PbrRenderContext
from mujoco_warp_pbr_amd.types import RenderBase
from mujoco_warp_pbr_amd.types import RenderOrder
from mujoco_warp_pbr_amd.types import Tier
from mujoco_warp_pbr_amd.types import ToneMapper
from mujoco_warp_pbr_amd.types import UnsetMaterial

DEFAULT_EV100 = 14.96578428466209
DFG_LUT_SIZE = 128
DFG_LUT_SAMPLES = 1024
MAX_PATH_VERTICES = 63
_MATERIAL_KEYS = ("metallic", "roughness")
_TRAVERSALS = ("software", "optix", "auto")
_MESH_BVH_FRAMES = ("auto", "mujoco")
_MESH_BVH_CONSTRUCTORS = ("cubql", "sah", "lbvh")


def _as_enum(cls, value, name: str):
  """`value` as a member of the IntEnum `cls` (members or their integer values)."""
  members = {int(m): m for m in cls}
  if isinstance(value, (int, np.integer)) and int(value) in members:
    return members[int(value)]
  listing = ", ".join(f"{cls.__name__}.{m.name}" for m in cls)
  raise ValueError(f"{name} must be one of {listing}, got {value!r}")


# Argument rules of create_render_context, checked in this order: (argument, test, message). The
# messages are the reference's (io.py:471-509; its tests/test_context_validation.py pins them).
def _at_least(low):
  return (lambda v: v >= low), f"must be at least {low}, got {{v}}"


_ARGUMENT_RULES = (
  ("samples_per_pixel",) + _at_least(1),
  ("shadow_samples",) + _at_least(1),
  ("max_layers",) + _at_least(1),
  ("path_samples",) + _at_least(1),
  ("max_depth",) + _at_least(0),
  ("rr_depth",) + _at_least(0),
  ("denoise_iterations",) + _at_least(0),
  ("shadow_light_fraction", lambda v: 0.0 <= v <= 1.0, "must be in [0, 1], got {v}"),
  ("denoiser", lambda v: v in ("atrous", "learned"), 'must be "atrous" or "learned", got {v!r}'),
  ("path_kernel", lambda v: v in ("auto", "wavefront", "megakernel"), 'must be "auto", "wavefront" or "megakernel", got {v!r}'),
  ("mesh_bvh_frame", lambda v: v in _MESH_BVH_FRAMES, 'must be "auto" or "mujoco", got {v!r}'),
  ("mesh_bvh_constructor", lambda v: v in _MESH_BVH_CONSTRUCTORS, f"must be one of {_MESH_BVH_CONSTRUCTORS}, got {{v!r}}"),
  ("traversal", lambda v: v in _TRAVERSALS, 'must be "software", "optix" or "auto", got {v!r}'),
  ("block_dim", lambda v: v is None or v >= 1, "must be at least 1, got {v}"),
  ("max_registers", lambda v: v is None or v >= 0, "must be at least 0, got {v}"),
  ("nworld", lambda v: v >= 1, "must be at least 1, got {v}"),
)


def _rendered_mesh_triangles(mjm, rendered) -> int:
  """Triangles of the rendered mesh geoms, counted once a geom (instances of one mesh add up)."""
  ids = np.asarray(rendered)
  if ids.dtype == bool:
    ids = np.nonzero(ids)[0]
  mesh = ids[mjm.geom_type[ids] == mdl.GEOM_MESH]
  dataid = mjm.geom_dataid[mesh]
  dataid = dataid[dataid >= 0]
  return int(np.asarray(mjm.mesh_facenum, np.int64)[dataid].sum()) if dataid.size else 0


def _check_arguments(args: dict) -> None:
  for name, ok, message in _ARGUMENT_RULES:
    value = args[name]
    if not ok(value):
      raise ValueError(f"{name} " + message.format(v=value))


def _check_default_material(default_material) -> None:
  """Keys among metallic / roughness, values in [0, 1]."""
  for key, value in dict(default_material or {}).items():
    if key not in _MATERIAL_KEYS:
      raise ValueError(f"default_material keys must be among {_MATERIAL_KEYS}, got {key!r}")
    if not 0.0 <= value <= 1.0:
      raise ValueError(f"default_material[{key!r}] must be in [0, 1], got {value}")


def _envmap_images(envmap) -> list[np.ndarray]:
  """The radiance images of `envmap`: None, one map or a list of maps, each a path to an RGBE
  file or a (height, width, 3) array (a fourth channel is ignored)."""
  if envmap is None:
    return []
  maps = list(envmap) if isinstance(envmap, (list, tuple)) else [envmap]

  def image(i, item):
    if isinstance(item, (str, bytes, os.PathLike)):
      return envmap_lib.load_hdr(os.fsdecode(item))
    arr = np.asarray(item)
    if arr.ndim == 3 and arr.shape[2] in (3, 4) and min(arr.shape[:2]) >= 1:
      return arr
    raise ValueError(f"envmap[{i}] must be a path or a (height, width, 3) radiance array, got shape {arr.shape}")

  return [image(i, item) for i, item in enumerate(maps)]


def _per_camera(value, n: int, name: str, default: bool) -> np.ndarray:
  if value is None:
    return np.full(n, default, bool)
  if isinstance(value, (bool, np.bool_)):
    return np.full(n, bool(value), bool)
  value = list(value)
  if len(value) != n:
    raise ValueError(f"{name} must have one entry per rendered camera ({n}), got {len(value)}")
  return np.array(value, bool)


def _envmap_images(envmap) -> list[np.ndarray]:
  if envmap is None:
    return []
  items = envmap if isinstance(envmap, (list, tuple)) else [envmap]
  images = []
  for i, item in enumerate(items):
    if isinstance(item, (str, bytes, os.PathLike)):
      images.append(envmap_lib.load_hdr(os.fsdecode(item)))
      continue
    image = np.asarray(item)
    if image.ndim != 3 or image.shape[2] not in (3, 4) or image.shape[0] < 1 or image.shape[1] < 1:
      raise ValueError(f"envmap[{i}] must be a path or a (height, width, 3) radiance array, got shape {image.shape}")
    images.append(image)
  return images


def create_render_context(
  mjm,
  nworld: int = 1,
  cam_res=None,
  render_rgb=None,
  render_depth=None,
  render_seg=None,
  enabled_geom_groups: Sequence[int] = (0, 1, 2),
  cam_active=None,
  background_color=(0.0, 0.0, 0.0, 1.0),
  use_fast_math: bool = True,
  use_shadows: bool = True,
  shadow_light_fraction: float = 0.0,
  enable_backface_culling: bool = True,
  enable_vertex_normals: bool = True,
  light_model=LightModel.PHYSICAL,
  tone_mapper=ToneMapper.PBR_NEUTRAL,
  ev100: float = DEFAULT_EV100,
  unset_material=UnsetMaterial.GLTF,
  default_material: dict | None = None,
  use_textures: bool = True,
  samples_per_pixel: int = 1,
  specular_aa: bool = True,
  shadow_samples: int = 1,
  max_layers: int = 4,
  enable_reflections: bool = True,
  envmap=None,
  envmap_intensity: float | None = None,
  envmap_rotation: float = 0.0,
  envmap_background: bool = True,
  use_skybox: bool = True,
  render_order=RenderOrder.TILED,
  block_dim: int | None = None,
  max_registers: int | None = None,
  mesh_bvh_constructor: str = "cubql",
  mesh_bvh_frame: str = "auto",
  ordered_traversal: bool | None = None,
  tier=Tier.REALTIME,
  path_samples: int = 64,
  max_depth: int = 4,
  rr_depth: int = 3,
  denoise: bool = False,
  denoise_iterations: int = 3,
  render_hdr: bool = False,
  traversal: str = "software",
  render_guides: bool = False,
  camera_rays=None,
  device="cuda",
  precompute: bool = True,
  denoiser: str = "atrous",
  denoiser_weights: str | None = None,
  path_kernel: str = "auto",
) -> PbrRenderContext:
  """Creates a PBR render context on `device` (reference io.py:331-634; same arguments).

  The model is a `mujoco.MjModel` (its poses from mj_forward at qpos0, as the reference
  computes them) or a `HostModel`. Differences from the reference: `render_rgb/depth/seg=None`
  render all three outputs; `mesh_bvh_frame="auto"` re-frames elongated meshes as the reference
  does, but every BVH is this renderer's median-split tree, so `mesh_bvh_constructor`,
  `ordered_traversal`, `block_dim`, `max_registers` and `use_fast_math` are accepted for
  compatibility only: traversal is always nearest-first.
  `traversal="optix"` is not available (no OptiX on this hardware) and falls back to
  "software" with a warning, as the reference does without OptiX.

  `denoiser` ("atrous" or "learned") picks the filter `denoise=True` applies: the reference's
  a-trous filter, or the learned U-Net (weights `denoiser_weights`, default the packaged ones)
  with the a-trous filter as fallback for worlds outside its training range. The learned
  denoiser needs camera sides that are multiples of 64.

  `path_kernel` picks the path tracer's kernel: "wavefront" (a persistent kernel that refills
  lanes as paths end), "megakernel" (one thread per pixel, the reference's scheme) or "auto"
  (the default: the wavefront kernel where it measured faster, everywhere but at one sample a
  pixel on scenes of more than 65,536 rendered mesh triangles).
  Both give identical images; contexts with
  more cameras than one launch takes, and renders of hundreds of samples a pixel (sample regions
  beyond 2 GiB), use the megakernel.

  `precompute=False` leaves the GPU-built tables (DFG LUT, texture and radiance mip chains,
  prefiltered environment stacks, SH9 coefficients, sampling tables) zero; it exists so host-side
  test harnesses can create contexts on device="cpu" and fill the tables themselves.

  Raises:
    ValueError: for invalid arguments, or a model with nothing to render.
  """
  if precompute and torch.device(device).type != "cuda":
    raise ValueError("create_render_context builds its tables with HIP kernels: use a GPU device (or precompute=False)")
  mjm = as_host_model(mjm)
  light_model = _as_enum(LightModel, light_model, "light_model")
  tone_mapper = _as_enum(ToneMapper, tone_mapper, "tone_mapper")
  unset_material = _as_enum(UnsetMaterial, unset_material, "unset_material")
  render_order = _as_enum(RenderOrder, render_order, "render_order")
  tier = _as_enum(Tier, tier, "tier")
  _check_arguments(locals())
  if tier == Tier.PATH and max_depth + 4 * max_layers >= MAX_PATH_VERTICES:
    raise ValueError(
      f"max_depth + 4 * max_layers must be below {MAX_PATH_VERTICES} for the path tracer's random "
      f"number layout, got {max_depth} + 4 * {max_layers}"
    )
  _check_default_material(default_material)
  if denoise and tier != Tier.PATH:
    raise ValueError("denoise applies to the path-traced tier (Tier.PATH)")
  if traversal == "optix":
    warnings.warn('traversal="optix" falls back to "software": no OptiX on this device', stacklevel=2)
  traversal = "software"
  mjm.validate()

  image_lights = np.nonzero(mjm.light_type == mdl.LIGHT_IMAGE)[0] if mjm.nlight else []
  if len(image_lights):
    warnings.warn(f"image lights are not supported and are ignored: {list(image_lights)}", stacklevel=2)
  rendered = np.isin(mjm.geom_group, list(enabled_geom_groups))
  if not np.any(rendered):
    raise ValueError("the model has no geoms in the enabled geom groups to render")

  # cameras
  if cam_active is None:
    cams = list(range(mjm.ncam))
  else:
    if len(cam_active) and isinstance(cam_active[0], (bool, np.bool_)):
      cams = [i for i, a in enumerate(cam_active) if a]
    else:
      cams = [int(c) for c in cam_active]
  if not cams:
    raise ValueError("the model has no active camera to render")
  nrender = len(cams)
  if cam_res is None:
    res = [tuple(int(x) for x in mjm.cam_resolution[c]) for c in cams]
  elif isinstance(cam_res, tuple) and len(cam_res) == 2 and all(isinstance(x, (int, np.integer)) for x in cam_res):
    res = [tuple(cam_res)] * nrender
  else:
    res = [tuple(r) for r in cam_res]
  if len(res) != nrender:
    raise ValueError(f"cam_res must give one resolution per rendered camera ({nrender}), got {len(res)}")
  for r in res:
    if r[0] < 1 or r[1] < 1:
      raise ValueError(f"camera resolutions must be positive, got {r}")
    if denoise and denoiser == "learned" and (r[0] % 64 or r[1] % 64):
      raise ValueError(f"the learned denoiser needs camera sides that are multiples of 64, got {r}")
  want_rgb = _per_camera(render_rgb, nrender, "render_rgb", True)
  want_depth = _per_camera(render_depth, nrender, "render_depth", True)
  want_seg = _per_camera(render_seg, nrender, "render_seg", True)
  npix = [w * h for w, h in res]

  def adrs(flags):
    a, off = [], 0
    for f, n in zip(flags, npix):
      a.append(off if f else -1)
      off += n if f else 0
    return np.array(a, np.int32), max(off, 1)

  rgb_adr, nrgb = adrs(want_rgb)
  depth_adr, ndepth = adrs(want_depth)
  seg_adr, nseg = adrs(want_seg)

  dev = torch.device(device)
  enabled = np.nonzero(rendered)[0].astype(np.int32)
  # Ground planes are tested by every ray before the scene BVH instead of sitting in it, where a
  # plane's box (unbounded, or a floor as wide as the scene) would cover every ancestor node:
  # planes of unbounded extent (a nonpositive half-size) always, and every plane when the model
  # has at most MAX_DIRECT_PLANES of them (a plane test costs about one box test).
  gsize = np.asarray(mjm.geom_size, np.float32).reshape(-1, 3)
  planes = mjm.geom_type[enabled] == mdl.GEOM_PLANE
  direct = planes & ((gsize[enabled, 0] <= 0) | (gsize[enabled, 1] <= 0))
  if np.count_nonzero(planes) <= MAX_DIRECT_PLANES:
    direct = planes.copy()
  if np.all(direct):
    direct[:] = False
  bvh_ids, direct_ids = enabled[~direct], enabled[direct]
  depth = bvh_lib.scene_depth(len(bvh_ids))
  tlas_box, tlas_leaf = bvh_lib.alloc_scene_bvh(nworld, depth, dev)
  frames = bvh_lib.select_mesh_frames(mjm) if mesh_bvh_frame == "auto" else {}
  meshes = bvh_lib.build_mesh_bvhs(mjm, dev, precompute=precompute, frames=frames)

  base = RenderBase(
    nworld=nworld,
    nrender=nrender,
    cam_id_map=np.array(cams, np.int32),
    cam_res=np.array(res, np.int32),
    render_rgb=want_rgb,
    render_depth=want_depth,
    render_seg=want_seg,
    rgb_adr=rgb_adr,
    depth_adr=depth_adr,
    seg_adr=seg_adr,
    total_rays=int(sum(npix)),
    rgb_data=torch.zeros((nworld, nrgb), dtype=torch.int32, device=dev),
    depth_data=torch.zeros((nworld, ndepth), dtype=torch.float32, device=dev),
    seg_data=torch.zeros((nworld, nseg, 2), dtype=torch.int32, device=dev),
    znear=0.01,
    zfar=0.0,
    enabled_geom_ids=torch.as_tensor(enabled).to(dev),
    tlas_depth=depth,
    tlas_box=tlas_box,
    tlas_leaf=tlas_leaf,
    blas_depth=torch.as_tensor(meshes.depth).to(dev),
    blas_node_adr=torch.as_tensor(meshes.node_adr).to(dev),
    blas_tri_adr=torch.as_tensor(meshes.tri_adr).to(dev),
    blas_box=meshes.box,
    blas_tri=meshes.tri,
    mesh_aabb=torch.as_tensor(meshes.aabb).to(dev),
    enable_backface_culling=enable_backface_culling,
    enable_vertex_normals=enable_vertex_normals,
    use_fast_math=use_fast_math,
    headlight_active=bool(mjm.headlight_active),
    headlight_diffuse=tuple(float(x) for x in mjm.headlight_diffuse),
    headlight_specular=tuple(float(x) for x in mjm.headlight_specular),
    has_spot_lights=bool(mjm.nlight and np.any(mjm.light_type == mdl.LIGHT_SPOT)),
    has_meshes=bool(np.any(np.isin(mjm.geom_type[rendered], (mdl.GEOM_MESH, mdl.GEOM_SDF, mdl.GEOM_HFIELD)))),
    mesh_triangles=_rendered_mesh_triangles(mjm, rendered),
  )

  base.bvh_geom_ids = torch.as_tensor(bvh_ids).to(dev)
  base.mesh_frame, base.trace_mesh_vert, base.trace_mesh_normal = meshes.frame, meshes.vert, meshes.normal
  base.mesh_bvhs = meshes
  base.direct_geom_ids = torch.as_tensor(direct_ids).to(dev)
  base.world_stage = torch.zeros((nworld, world_stage_floats(depth, mjm.ngeom, mjm.nlight, len(direct_ids))),
                                 dtype=torch.float32, device=dev)

  physical = light_model == LightModel.PHYSICAL
  lighting = material.resolve_lighting(mjm, light_model, ev100)
  surf = _resolve_surfaces(mjm, rendered, unset_material, default_material, use_textures, specular_aa and tier == Tier.REALTIME,
                           enable_reflections)
  metallic, roughness, reflectance, features = surf.metallic, surf.roughness, surf.reflectance, surf.features
  has_transparency, has_reflection = surf.has_transparency, surf.has_reflection
  textures = texture.upload_textures(mjm, surf.used_textures, dev, precompute=precompute)
  background_linear = tuple(float(x) for x in material.srgb_to_linear(background_color[:3]))
  env, has_env, any_env = _resolve_environment(mjm, envmap, use_skybox, physical, dev, precompute)
  if envmap_intensity is None:
    envmap_intensity = 1.0 / lighting.exposure
  if has_env and physical:
    lighting.ambient[:] = 0.0  # image-based lighting replaces the headlight's ambient term

  dfg = torch.zeros((DFG_LUT_SIZE, DFG_LUT_SIZE, 2), dtype=torch.float32, device=dev)
  if precompute:
    _abi.call("mjpbr_dfg_lut", _abi.ptr(dfg), DFG_LUT_SIZE, DFG_LUT_SAMPLES, envmap_lib._stream(dev))

  npixel = int(nrgb)
  aov = npixel if denoise else 1

  def per_world(values, dtype=torch.float32):
    return torch.as_tensor(np.asarray(values, dtype=np.float32).reshape(1, -1)).to(dev)

  def zeros(shape, dtype=torch.float32):
    return torch.zeros(shape, dtype=dtype, device=dev)

  rc = PbrRenderContext(
    base=base,
    light_model=int(light_model),
    tone_mapper=int(tone_mapper),
    use_shadows=use_shadows,
    shadow_light_fraction=shadow_light_fraction,
    geom_metallic=per_world(metallic),
    geom_roughness=per_world(roughness),
    geom_reflectance=per_world(reflectance),
    mat_reflectance=per_world(mjm.mat_reflectance if mjm.nmat else np.zeros(1)),
    light_intensity=per_world(lighting.light_intensity if mjm.nlight else np.zeros(1)),
    light_range=per_world(mjm.light_range if mjm.nlight else np.zeros(1)),
    light_bulbradius=per_world(mjm.light_bulbradius if mjm.nlight else np.zeros(1)),
    headlight_intensity=lighting.headlight_intensity,
    exposure=torch.tensor([lighting.exposure], dtype=torch.float32, device=dev),
    ambient_radiance=torch.as_tensor(lighting.ambient.reshape(1, 3).astype(np.float32)).to(dev),
    dfg_lut=dfg,
    background_linear=background_linear,
    has_env=has_env,
    env_background=any_env and envmap_background,
    env=env,
    env_texel_angle=envmap_lib.texel_angle(env.width),
    env_id=torch.zeros(1, dtype=torch.int32, device=dev),
    env_intensity=torch.tensor([envmap_intensity], dtype=torch.float32, device=dev),
    env_yaw=torch.tensor([math.radians(envmap_rotation)], dtype=torch.float32, device=dev),
    textures=textures,
    features=features,
    samples_per_pixel=samples_per_pixel,
    shadow_samples=shadow_samples,
    has_transparency=has_transparency,
    max_layers=max_layers,
    has_reflection=has_reflection,
    render_order=int(render_order),
    block_dim=block_dim,
    max_registers=max_registers or 0,
    ordered_traversal=True,
    tier=int(tier),
    path_samples=path_samples,
    max_depth=max_depth,
    rr_depth=rr_depth,
    denoise=denoise,
    denoise_iterations=denoise_iterations,
    aov_albedo=zeros((nworld, aov, 3), torch.int16),
    aov_normal=zeros((nworld, aov, 3), torch.int16),
    aov_depth=zeros((nworld, aov)),
    aov_emission=zeros((nworld, aov, 3)),
    aov_variance=zeros((nworld, aov)),
    denoise_buffers=(zeros((nworld, aov, 4)), zeros((nworld, aov, 4)), zeros((nworld, aov, 4))),
    render_hdr=render_hdr,
    hdr_data=zeros((nworld, npixel if render_hdr or denoise else 1, 3)),
    render_guides=render_guides,
    guide_albedo=zeros((nworld, npixel if render_guides else 1, 3)),
    guide_normal=zeros((nworld, npixel if render_guides else 1, 3)),
    traversal=traversal,
    denoiser=denoiser,
    denoiser_weights=denoiser_weights,
    path_kernel=path_kernel,
  )
  if camera_rays is not None:
    from mujoco_warp_pbr_amd.rays import set_camera_rays

    set_camera_rays(rc, camera_rays)
  return rc


# Planes tested before the scene BVH when a model has at most this many (see create_render_context).
MAX_DIRECT_PLANES = 4


def world_stage_floats(tlas_depth: int, ngeom: int, nlight: int, ndirect: int = 0) -> int:
  """Floats one world takes in the per-world record buffer (the library's layout)."""
  S = _abi.structs()
  sc = S["MjpbrScene"]()
  sc.geo.tlas_depth = tlas_depth
  sc.geo.ngeom = ngeom
  sc.geo.ndirect = ndirect
  sc.lit.nlight = nlight
  n = _abi.library().mjpbr_world_stage_floats(_abi.byref(sc))
  if n < 0:
    raise _abi.MjpbrError(_abi.library().mjpbr_last_error().decode())
  return int(n)


def build_scene_bvh(m, d, rc) -> None:
  """Builds the per-world scene BVH (leaf assignment and boxes) from the current poses."""
  bvh_lib.refit_bvh(m, d, rc, rebuild=True)


@dataclasses.dataclass
class _Surfaces:
  metallic: np.ndarray
  roughness: np.ndarray
  reflectance: np.ndarray
  features: material.SurfaceFeatures
  used_textures: np.ndarray  # (max(ntex, 1),) bool
  has_transparency: bool
  has_reflection: bool


def _resolve_surfaces(mjm, rendered, unset_material, default_material, use_textures, specular_aa, enable_reflections) -> _Surfaces:
  """Per-geom PBR parameters, the texture roles and textures the rendered materials use, and
  whether translucency (an rgba or alpha texture below 1) or mirrors appear at all."""
  metallic, roughness, reflectance = material.resolve_geom_materials(mjm, unset_material, default_material)
  mats = np.unique(mjm.geom_matid[rendered & (mjm.geom_matid >= 0)])
  features = material.SurfaceFeatures(mjm, use_textures, specular_aa, mats)
  used = np.zeros(max(mjm.ntex, 1), dtype=bool)
  if features.textures:
    ids = np.asarray(mjm.mat_texid)[mats].ravel()
    used[ids[ids >= 0]] = True
  alpha = material.effective_geom_rgba(mjm)[rendered, 3]
  translucent = bool(np.any((alpha > 0.0) & (alpha < 1.0))) or features.alpha_map
  mirrors = enable_reflections and bool(mjm.nmat and np.any(np.asarray(mjm.mat_reflectance)[mats] > 0.0))
  return _Surfaces(metallic, roughness, reflectance, features, used, translucent, mirrors)


def _resolve_environment(mjm, envmap, use_skybox: bool, physical: bool, dev, precompute: bool):
  """(environments, has_env, any_env) from the explicit maps, else the model's skybox, else
  none (constant ambient) (reference io.py:267-285). A skybox lights the scene only under the
  physical light model; with MuJoCo's lights it is background only (has_env False)."""
  images = _envmap_images(envmap)
  lights_scene = True
  if not images and use_skybox:
    sky = envmap_lib.skybox_image(mjm)
    if sky is not None:
      images, lights_scene = [sky], physical
  if not images:
    return envmap_lib.empty_environments(dev), False, False
  return envmap_lib.build_environments(images, dev, precompute=precompute), lights_scene, True
