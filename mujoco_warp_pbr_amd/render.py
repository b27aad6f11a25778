"""render(): marshals the context into the C ABI structs and launches the HIP megakernels
(reference render.py:430-460, pathtrace.py:426-438, launch.py:89-126, kernel_inputs.py).

Every pointer handed to libmjpbr_hip.so is a tensor's data_ptr(); work is enqueued on the
current torch stream of the data's device, so render() performs no host synchronisation.
"""

from __future__ import annotations

import ctypes
import functools
import warnings

import numpy as np
import torch

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd.types import PbrRenderContext
from mujoco_warp_pbr_amd.types import Tier

NTEXROLE = 10


def _stream_of(t: torch.Tensor) -> int:
  if t.is_cuda:
    return int(torch.cuda.current_stream(t.device).cuda_stream)
  return 0


def _n(t: torch.Tensor) -> int:
  return int(t.shape[0])


def _check(t: torch.Tensor, dtype, name: str):
  if t.dtype != dtype:
    raise ValueError(f"{name} must have dtype {dtype}, got {t.dtype}")
  if not t.is_contiguous():
    raise ValueError(f"{name} must be contiguous")
  return t


def geometry_struct(m, d, rc: PbrRenderContext):
  """The MjpbrGeometry of the current frame. Every entry point builds it, so the check that the
  data's world count matches the context's buffers (sized for rc.base.nworld) sits here."""
  base = rc.base
  if d.nworld != base.nworld:
    raise ValueError(f"data has {d.nworld} worlds, the context {base.nworld}")
  S = _abi.structs()
  g = S["MjpbrGeometry"]()
  g.nworld = d.nworld
  g.ngeom = m.ngeom
  g.nmesh = m.nmesh
  g.geom_type = _abi.ptr(_check(m.geom_type, torch.int32, "geom_type"))
  g.geom_dataid = _abi.ptr(_check(m.geom_dataid, torch.int32, "geom_dataid"))
  g.geom_matid = _abi.ptr(_check(m.geom_matid, torch.int32, "geom_matid"))
  g.geom_size = _abi.ptr(_check(m.geom_size, torch.float32, "geom_size"))
  g.geom_rgba = _abi.ptr(_check(m.geom_rgba, torch.float32, "geom_rgba"))
  g.mat_rgba = _abi.ptr(_check(m.mat_rgba, torch.float32, "mat_rgba"))
  g.geom_xpos = _abi.ptr(_check(d.geom_xpos, torch.float32, "geom_xpos"))
  g.geom_xmat = _abi.ptr(_check(d.geom_xmat, torch.float32, "geom_xmat"))
  g.geom_dataid_n = _n(m.geom_dataid)
  g.geom_matid_n = _n(m.geom_matid)
  g.geom_size_n = _n(m.geom_size)
  g.geom_rgba_n = _n(m.geom_rgba)
  g.mat_rgba_n = _n(m.mat_rgba)
  g.nmat = m.nmat
  g.tlas_depth = base.tlas_depth
  bvh_ids = base.enabled_geom_ids if base.bvh_geom_ids is None else base.bvh_geom_ids
  g.nbvh = int(bvh_ids.shape[0])
  g.tlas_box = _abi.ptr(base.tlas_box)
  g.tlas_leaf = _abi.ptr(base.tlas_leaf)
  g.bvh_geom = _abi.ptr(bvh_ids)
  if base.direct_geom_ids is not None and base.direct_geom_ids.numel():
    g.direct_geom = _abi.ptr(base.direct_geom_ids)
    g.ndirect = int(base.direct_geom_ids.shape[0])
  g.blas_depth = _abi.ptr(base.blas_depth)
  g.blas_node_adr = _abi.ptr(base.blas_node_adr)
  g.blas_tri_adr = _abi.ptr(base.blas_tri_adr)
  g.blas_box = _abi.ptr(base.blas_box)
  g.blas_tri = _abi.ptr(base.blas_tri)
  g.mesh_aabb = _abi.ptr(base.mesh_aabb)
  g.world_stage = _abi.ptr(base.world_stage)
  g.mesh_frame = _abi.ptr(base.mesh_frame)
  return g


def camera_chunks(rc: PbrRenderContext) -> list[tuple[int, int]]:
  """The rendering cameras as launches of at most MJPBR_MAX_CAMS consecutive ones: [(first, end)]."""
  n, step = int(rc.base.nrender), _max_cams()
  return [(a, min(a + step, n)) for a in range(0, n, step)]


@functools.lru_cache(maxsize=None)
def _max_cams() -> int:
  """MJPBR_MAX_CAMS of include/mjpbr.h (read once: render() runs on the hot path)."""
  return _abi._defines(_abi._header_text())["MJPBR_MAX_CAMS"]


def scene_struct(m, d, rc: PbrRenderContext, cams: tuple[int, int] | None = None):
  """The MjpbrScene of the current frame; `cams` = (first, end) limits its cameras to one launch's
  (camera_chunks), all of them by default."""
  S = _abi.structs()
  sc = S["MjpbrScene"]()
  sc.geo = geometry_struct(m, d, rc)
  # materials
  base = rc.base
  mt = sc.mats
  mt.geom_metallic = _abi.ptr(_check(rc.geom_metallic, torch.float32, "geom_metallic"))
  mt.geom_roughness = _abi.ptr(_check(rc.geom_roughness, torch.float32, "geom_roughness"))
  mt.geom_reflectance = _abi.ptr(_check(rc.geom_reflectance, torch.float32, "geom_reflectance"))
  mt.mat_reflectance = _abi.ptr(_check(rc.mat_reflectance, torch.float32, "mat_reflectance"))
  mt.mat_emission = _abi.ptr(_check(m.mat_emission, torch.float32, "mat_emission"))
  mt.mat_texid = _abi.ptr(_check(m.mat_texid, torch.int32, "mat_texid"))
  mt.mat_texuniform = _abi.ptr(_check(m.mat_texuniform, torch.int32, "mat_texuniform"))
  mt.mat_texrepeat = _abi.ptr(_check(m.mat_texrepeat, torch.float32, "mat_texrepeat"))
  mt.geom_metallic_n = _n(rc.geom_metallic)
  mt.geom_roughness_n = _n(rc.geom_roughness)
  mt.geom_reflectance_n = _n(rc.geom_reflectance)
  mt.mat_reflectance_n = _n(rc.mat_reflectance)
  mt.mat_emission_n = _n(m.mat_emission)
  mt.mat_texid_n = _n(m.mat_texid)
  mt.mat_texuniform_n = _n(m.mat_texuniform)
  mt.mat_texrepeat_n = _n(m.mat_texrepeat)
  for name in (
    "mesh_vertadr",
    "mesh_faceadr",
    "mesh_facenum",
    "mesh_normaladr",
    "mesh_texcoordadr",
    "mesh_vert",
    "mesh_normal",
    "mesh_face",
    "mesh_facenormal",
    "mesh_facetexcoord",
    "mesh_texcoord",
  ):
    setattr(mt, name, _abi.ptr(getattr(m, name)))
  if base.trace_mesh_vert is not None:  # re-framed mesh BVHs: shade from the traced-frame copies
    mt.mesh_vert = _abi.ptr(base.trace_mesh_vert)
    mt.mesh_normal = _abi.ptr(base.trace_mesh_normal)
  tex = rc.textures
  mt.ntex = tex.ntex
  mt.tex_level_adr = _abi.ptr(tex.level_adr)
  mt.tex_desc = _abi.ptr(tex.desc)
  mt.tex_data = _abi.ptr(tex.data)
  # lights
  lt = sc.lit
  lt.nlight = m.nlight
  for name in ("light_type", "light_castshadow", "light_active"):
    setattr(lt, name, _abi.ptr(_check(getattr(m, name), torch.int32, name)))
    setattr(lt, name + "_n", _n(getattr(m, name)))
  for name in ("light_attenuation", "light_cutoff", "light_exponent", "light_diffuse", "light_specular"):
    setattr(lt, name, _abi.ptr(_check(getattr(m, name), torch.float32, name)))
    setattr(lt, name + "_n", _n(getattr(m, name)))
  for name in ("light_intensity", "light_range", "light_bulbradius"):
    t = _check(getattr(rc, name), torch.float32, name)
    setattr(lt, name, _abi.ptr(t))
    setattr(lt, name + "_n", _n(t))
  lt.light_xpos = _abi.ptr(_check(d.light_xpos, torch.float32, "light_xpos"))
  lt.light_xdir = _abi.ptr(_check(d.light_xdir, torch.float32, "light_xdir"))
  # cameras
  base = rc.base
  cm = sc.cams
  first, end = _chunk(rc, cams)
  cm.nrender = end - first
  cm.ncam = m.ncam
  cm.cam_fovy_n = _n(m.cam_fovy)
  cm.cam_intrinsic_n = _n(m.cam_intrinsic)
  cm.total_rays = base.total_rays
  cm.ray_base = sum(int(base.cam_res[i][0]) * int(base.cam_res[i][1]) for i in range(first))
  for k, i in enumerate(range(first, end)):
    cm.cam_res[k][0] = int(base.cam_res[i][0])
    cm.cam_res[k][1] = int(base.cam_res[i][1])
    cm.cam_id_map[k] = int(base.cam_id_map[i])
  cm.cam_projection = _abi.ptr(m.cam_projection)
  cm.cam_fovy = _abi.ptr(_check(m.cam_fovy, torch.float32, "cam_fovy"))
  cm.cam_sensorsize = _abi.ptr(m.cam_sensorsize)
  cm.cam_intrinsic = _abi.ptr(m.cam_intrinsic)
  cm.cam_xpos = _abi.ptr(_check(d.cam_xpos, torch.float32, "cam_xpos"))
  cm.cam_xmat = _abi.ptr(_check(d.cam_xmat, torch.float32, "cam_xmat"))
  cm.ray_origin = _abi.ptr(rc.ray_origin)
  cm.ray_direction = _abi.ptr(rc.ray_direction)
  # environment
  ev = sc.env
  env = rc.env
  ev.exposure = _abi.ptr(_check(rc.exposure, torch.float32, "exposure"))
  ev.ambient = _abi.ptr(_check(rc.ambient_radiance, torch.float32, "ambient_radiance"))
  ev.exposure_n = _n(rc.exposure)
  ev.ambient_n = _n(rc.ambient_radiance)
  for k in range(3):
    ev.background[k] = float(rc.background_linear[k])
  ev.dfg_size = int(rc.dfg_lut.shape[0])
  ev.dfg_lut = _abi.ptr(rc.dfg_lut)
  ev.nenv = env.nenv
  ev.env_id_n = _n(rc.env_id)
  ev.env_intensity_n = _n(rc.env_intensity)
  ev.env_yaw_n = _n(rc.env_yaw)
  ev.env_level_adr = _abi.ptr(env.level_adr)
  ev.env_desc = _abi.ptr(env.desc)
  ev.env_radiance = _abi.ptr(env.radiance)
  ev.pf_levels, ev.pf_height, ev.pf_width = (int(x) for x in env.prefiltered.shape[1:4])
  ev.env_prefiltered = _abi.ptr(env.prefiltered)
  ev.env_sh = _abi.ptr(env.sh)
  ev.env_id = _abi.ptr(_check(rc.env_id, torch.int32, "env_id"))
  ev.env_intensity = _abi.ptr(_check(rc.env_intensity, torch.float32, "env_intensity"))
  ev.env_yaw = _abi.ptr(_check(rc.env_yaw, torch.float32, "env_yaw"))
  ev.cdf_h, ev.cdf_w = int(env.prob.shape[1]), int(env.prob.shape[2])
  ev.env_prob = _abi.ptr(env.prob)
  ev.env_marginal = _abi.ptr(env.marginal)
  ev.env_conditional = _abi.ptr(env.conditional)
  return sc


def settings_struct(rc: PbrRenderContext):
  """MjpbrSettings of the context (reference kernel_inputs.py:281-318 `settings_values`)."""
  S = _abi.structs()
  s = S["MjpbrSettings"]()
  base = rc.base
  f = rc.features
  s.render_order = int(rc.render_order)
  s.samples_per_axis = int(rc.samples_per_pixel)
  s.znear = float(base.znear)
  s.zfar = float(base.zfar)
  s.tone_mapper = int(rc.tone_mapper)
  s.render_hdr = int(bool(rc.render_hdr))
  s.any_rgb = int(bool(np.any(base.render_rgb)))
  s.light_model = int(rc.light_model)
  s.has_spot = int(bool(base.has_spot_lights))
  s.use_shadows = int(bool(rc.use_shadows))
  s.shadow_samples = int(rc.shadow_samples)
  s.shadow_light_fraction = float(rc.shadow_light_fraction)
  s.headlight_active = int(bool(base.headlight_active))
  for k in range(3):
    s.headlight_diffuse[k] = float(base.headlight_diffuse[k])
    s.headlight_specular[k] = float(base.headlight_specular[k])
  s.headlight_intensity = float(rc.headlight_intensity)
  s.has_env = int(bool(rc.has_env))
  s.env_background = int(bool(rc.env_background))
  s.inv_texel_angle = 1.0 / float(rc.env_texel_angle)
  s.textures = int(bool(f.textures))
  s.base_color_map = int(bool(f.base_color_map))
  s.normal_map = int(bool(f.normal_map))
  s.orm_map = int(bool(f.orm_map))
  s.emissive_map = int(bool(f.emissive_map))
  s.opacity_map = int(bool(f.opacity_map))
  s.specular_aa = int(bool(f.specular_aa))
  s.vertex_normals = int(bool(base.enable_vertex_normals))
  s.has_reflection = int(bool(rc.has_reflection))
  s.has_transparency = int(bool(rc.has_transparency))
  s.max_layers = int(rc.max_layers)
  s.denoise = int(bool(rc.denoise))
  s.cull_backfaces = int(bool(base.enable_backface_culling))
  s.render_guides = int(bool(rc.render_guides))
  s.path_samples = int(rc.path_samples)
  s.max_depth = int(rc.max_depth)
  s.rr_depth = int(rc.rr_depth)
  s.has_meshes = int(bool(base.has_meshes))
  s.mesh_triangles = min(int(base.mesh_triangles), 2**31 - 1)
  return s


def _check_output_buffers(rc: PbrRenderContext) -> None:
  """The kernels write HDR, guides and denoiser features for every RGB pixel when the matching
  flag is set; a flag switched on after the context was created (its buffer then holds one
  pixel a world) is refused here instead of being written past."""
  base = rc.base
  npix = int(base.rgb_data.shape[1])
  need = [("render_hdr" if rc.render_hdr else "denoise", bool(rc.render_hdr) or bool(rc.denoise), rc.hdr_data)]
  need.append(("render_guides", bool(rc.render_guides), rc.guide_albedo))
  need.append(("render_guides", bool(rc.render_guides), rc.guide_normal))
  if rc.tier == Tier.PATH and rc.denoise:
    need += [("denoise", True, t) for t in (rc.aov_albedo, rc.aov_normal, rc.aov_depth, rc.aov_emission, rc.aov_variance)]
  for flag, on, t in need:
    if on and (int(t.shape[0]) < base.nworld or int(t.shape[1]) < npix):
      raise ValueError(f"rc.{flag} is set but its output buffers were not allocated for it: create the context with {flag}=True")


def _chunk(rc: PbrRenderContext, cams: tuple[int, int] | None) -> tuple[int, int]:
  n = int(rc.base.nrender)
  first, end = (0, n) if cams is None else (int(cams[0]), int(cams[1]))
  if not 0 <= first < end <= n or end - first > _max_cams():
    raise ValueError(f"a launch renders 1 to {_max_cams()} of the context's {n} cameras, got [{first}, {end})")
  return first, end


def outputs_struct(rc: PbrRenderContext, cams: tuple[int, int] | None = None):
  """The MjpbrOutputs of the context; `cams` as in scene_struct."""
  _check_output_buffers(rc)
  S = _abi.structs()
  o = S["MjpbrOutputs"]()
  base = rc.base
  first, end = _chunk(rc, cams)
  for k, i in enumerate(range(first, end)):
    o.render_rgb[k] = int(bool(base.render_rgb[i]))
    o.render_depth[k] = int(bool(base.render_depth[i]))
    o.render_seg[k] = int(bool(base.render_seg[i]))
    o.rgb_adr[k] = int(base.rgb_adr[i])
    o.depth_adr[k] = int(base.depth_adr[i])
    o.seg_adr[k] = int(base.seg_adr[i])
  o.rgb_stride = int(base.rgb_data.shape[1])
  o.depth_stride = int(base.depth_data.shape[1])
  o.seg_stride = int(base.seg_data.shape[1])
  o.hdr_stride = int(rc.hdr_data.shape[1])
  o.guide_stride = int(rc.guide_albedo.shape[1])
  o.aov_stride = int(rc.aov_depth.shape[1])
  o.rgb = _abi.ptr(base.rgb_data)
  o.depth = _abi.ptr(base.depth_data)
  o.seg = _abi.ptr(base.seg_data)
  o.hdr = _abi.ptr(rc.hdr_data)
  o.albedo = _abi.ptr(rc.guide_albedo)
  o.normal = _abi.ptr(rc.guide_normal)
  o.aov_albedo = _abi.ptr(rc.aov_albedo)
  o.aov_normal = _abi.ptr(rc.aov_normal)
  o.aov_depth = _abi.ptr(rc.aov_depth)
  o.aov_emission = _abi.ptr(rc.aov_emission)
  o.aov_variance = _abi.ptr(rc.aov_variance)
  o.ray_counts = _abi.ptr(rc._host_cache.get("ray_counts"))
  return o


RAY_KINDS = ("segments", "light_shadow", "env_shadow")


def count_path_rays(m, d, rc: PbrRenderContext, seed: int = 0) -> dict:
  """Renders once with the path tracer's counting kernel instance (the same image as
  render(m, d, rc, seed)) and returns the rays it traced: path segments (closest-hit rays), light
  shadow rays and environment shadow rays, and their total. Path-traced tier only."""
  if rc.tier != Tier.PATH:
    raise ValueError("count_path_rays needs a path-traced context (tier=Tier.PATH)")
  counts = torch.zeros(3, dtype=torch.int64, device=rc.base.rgb_data.device)
  rc._host_cache["ray_counts"] = counts
  try:
    render(m, d, rc, seed=seed)
  finally:
    rc._host_cache.pop("ray_counts", None)
  out = dict(zip(RAY_KINDS, (int(x) for x in counts.cpu().tolist())))
  out["total"] = sum(out[k] for k in RAY_KINDS)
  return out


def check_env_id(rc: PbrRenderContext) -> None:
  """Raises if a (replaced) env_id selects a missing environment map (kernel_inputs.py:480-490).

  Checked once per env_id tensor (it costs a device-to-host copy); skipped under graph capture.
  """
  env_id = rc.env_id
  if rc._host_cache.get("env_id") is env_id:
    return
  if env_id.is_cuda and torch.cuda.is_current_stream_capturing():
    return
  ids = env_id.cpu().numpy()
  nenv = rc.env.nenv
  if ids.size and (ids.min() < 0 or ids.max() >= nenv):
    raise ValueError(f"env_id must be in [0, {nenv}) (one per environment map), got values in [{ids.min()}, {ids.max()}]")
  rc._host_cache["env_id"] = env_id


def warn_ignored_overrides(m, rc: PbrRenderContext) -> None:
  """Warns once per replaced array when it needs a flag that is off (launch.py:59-87).

  has_transparency and has_reflection are derived from the model when the context is created;
  a replaced m.geom_rgba with translucent alphas, or a replaced rc.mat_reflectance with
  reflective entries, renders as before until the flag is set. The arrays seen by the first
  render are the model's own and are not checked. Skipped under graph capture.
  """
  if m.geom_rgba.is_cuda and torch.cuda.is_current_stream_capturing():
    return
  cache = rc._host_cache
  checks = (
    ("geom_rgba", m.geom_rgba, rc.has_transparency, lambda a: bool(((a[..., 3] > 0.0) & (a[..., 3] < 1.0)).any()), "has_transparency"),
    ("mat_reflectance", rc.mat_reflectance, rc.has_reflection, lambda a: bool((a > 0.0).any()), "has_reflection"),
  )
  for name, array, enabled, needs, flag in checks:
    seen = cache.get(name)
    cache[name] = array
    if seen is None or seen is array or enabled:
      continue
    if needs(array):
      warnings.warn(f"the replaced {name} is ignored by the renderer: set rc.{flag} = True to render it", stacklevel=3)


def render_path_samples(m, d, rc: PbrRenderContext, seed: int = 0) -> None:
  """The path tracer alone: noisy HDR (and, with `rc.denoise`, the first-hit AOVs) into the
  context's buffers, no filter. Training-data generation for the learned denoiser uses it."""
  if rc.tier != Tier.PATH:
    raise ValueError("render_path_samples needs a path-traced context (Tier.PATH)")
  check_env_id(rc)
  warn_ignored_overrides(m, rc)
  _path_trace(m, d, rc, settings_struct(rc), seed, _stream_of(d.geom_xpos))


# rc.path_kernel "auto" takes the wavefront kernel (lanes refill with new paths, and a wave's new
# paths share few pixels) except at 1 sample a pixel on scenes of more than 64k rendered mesh
# triangles. Measured at 1024 worlds (ms, per-pixel kernel vs wavefront): the reference suite's
# scenes at 16 spp + a-trous 97.2 vs 77.5 and 121.7 vs 109.3; the full-feature scene (4k mesh
# triangles) at 4 spp 256^2 275.3 vs 239.0, at 1 spp 128^2 22.7 vs 21.2 (work items of four
# tiles); the mesh-heavy scene (200k) at 4 spp 128^2 192.0 vs 174.8, at 1 spp 256^2 182.9 vs 205.3.
AUTO_MEGAKERNEL_MESH_TRIANGLES = 1 << 16


def _wavefront_pays(rc: PbrRenderContext) -> bool:
  if rc.path_kernel == "wavefront":
    return True
  if rc.path_kernel == "megakernel":
    return False
  return int(rc.path_samples) >= 2 or int(rc.base.mesh_triangles) <= AUTO_MEGAKERNEL_MESH_TRIANGLES


def _path_trace(m, d, rc: PbrRenderContext, s, seed: int, stream: int) -> None:
  """The path tracer's launches of render(): the wavefront kernel (rc.path_kernel, cameras in one
  launch), else the per-pixel megakernel a launch per camera_chunks entry."""
  chunks = camera_chunks(rc)
  if len(chunks) == 1 and _wavefront_pays(rc):
    from mujoco_warp_pbr_amd import wavefront

    if wavefront.render_context(m, d, rc, s, seed, stream):
      return
  for cams in chunks:
    sc = scene_struct(m, d, rc, cams)
    out = outputs_struct(rc, cams)
    _abi.call("mjpbr_pt_render", _abi.byref(sc), _abi.byref(s), _abi.byref(out), ctypes.c_int32(int(seed)), stream)


def render(m, d, rc: PbrRenderContext, seed: int = 0) -> None:
  """Renders the current frame into the context's output buffers (reference render.py:430-460).

  Call `refit_bvh(m, d, rc)` first whenever geometry moved.
  """
  if d.nworld != rc.base.nworld:
    raise ValueError(f"data has {d.nworld} worlds, the context {rc.base.nworld}")
  check_env_id(rc)
  warn_ignored_overrides(m, rc)
  s = settings_struct(rc)
  stream = _stream_of(d.geom_xpos)
  if rc.tier == Tier.PATH:
    _path_trace(m, d, rc, s, seed, stream)
  else:
    # contexts with more cameras than one launch takes render them in consecutive launches
    for cams in camera_chunks(rc):
      sc = scene_struct(m, d, rc, cams)
      out = outputs_struct(rc, cams)
      _abi.call("mjpbr_rt_render", _abi.byref(sc), _abi.byref(s), _abi.byref(out), ctypes.c_int32(int(seed)), stream)
  if rc.tier == Tier.PATH and rc.denoise:
    if rc.denoiser == "learned":
      from mujoco_warp_pbr_amd import learned

      learned.denoise(rc, d.nworld, stream)
    else:
      from mujoco_warp_pbr_amd import denoise

      denoise.denoise(rc, d.nworld, stream)
