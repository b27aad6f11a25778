"""À-trous denoising of the path-traced tier (reference denoise.py:204-248)."""

from __future__ import annotations

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd.types import PbrRenderContext


def atrous_params(rc: PbrRenderContext, nworld: int, cams: tuple[int, int] | None = None):
  """MjpbrAtrousParams of a context: camera layout of the RGB buffer and the filter sigmas;
  `cams` = (first, end) limits it to one launch's cameras (render.camera_chunks)."""
  S = _abi.structs()
  p = S["MjpbrAtrousParams"]()
  base = rc.base
  first, end = (0, int(base.nrender)) if cams is None else cams
  p.nworld = nworld
  p.nrender = end - first
  for k, i in enumerate(range(first, end)):
    p.cam_res[k][0] = int(base.cam_res[i][0])
    p.cam_res[k][1] = int(base.cam_res[i][1])
    p.rgb_adr[k] = int(base.rgb_adr[i])
  p.stride = int(base.rgb_data.shape[1])
  p.iterations = int(rc.denoise_iterations)
  p.sigma_color = float(rc.denoise_sigma_color)
  p.sigma_normal = float(rc.denoise_sigma_normal)
  p.sigma_depth = float(rc.denoise_sigma_depth)
  p.sigma_albedo = float(rc.denoise_sigma_albedo)
  p.tone_mapper = int(rc.tone_mapper)
  p.render_hdr = int(bool(rc.render_hdr))
  p.exposure = _abi.ptr(rc.exposure)
  p.exposure_n = int(rc.exposure.shape[0])
  return p


def denoise(rc: PbrRenderContext, nworld: int, stream: int, world_select=None) -> None:
  """Filters the noisy HDR in `rc.hdr_data` (with the first-hit AOVs) into the RGB (and HDR)
  outputs, in place, on `stream`, one launch per camera_chunks launch; `world_select` (int32
  device tensor, one entry a world) filters only the worlds whose entry is non-zero."""
  from mujoco_warp_pbr_amd.render import camera_chunks
  from mujoco_warp_pbr_amd.render import outputs_struct

  a, b, guides = rc.denoise_buffers
  for cams in camera_chunks(rc):
    p = atrous_params(rc, nworld, cams)
    if world_select is not None:
      p.world_select = _abi.ptr(world_select)
    out = outputs_struct(rc, cams)
    _abi.call("mjpbr_atrous", _abi.byref(p), _abi.byref(out), _abi.ptr(a), _abi.ptr(b), _abi.ptr(guides), stream)
