"""À-trous denoising of the path-traced tier (reference denoise.py:204-248)."""

from __future__ import annotations

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd.types import PbrRenderContext


def atrous_params(rc: PbrRenderContext, nworld: int):
  """MjpbrAtrousParams of a context: camera layout of the RGB buffer and the filter sigmas."""
  S = _abi.structs()
  p = S["MjpbrAtrousParams"]()
  base = rc.base
  p.nworld = nworld
  p.nrender = base.nrender
  for i in range(base.nrender):
    p.cam_res[i][0] = int(base.cam_res[i][0])
    p.cam_res[i][1] = int(base.cam_res[i][1])
    p.rgb_adr[i] = int(base.rgb_adr[i])
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


def denoise(rc: PbrRenderContext, nworld: int, out, stream: int) -> None:
  """Filters the noisy HDR in `rc.hdr_data` (with the first-hit AOVs) into the RGB (and HDR)
  outputs, in place, on `stream`."""
  p = atrous_params(rc, nworld)
  a, b, guides = rc.denoise_buffers
  _abi.call("mjpbr_atrous", _abi.byref(p), _abi.byref(out), _abi.ptr(a), _abi.ptr(b), _abi.ptr(guides), stream)
