"""Per-camera readback (reference output.py:34-152).

The get_* functions copy camera `camera_index` of every world (the first `out.shape[0]` worlds)
into a caller-allocated `out`, whose shape and dtype are checked; nothing synchronises with the
host. The *_view functions return views of the same pixels in the context's buffers instead.
"""

from __future__ import annotations

import torch

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd.render import _stream_of
from mujoco_warp_pbr_amd.types import PbrRenderContext


def _camera(rc: PbrRenderContext, camera_index: int, adr_name: str, what: str):
  base = rc.base
  if not 0 <= camera_index < base.nrender:
    raise ValueError(f"camera_index must be in [0, {base.nrender}), got {camera_index}")
  adr = int(getattr(base, adr_name)[camera_index])
  if adr < 0:
    raise ValueError(f"camera {camera_index} does not render {what}")
  w, h = (int(x) for x in base.cam_res[camera_index])
  return adr, w, h


def _check_out(out: torch.Tensor, rc: PbrRenderContext, h: int, w: int, tail: tuple, dtype, name: str):
  n = out.shape[0] if out.dim() else 0
  shape = (n, h, w, *tail)
  if out.dim() != len(shape) or tuple(out.shape) != shape or n < 1 or n > rc.base.nworld:
    per_pixel = f" of {tail[0]}-vectors" if tail else ""
    raise ValueError(f"{name} must have shape (<= {rc.base.nworld}, {h}, {w}){per_pixel}, got {tuple(out.shape)}")
  if out.dtype != dtype:
    raise ValueError(f"{name} must have dtype {dtype}, got {out.dtype}")
  return n


def get_rgb(rc: PbrRenderContext, camera_index: int, rgb_out: torch.Tensor) -> None:
  """Display RGB in [0, 1] of one camera, (n, h, w, 3) float32: the packed sRGB8 bytes / 255
  (rgb_data holds 0xAARRGGBB words, MuJoCo Warp's layout)."""
  adr, w, h = _camera(rc, camera_index, "rgb_adr", "rgb")
  n = _check_out(rgb_out, rc, h, w, (3,), torch.float32, "rgb_out")
  bits = rc.base.rgb_data[:n, adr : adr + w * h]
  for c, shift in enumerate((16, 8, 0)):
    rgb_out[..., c] = ((bits >> shift) & 0xFF).reshape(n, h, w).to(torch.float32) / 255.0


def get_depth(rc: PbrRenderContext, camera_index: int, depth_scale: float, depth_out: torch.Tensor) -> None:
  """Planar depth / depth_scale clamped to [0, 1] of one camera, (n, h, w) float32 (reference
  output.py:41-45; background pixels hold depth 0)."""
  adr, w, h = _camera(rc, camera_index, "depth_adr", "depth")
  n = _check_out(depth_out, rc, h, w, (), torch.float32, "depth_out")
  depth = rc.base.depth_data[:n, adr : adr + w * h].reshape(n, h, w)
  torch.clamp(depth / float(depth_scale), 0.0, 1.0, out=depth_out)


def get_segmentation(rc: PbrRenderContext, camera_index: int, segmentation_out: torch.Tensor) -> None:
  """(geom id, object type) of one camera, (n, h, w, 2) int32; (-1, -1) for the background."""
  adr, w, h = _camera(rc, camera_index, "seg_adr", "segmentation")
  n = _check_out(segmentation_out, rc, h, w, (2,), torch.int32, "segmentation_out")
  segmentation_out.copy_(rc.base.seg_data[:n, adr : adr + w * h].reshape(n, h, w, 2))


def _get_rgb_image(rc: PbrRenderContext, buffer: torch.Tensor, camera_index: int, out: torch.Tensor, name: str) -> None:
  adr, w, h = _camera(rc, camera_index, "rgb_adr", "rgb")
  n = _check_out(out, rc, h, w, (3,), torch.float32, name)
  if not out.is_contiguous():
    raise ValueError(f"{name} must be contiguous")
  _abi.call(
    "mjpbr_extract_rgb_image",
    _abi.ptr(buffer),
    int(buffer.shape[1]),
    adr,
    w,
    h,
    n,
    _abi.ptr(out),
    _stream_of(buffer),
  )


def _check_guides(rc: PbrRenderContext) -> None:
  if not rc.render_guides:
    raise ValueError("create the render context with render_guides=True to read the albedo / normal guides")


def get_hdr(rc: PbrRenderContext, camera_index: int, hdr_out: torch.Tensor) -> None:
  """Linear HDR radiance (before exposure and tone mapping), (n, h, w, 3); needs render_hdr."""
  if not rc.render_hdr:
    raise ValueError("create the render context with render_hdr=True to read HDR output")
  _get_rgb_image(rc, rc.hdr_data, camera_index, hdr_out, "hdr_out")


def get_albedo(rc: PbrRenderContext, camera_index: int, albedo_out: torch.Tensor) -> None:
  """First-hit linear base colour, (n, h, w, 3); needs render_guides."""
  _check_guides(rc)
  _get_rgb_image(rc, rc.guide_albedo, camera_index, albedo_out, "albedo_out")


def get_normal(rc: PbrRenderContext, camera_index: int, normal_out: torch.Tensor) -> None:
  """First-hit world shading normal, (n, h, w, 3); needs render_guides."""
  _check_guides(rc)
  _get_rgb_image(rc, rc.guide_normal, camera_index, normal_out, "normal_out")


# ----------------------------------------------------------------------------- zero-copy views
# Views of one camera's pixels inside the context's own buffers: no copy, no launch; they alias
# what the next render writes. Torch views with the buffers' strides, so torch.utils.dlpack
# hands them to other frameworks without a copy either.


def _pixels(buffer: torch.Tensor, adr: int, w: int, h: int) -> torch.Tensor:
  return buffer[:, adr : adr + w * h].unflatten(1, (h, w))


def bgra_view(rc: PbrRenderContext, camera_index: int) -> torch.Tensor:
  """(nworld, h, w, 4) uint8 view of the packed display colour: each 0xAARRGGBB word is the bytes
  B, G, R, A in memory (little endian), so channels are (B, G, R, A), A = 255."""
  adr, w, h = _camera(rc, camera_index, "rgb_adr", "rgb")
  return rc.base.rgb_data[:, adr : adr + w * h].view(torch.uint8).unflatten(1, (h, w, 4))


def packed_rgb_view(rc: PbrRenderContext, camera_index: int) -> torch.Tensor:
  """(nworld, h, w) int32 view of the packed 0xAARRGGBB words (MuJoCo Warp's rgb_data layout)."""
  adr, w, h = _camera(rc, camera_index, "rgb_adr", "rgb")
  return _pixels(rc.base.rgb_data, adr, w, h)


def depth_view(rc: PbrRenderContext, camera_index: int) -> torch.Tensor:
  """(nworld, h, w) float32 view of the planar depth (0 for background; unscaled, unclamped)."""
  adr, w, h = _camera(rc, camera_index, "depth_adr", "depth")
  return _pixels(rc.base.depth_data, adr, w, h)


def segmentation_view(rc: PbrRenderContext, camera_index: int) -> torch.Tensor:
  """(nworld, h, w, 2) int32 view of (geom id, object type); (-1, -1) for the background."""
  adr, w, h = _camera(rc, camera_index, "seg_adr", "segmentation")
  return rc.base.seg_data[:, adr : adr + w * h].unflatten(1, (h, w))


def _vec3_view(rc: PbrRenderContext, buffer: torch.Tensor, camera_index: int) -> torch.Tensor:
  adr, w, h = _camera(rc, camera_index, "rgb_adr", "rgb")
  return buffer[:, adr : adr + w * h].unflatten(1, (h, w))


def hdr_view(rc: PbrRenderContext, camera_index: int) -> torch.Tensor:
  """(nworld, h, w, 3) float32 view of the linear HDR radiance; needs render_hdr."""
  if not rc.render_hdr:
    raise ValueError("create the render context with render_hdr=True to read HDR output")
  return _vec3_view(rc, rc.hdr_data, camera_index)


def albedo_view(rc: PbrRenderContext, camera_index: int) -> torch.Tensor:
  """(nworld, h, w, 3) float32 view of the first-hit base colour; needs render_guides."""
  _check_guides(rc)
  return _vec3_view(rc, rc.guide_albedo, camera_index)


def normal_view(rc: PbrRenderContext, camera_index: int) -> torch.Tensor:
  """(nworld, h, w, 3) float32 view of the first-hit shading normal; needs render_guides."""
  _check_guides(rc)
  return _vec3_view(rc, rc.guide_normal, camera_index)
