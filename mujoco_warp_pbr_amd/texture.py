"""Material textures (reference texture.py:41-101): one device byte pool holding every used 2D
texture's full mip chain, built on the GPU from the uploaded RGBA8 texels.

Colour-role (sRGB) textures become linear RGBA16F, so filtering and mips are linear; data textures
(normal, ORM, ...) stay normalised RGBA8. Level l of texture t starts at byte `level_adr[t, l]`
(16-byte aligned); `desc[t] = (width, height, nlevels, format)` with format 0 = RGBA8,
1 = RGBA16F. Unused and non-2D textures hold one white RGBA8 texel.
"""

from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import torch

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd.material import srgb_to_linear

MAX_LEVELS = 16
FORMAT_RGBA8 = 0
FORMAT_RGBA16F = 1
_COLOR_ROLES = [mdl.ROLE_RGB, mdl.ROLE_RGBA, mdl.ROLE_EMISSIVE]
# sRGB byte -> linear f16 (the decode the GPU applies to colour textures' level 0)
SRGB_TABLE = srgb_to_linear(np.arange(256) / 255.0).astype(np.float16)


def is_srgb(mjm: mdl.HostModel, tex_id: int) -> bool:
  """texture.py:54-61: explicit colour space, else colour roles are sRGB."""
  cs = int(mjm.tex_colorspace[tex_id])
  if cs == mdl.COLORSPACE_SRGB:
    return True
  if cs == mdl.COLORSPACE_LINEAR:
    return False
  return bool(np.isin(tex_id, mjm.mat_texid[:, _COLOR_ROLES])) if mjm.nmat else False


def level_shapes(width: int, height: int) -> list[tuple[int, int]]:
  """(width, height) of each mip level down to 1 x 1; an axis of one texel stays one texel."""
  shapes = [(width, height)]
  while shapes[-1] != (1, 1) and len(shapes) < MAX_LEVELS:
    w, h = shapes[-1]
    shapes.append((max(w // 2, 1), max(h // 2, 1)))
  return shapes


@dataclasses.dataclass
class TextureSource:
  """Level-0 texels of one pooled texture (what the mip chain is built from)."""

  tex_id: int
  rgba8: np.ndarray  # (h, w, 4) uint8
  fmt: int


@dataclasses.dataclass
class TexturePool:
  ntex: int
  data: torch.Tensor  # uint8 pool
  level_adr: torch.Tensor  # (ntex, MAX_LEVELS) int64 byte offsets
  desc: torch.Tensor  # (ntex, 4) int32
  size: np.ndarray  # (ntex, 2) float32 (width, height)
  sources: list  # TextureSource of every pooled 2D texture
  host_level_adr: np.ndarray  # level_adr on the host


def _rgba8(mjm: mdl.HostModel, t: int) -> np.ndarray:
  w, h, nc = int(mjm.tex_width[t]), int(mjm.tex_height[t]), int(mjm.tex_nchannel[t])
  texels = np.asarray(mjm.tex_data[mjm.tex_adr[t] : mjm.tex_adr[t] + w * h * nc], np.uint8).reshape(h, w, nc)
  out = np.full((h, w, 4), 255, np.uint8)
  out[..., : min(nc, 4)] = texels[..., :4]
  if nc == 1:
    out[..., 1:3] = texels
  return out


def upload_textures(mjm: mdl.HostModel, used: np.ndarray, device, precompute: bool = True) -> TexturePool:
  """Pool of mipmapped textures for the 2D textures in `used` (GPU-built mips with precompute;
  otherwise the pooled textures' levels are left zero for the caller to fill)."""
  ntex = max(mjm.ntex, 1)
  level_adr = np.zeros((ntex, MAX_LEVELS), np.int64)
  desc = np.zeros((ntex, 4), np.int32)
  size = np.ones((ntex, 2), np.float32)
  sources: list[TextureSource] = []
  white: list[int] = []
  offset = 0
  for t in range(ntex):
    pooled = t < mjm.ntex and bool(used[t]) and mjm.tex_type[t] == mdl.TEXTURE_2D
    if pooled:
      rgba = _rgba8(mjm, t)
      fmt = FORMAT_RGBA16F if is_srgb(mjm, t) else FORMAT_RGBA8
      sources.append(TextureSource(t, rgba, fmt))
      shapes = level_shapes(rgba.shape[1], rgba.shape[0])
      size[t] = (rgba.shape[1], rgba.shape[0])
    else:
      fmt, shapes = FORMAT_RGBA8, [(1, 1)]
      white.append(t)
    desc[t] = (shapes[0][0], shapes[0][1], len(shapes), fmt)
    texel = 8 if fmt == FORMAT_RGBA16F else 4
    for lv, (w, h) in enumerate(shapes):
      level_adr[t, lv] = offset
      offset += -(-w * h * texel // 16) * 16
  pool = torch.zeros(max(offset, 16), dtype=torch.uint8, device=device)
  for t in white:
    pool[int(level_adr[t, 0]) : int(level_adr[t, 0]) + 4] = 255
  tp = TexturePool(
    ntex=ntex,
    data=pool,
    level_adr=torch.as_tensor(level_adr).to(device),
    desc=torch.as_tensor(desc).to(device),
    size=size,
    sources=sources,
    host_level_adr=level_adr,
  )
  if precompute and sources:
    _build_mips(tp)
  return tp


def _build_mips(tp: TexturePool) -> None:
  dev = tp.data.device
  stream = int(torch.cuda.current_stream(dev).cuda_stream)
  table = torch.as_tensor(SRGB_TABLE.view(np.int16)).to(dev)
  largest = max(s.rgba8.shape[0] * s.rgba8.shape[1] for s in tp.sources)
  work = torch.empty(8 * largest, dtype=torch.float64, device=dev)
  for s in tp.sources:
    h, w = s.rgba8.shape[:2]
    n = int(tp.desc[s.tex_id, 2])
    src = torch.as_tensor(np.ascontiguousarray(s.rgba8)).to(dev)
    adr = (ctypes.c_int64 * n)(*(int(x) for x in tp.host_level_adr[s.tex_id, :n]))  # read on the host by the launcher
    _abi.call(
      "mjpbr_texture_mips",
      _abi.ptr(src),
      w,
      h,
      s.fmt,
      _abi.ptr(table),
      _abi.ptr(tp.data),
      ctypes.cast(adr, ctypes.c_void_p),
      n,
      _abi.ptr(work),
      work.numel(),
      stream,
    )
