"""Material textures: decoding and mip chains in one device byte pool (reference
texture.py:41-101).

sRGB (colour-role) textures become linear float16 RGBA (filtering and mips are then linear);
data textures (normal, ORM, ...) stay normalised RGBA8. Each texture gets a full mip chain
made by 2x2 box filtering. Level l of texture t starts at byte `level_adr[t, l]`;
`desc[t] = (width, height, nlevels, format)` with format 0 = RGBA8, 1 = RGBA16F.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd.material import srgb_to_linear

MAX_LEVELS = 16
FORMAT_RGBA8 = 0
FORMAT_RGBA16F = 1
_COLOR_ROLES = [mdl.ROLE_RGB, mdl.ROLE_RGBA, mdl.ROLE_EMISSIVE]
SRGB_TABLE = srgb_to_linear(np.arange(256) / 255.0).astype(np.float16)


def is_srgb(mjm: mdl.HostModel, tex_id: int) -> bool:
  """texture.py:54-61: explicit colour space, else colour roles are sRGB."""
  cs = int(mjm.tex_colorspace[tex_id])
  if cs == mdl.COLORSPACE_SRGB:
    return True
  if cs == mdl.COLORSPACE_LINEAR:
    return False
  return bool(np.isin(tex_id, mjm.mat_texid[:, _COLOR_ROLES])) if mjm.nmat else False


def downsample2(image: np.ndarray) -> np.ndarray:
  """One mip step: 2x2 box filter (a 1-texel axis is kept), computed in float64."""
  h, w = image.shape[:2]
  img = image.astype(np.float64)
  if w > 1:
    w2 = w // 2
    img = 0.5 * (img[:, 0 : 2 * w2 : 2] + img[:, 1 : 2 * w2 : 2])
  if h > 1:
    h2 = h // 2
    img = 0.5 * (img[0 : 2 * h2 : 2] + img[1 : 2 * h2 : 2])
  return img


def mip_chain(image: np.ndarray, fmt: int) -> list[np.ndarray]:
  """Full mip chain down to 1x1, each level stored in `fmt` (uint8 RGBA or float16 RGBA)."""
  levels = []
  cur = image.astype(np.float64)
  while True:
    if fmt == FORMAT_RGBA8:
      levels.append(np.clip(np.floor(cur + 0.5), 0, 255).astype(np.uint8))
    else:
      levels.append(cur.astype(np.float16))
    h, w = cur.shape[:2]
    if (h == 1 and w == 1) or len(levels) == MAX_LEVELS:
      break
    cur = downsample2(cur)
  return levels


@dataclasses.dataclass
class TexturePool:
  ntex: int
  data: torch.Tensor  # uint8 pool
  level_adr: torch.Tensor  # (ntex, MAX_LEVELS) int64 byte offsets
  desc: torch.Tensor  # (ntex, 4) int32
  size: np.ndarray  # (ntex, 2) float32 (width, height)


def _rgba8(mjm: mdl.HostModel, t: int) -> np.ndarray:
  w, h, nc = int(mjm.tex_width[t]), int(mjm.tex_height[t]), int(mjm.tex_nchannel[t])
  data = mjm.tex_data[mjm.tex_adr[t] : mjm.tex_adr[t] + w * h * nc].reshape(h, w, nc)
  rgba = np.full((h, w, 4), 255, dtype=np.uint8)
  if nc == 1:
    rgba[..., :3] = data
  else:
    rgba[..., :nc] = data
  return rgba


def upload_textures(mjm: mdl.HostModel, used: np.ndarray, device) -> TexturePool:
  """Pool of mipmapped textures for the 2D textures in `used`; others get a 1x1 white texel."""
  chunks: list[np.ndarray] = []
  offset = 0
  ntex = max(mjm.ntex, 1)
  level_adr = np.zeros((ntex, MAX_LEVELS), np.int64)
  desc = np.zeros((ntex, 4), np.int32)
  size = np.ones((ntex, 2), np.float32)
  for t in range(ntex):
    if t >= mjm.ntex or not used[t] or mjm.tex_type[t] != mdl.TEXTURE_2D:
      levels, fmt = [np.full((1, 1, 4), 255, np.uint8)], FORMAT_RGBA8
    else:
      rgba = _rgba8(mjm, t)
      if is_srgb(mjm, t):
        texels = np.concatenate([SRGB_TABLE[rgba[..., :3]].astype(np.float64), rgba[..., 3:] / 255.0], axis=-1)
        fmt = FORMAT_RGBA16F
        texels = texels.astype(np.float16).astype(np.float64)
      else:
        texels, fmt = rgba.astype(np.float64), FORMAT_RGBA8
      levels = mip_chain(texels, fmt)
      size[t] = (rgba.shape[1], rgba.shape[0])
    desc[t] = (levels[0].shape[1], levels[0].shape[0], len(levels), fmt)
    for lv, img in enumerate(levels):
      raw = np.ascontiguousarray(img).view(np.uint8).ravel()
      level_adr[t, lv] = offset
      chunks.append(raw)
      offset += raw.size
      pad = (-offset) % 16
      if pad:
        chunks.append(np.zeros(pad, np.uint8))
        offset += pad
  pool = np.concatenate(chunks) if chunks else np.zeros(16, np.uint8)
  return TexturePool(
    ntex=ntex,
    data=torch.as_tensor(pool).to(device),
    level_adr=torch.as_tensor(level_adr).to(device),
    desc=torch.as_tensor(desc).to(device),
    size=size,
  )
