"""Host restatement of the context tables the GPU builds (TEST INFRASTRUCTURE ONLY, like the rest
of oracle/): texture and radiance mip chains, SH9 irradiance coefficients and importance-sampling
tables of environment maps (reference texture.py:63-101, envmap.py:145-151, :224-249, :402-430).
Used as the checker of the HIP table kernels and to fill the tables of contexts created on the
host (precompute=False)."""

from __future__ import annotations

import numpy as np

LUMINANCE = np.array([0.2126, 0.7152, 0.0722])


def box_halve(img: np.ndarray) -> np.ndarray:
  """One mip step in float64: average column pairs, then row pairs; an axis of one texel stays."""
  out = np.asarray(img, np.float64)
  if out.shape[1] > 1:
    n = out.shape[1] // 2
    out = 0.5 * (out[:, 0 : 2 * n : 2] + out[:, 1 : 2 * n : 2])
  if out.shape[0] > 1:
    n = out.shape[0] // 2
    out = 0.5 * (out[0 : 2 * n : 2] + out[1 : 2 * n : 2])
  return out


def mip_chain(level0: np.ndarray, rgba16f: bool, max_levels: int = 16) -> list[np.ndarray]:
  """Mip levels down to 1 x 1: each cascaded in float64 from the unrounded level above and
  rounded once into the format (uint8: floor(v + 0.5); float16: round to nearest even)."""
  cur = np.asarray(level0, np.float64)
  levels = []
  while True:
    levels.append(cur.astype(np.float16) if rgba16f else np.clip(np.floor(cur + 0.5), 0, 255).astype(np.uint8))
    if cur.shape[:2] == (1, 1) or len(levels) == max_levels:
      return levels
    cur = box_halve(cur)


def texture_level0(rgba8: np.ndarray, rgba16f: bool, srgb_table: np.ndarray) -> np.ndarray:
  """Level 0 as stored: RGBA8 as is, or colour through the sRGB table and alpha / 255 in f16."""
  if not rgba16f:
    return rgba8.astype(np.float64)
  rgb = srgb_table[rgba8[..., :3]].astype(np.float64)
  alpha = (rgba8[..., 3:] / 255.0).astype(np.float16).astype(np.float64)
  return np.concatenate([rgb, alpha], axis=-1)


def radiance_level0(image: np.ndarray) -> np.ndarray:
  h, w = image.shape[:2]
  rgba = np.concatenate([np.asarray(image, np.float64)[..., :3], np.ones((h, w, 1))], axis=-1)
  return rgba.astype(np.float16).astype(np.float64)


def sh9(image: np.ndarray) -> np.ndarray:
  """(9, 3): the radiance projected on the real SH basis (Ramamoorthi & Hanrahan 2001, constants
  to six digits), each texel weighted by its solid angle, bands scaled by the Lambertian
  convolution over pi (1, 2/3, 1/4): evaluating it gives irradiance / pi."""
  img = np.asarray(image, np.float64)[..., :3]
  h, w = img.shape[:2]
  theta = np.pi * (np.arange(h) + 0.5) / h
  phi = 2.0 * np.pi * ((np.arange(w) + 0.5) / w - 0.5)
  sin_t = np.sin(theta)[:, None]
  x = sin_t * np.cos(phi)[None, :]
  y = sin_t * np.sin(phi)[None, :]
  z = np.cos(theta)[:, None] * np.ones((1, w))
  basis = [
    0.282095 * np.ones((h, w)),
    0.488603 * y,
    0.488603 * z,
    0.488603 * x,
    1.092548 * x * y,
    1.092548 * y * z,
    0.315392 * (3.0 * z * z - 1.0),
    1.092548 * x * z,
    0.546274 * (x * x - y * y),
  ]
  d_omega = (2.0 * np.pi / w) * (np.pi / h) * sin_t
  conv = [1.0] + [2.0 / 3.0] * 3 + [0.25] * 5
  return np.stack([conv[k] * np.tensordot(basis[k] * d_omega, img, axes=([0, 1], [0, 1])) for k in range(9)]).astype(np.float32)


def block_mean_to_width(image: np.ndarray, max_width: int) -> np.ndarray:
  """Mean over 2^k x 2^k blocks (trailing rows / columns dropped) with the smallest k that brings
  the width to at most max_width."""
  img = np.asarray(image, np.float64)[..., :3]
  k = 0
  while (img.shape[1] >> k) > max_width:
    k += 1
  b = 1 << k
  h, w = img.shape[0] >> k, img.shape[1] >> k
  return img[: h * b, : w * b].reshape(h, b, w, b, 3).mean(axis=(1, 3))


def sampling_tables(image: np.ndarray, max_width: int = 256):
  """(prob (h, w), marginal (h + 1,), conditional (h, w + 1)), float32: p over equirect cells
  proportional to luminance sin(theta) + 1e-6 mean luminance, CDF ends exactly 1."""
  small = block_mean_to_width(image, max_width)
  h = small.shape[0]
  lum = small @ LUMINANCE
  weight = lum * np.sin(np.pi * (np.arange(h) + 0.5) / h)[:, None] + 1e-6 * max(lum.mean(), 1e-12)
  prob = weight / weight.sum()
  rows = prob.sum(axis=1)
  marginal = np.zeros(h + 1)
  marginal[1:] = np.cumsum(rows)
  marginal[-1] = 1.0
  conditional = np.zeros((h, prob.shape[1] + 1))
  conditional[:, 1:] = np.cumsum(prob, axis=1) / np.maximum(rows, 1e-30)[:, None]
  conditional[:, -1] = 1.0
  return prob.astype(np.float32), marginal.astype(np.float32), conditional.astype(np.float32)


def fill_textures(pool) -> None:
  """Mip chains of every pooled texture of a host TexturePool, written into pool.data."""
  from mujoco_warp_pbr_amd.texture import FORMAT_RGBA16F
  from mujoco_warp_pbr_amd.texture import SRGB_TABLE

  data = pool.data.numpy()
  for s in pool.sources:
    f16 = s.fmt == FORMAT_RGBA16F
    for lv, img in enumerate(mip_chain(texture_level0(s.rgba8, f16, SRGB_TABLE), f16)):
      raw = np.ascontiguousarray(img).view(np.uint8).ravel()
      a = int(pool.host_level_adr[s.tex_id, lv])
      data[a : a + raw.size] = raw


def fill_environments(env) -> None:
  """Radiance mip chains, SH9 coefficients and sampling tables of a host Environments."""
  import torch

  rad = env.radiance.numpy().view(np.float16)
  adr = env.level_adr.numpy()
  for e, (im, tim) in enumerate(zip(env.images, env.table_images)):
    for lv, img in enumerate(mip_chain(radiance_level0(im), True)):
      a = int(adr[e, lv])
      rad[a : a + img.size] = img.ravel()
    env.sh[e].copy_(torch.from_numpy(sh9(im)))
    prob, marginal, conditional = sampling_tables(tim)
    env.prob[e].copy_(torch.from_numpy(prob))
    env.marginal[e].copy_(torch.from_numpy(marginal))
    env.conditional[e].copy_(torch.from_numpy(conditional))
