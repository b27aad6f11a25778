"""Environment maps: RGBE loading, radiance mip chains, GGX-prefiltered stacks (on the GPU),
SH9 irradiance and importance-sampling tables (reference envmap.py:57-576).

Per map we keep the radiance as an RGBA16F mip chain (background, sharp reflections), an
8-level GGX-prefiltered stack whose depth axis is perceptual roughness, and 9 SH coefficients
of the Lambertian radiance. All maps share the prefiltered stack and sampling-table shapes.
"""

from __future__ import annotations

import dataclasses
import math
import re
from pathlib import Path

import numpy as np
import torch

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd.material import LUMINANCE_WEIGHTS

PREFILTER_WIDTH = 256
PREFILTER_LEVELS = 8
PREFILTER_SAMPLES = 512
MAX_RADIANCE_WIDTH = 2048
CDF_WIDTH = 256
MAX_LEVELS = 16


def _decode_rle_scanline(data: bytes, pos: int, width: int, path: str, y: int):
  """Adaptive run-length encoded RGBE scanline at `pos` -> ((width, 4) uint8, end)."""
  line = bytearray(4 * width)
  n = len(data)
  for c in range(4):
    x = c * width
    end = x + width
    while x < end:
      if pos >= n:
        raise ValueError(f"{path}: truncated run-length data in scanline {y}")
      count = data[pos]
      pos += 1
      if count > 128:
        count -= 128
        if count > end - x or pos >= n:
          raise ValueError(f"{path}: bad run-length data in scanline {y}")
        line[x : x + count] = bytes((data[pos],)) * count
        pos += 1
      else:
        if count == 0 or count > end - x or pos + count > n:
          raise ValueError(f"{path}: bad run-length data in scanline {y}")
        line[x : x + count] = data[pos : pos + count]
        pos += count
      x += count
  return np.frombuffer(bytes(line), dtype=np.uint8).reshape(4, width).T, pos


def load_hdr(path) -> np.ndarray:
  """Reads a Radiance RGBE (.hdr) file as float32 (height, width, 3) (reference envmap.py:106-142).

  Raises:
    ValueError: not a '-Y H +X W' RGBE file, or truncated / malformed pixel data.
  """
  data = Path(path).read_bytes()
  if not data.startswith(b"#?"):
    raise ValueError(f"{path}: not a Radiance .hdr file (missing '#?' signature)")
  try:
    header_end = data.index(b"\n\n") + 2
    resolution_end = data.index(b"\n", header_end)
  except ValueError:
    raise ValueError(f"{path}: truncated .hdr header") from None
  match = re.match(rb"-Y (\d+) \+X (\d+)", data[header_end:resolution_end])
  if match is None:
    raise ValueError(f"{path}: only '-Y H +X W' oriented .hdr files are supported")
  height, width = int(match.group(1)), int(match.group(2))
  pos = resolution_end + 1
  rgbe = np.empty((height, width, 4), dtype=np.uint8)
  for y in range(height):
    rle = 8 <= width <= 0x7FFF and data[pos : pos + 2] == b"\x02\x02" and pos + 2 < len(data) and not data[pos + 2] & 0x80
    if not rle:
      if pos + 4 * width > len(data):
        raise ValueError(f"{path}: truncated pixel data in scanline {y}")
      rgbe[y] = np.frombuffer(data, dtype=np.uint8, count=4 * width, offset=pos).reshape(width, 4)
      pos += 4 * width
      continue
    if (data[pos + 2] << 8 | data[pos + 3]) != width:
      raise ValueError(f"{path}: scanline {y} width does not match the header")
    rgbe[y], pos = _decode_rle_scanline(data, pos + 4, width, str(path), y)
  exponent = rgbe[..., 3].astype(np.int32)
  scale = np.where(exponent > 0, np.ldexp(1.0, exponent - 136), 0.0).astype(np.float32)
  return rgbe[..., :3].astype(np.float32) * scale[..., None]


def save_hdr(path, image: np.ndarray) -> None:
  """Writes float (height, width, 3) radiance as a flat (not run-length encoded) RGBE file."""
  image = np.asarray(image, dtype=np.float64)
  h, w, _ = image.shape
  m = image.max(axis=-1)
  mant, expo = np.frexp(m)
  scale = np.where(m > 1e-32, mant * 256.0 / np.maximum(m, 1e-300), 0.0)
  rgbe = np.zeros((h, w, 4), np.uint8)
  rgbe[..., :3] = np.clip(image * scale[..., None], 0, 255).astype(np.uint8)
  rgbe[..., 3] = np.where(m > 1e-32, expo + 128, 0).astype(np.uint8)
  header = f"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y {h} +X {w}\n".encode()
  Path(path).write_bytes(header + rgbe.tobytes())


def downsample(image: np.ndarray, max_width: int) -> np.ndarray:
  """Box-filter halvings until the width is at most max_width (envmap.py:145-151)."""
  while image.shape[1] > max_width:
    h, w = image.shape[0] // 2 * 2, image.shape[1] // 2 * 2
    image = image[:h, :w].reshape(h // 2, 2, w // 2, 2, -1).mean(axis=(1, 3))
  return image


def sh9_irradiance(image: np.ndarray) -> np.ndarray:
  """SH9 coefficients (9, 3) of the Lambertian radiance of an equirect image (envmap.py:224-249)."""
  h, w, _ = image.shape
  theta = (np.arange(h) + 0.5) / h * np.pi
  phi = ((np.arange(w) + 0.5) / w - 0.5) * 2.0 * np.pi
  st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
  x, y, z = st * np.cos(phi)[None], st * np.sin(phi)[None], np.broadcast_to(ct, (h, w))
  basis = np.stack(
    [
      np.full((h, w), 0.282095),
      0.488603 * y,
      0.488603 * z,
      0.488603 * x,
      1.092548 * x * y,
      1.092548 * y * z,
      0.315392 * (3.0 * z * z - 1.0),
      1.092548 * x * z,
      0.546274 * (x * x - y * y),
    ]
  )
  solid_angle = (2.0 * np.pi / w) * (np.pi / h) * np.broadcast_to(st, (h, w))
  coeffs = np.einsum("khw,hwc->kc", basis * solid_angle, image.astype(np.float64))
  band = np.array([np.pi, 2 * np.pi / 3, 2 * np.pi / 3, 2 * np.pi / 3] + [np.pi / 4] * 5)
  return (coeffs * band[:, None] / np.pi).astype(np.float32)


def sampling_tables(image: np.ndarray, width: int = CDF_WIDTH):
  """(prob (h, w), marginal (h + 1,), conditional (h, w + 1)) of the distribution over equirect
  cells proportional to luminance * sin(theta) (envmap.py:402-430)."""
  small = downsample(np.asarray(image, dtype=np.float64)[..., :3], width)
  h = small.shape[0]
  lum = small @ np.array(LUMINANCE_WEIGHTS)
  sin_theta = np.sin((np.arange(h) + 0.5) / h * np.pi)[:, None]
  weight = lum * sin_theta + 1e-6 * max(float(lum.mean()), 1e-12)
  prob = weight / weight.sum()
  row = prob.sum(axis=1)
  marginal = np.concatenate([[0.0], np.cumsum(row)])
  conditional = np.concatenate([np.zeros((h, 1)), np.cumsum(prob, axis=1) / np.maximum(row[:, None], 1e-30)], axis=1)
  marginal[-1] = 1.0
  conditional[:, -1] = 1.0
  return prob.astype(np.float32), marginal.astype(np.float32), conditional.astype(np.float32)


def _downsampled_shape(shape, max_width: int):
  h, w = shape[0], shape[1]
  while w > max_width:
    h, w = h // 2, w // 2
  return h, w


def _resize_nearest(image: np.ndarray, shape) -> np.ndarray:
  h, w = shape
  ys = (np.arange(h) + 0.5) / h * image.shape[0]
  xs = (np.arange(w) + 0.5) / w * image.shape[1]
  return image[ys.astype(int)][:, xs.astype(int)]


def radiance_levels(image: np.ndarray) -> list[np.ndarray]:
  """RGBA16F mip chain of a radiance image (alpha 1), 2x2 box filtered down to 1x1."""
  from mujoco_warp_pbr_amd.texture import downsample2

  h, w, _ = image.shape
  cur = np.concatenate([image.astype(np.float64), np.ones((h, w, 1))], axis=-1).astype(np.float16).astype(np.float64)
  levels = []
  while True:
    levels.append(cur.astype(np.float16))
    if (cur.shape[0] == 1 and cur.shape[1] == 1) or len(levels) == MAX_LEVELS:
      break
    cur = downsample2(cur)
  return levels


@dataclasses.dataclass
class Environments:
  """Device resources of the environment maps (reference envmap.py:493-515)."""

  nenv: int
  radiance: torch.Tensor  # uint16 (f16 bits) pool
  level_adr: torch.Tensor  # (nenv, MAX_LEVELS) int64, offsets in halves
  desc: torch.Tensor  # (nenv, 4) int32: width, height, nlevels, 0
  prefiltered: torch.Tensor  # (nenv, L, H, W, 4) f16 bits
  sh: torch.Tensor  # (nenv, 9, 3)
  width: float
  prob: torch.Tensor
  marginal: torch.Tensor
  conditional: torch.Tensor
  images: list

  @property
  def pf_shape(self):
    return tuple(self.prefiltered.shape[1:4])


def _upload_radiance(images: list[np.ndarray], device):
  chunks, offset = [], 0
  level_adr = np.zeros((len(images), MAX_LEVELS), np.int64)
  desc = np.zeros((len(images), 4), np.int32)
  for e, im in enumerate(images):
    levels = radiance_levels(im)
    desc[e] = (im.shape[1], im.shape[0], len(levels), 1)  # format 1: RGBA16F
    for lv, img in enumerate(levels):
      raw = np.ascontiguousarray(img).view(np.uint16).ravel()
      level_adr[e, lv] = offset
      chunks.append(raw)
      offset += raw.size
  pool = np.concatenate(chunks)
  return (
    torch.as_tensor(pool.view(np.int16)).to(device),
    torch.as_tensor(level_adr).to(device),
    torch.as_tensor(desc).to(device),
  )


def build_environments(images: list[np.ndarray], device, nsample: int = PREFILTER_SAMPLES, prefilter: bool = True) -> Environments:
  """GPU resources of a list of equirect radiance maps; prefiltering runs on the GPU
  (prefilter=False leaves the stacks zero for a caller that fills them)."""
  imgs = [downsample(np.asarray(im, dtype=np.float32)[..., :3], MAX_RADIANCE_WIDTH) for im in images]
  radiance, level_adr, desc = _upload_radiance(imgs, device)
  pw = min(PREFILTER_WIDTH, min(im.shape[1] for im in imgs))
  pw = max(pw, 2)
  stack = torch.zeros((len(imgs), PREFILTER_LEVELS, pw // 2, pw, 4), dtype=torch.int16, device=device)
  stream = _stream(device)
  for e in range(len(imgs) if prefilter else 0):
    _abi.call(
      "mjpbr_env_prefilter",
      _abi.ptr(radiance),
      _abi.ptr(level_adr[e]),
      _abi.ptr(desc[e]),
      _abi.ptr(stack[e]),
      PREFILTER_LEVELS,
      pw // 2,
      pw,
      nsample,
      stream,
    )
  shape = None
  tables = []
  for im in imgs:
    if shape is not None and _downsampled_shape(im.shape, CDF_WIDTH) != shape:
      im = _resize_nearest(im, shape)
    t = sampling_tables(im)
    shape = shape or t[0].shape
    tables.append(t)
  return Environments(
    nenv=len(imgs),
    radiance=radiance,
    level_adr=level_adr,
    desc=desc,
    prefiltered=stack,
    sh=torch.as_tensor(np.stack([sh9_irradiance(im) for im in imgs])).to(device),
    width=float(max(im.shape[1] for im in imgs)),
    prob=torch.as_tensor(np.stack([t[0] for t in tables])).to(device),
    marginal=torch.as_tensor(np.stack([t[1] for t in tables])).to(device),
    conditional=torch.as_tensor(np.stack([t[2] for t in tables])).to(device),
    images=imgs,
  )


def empty_environments(device) -> Environments:
  """Placeholder resources of a context without environment map (constant ambient)."""
  radiance, level_adr, desc = _upload_radiance([np.zeros((1, 1, 3), np.float32)], device)
  return Environments(
    nenv=1,
    radiance=radiance,
    level_adr=level_adr,
    desc=desc,
    prefiltered=torch.zeros((1, 2, 1, 2, 4), dtype=torch.int16, device=device),
    sh=torch.zeros((1, 9, 3), dtype=torch.float32, device=device),
    width=1.0,
    prob=torch.ones((1, 1, 1), dtype=torch.float32, device=device),
    marginal=torch.tensor([[0.0, 1.0]], dtype=torch.float32, device=device),
    conditional=torch.tensor([[[0.0, 1.0]]], dtype=torch.float32, device=device),
    images=[],
  )


SKYBOX_WIDTH = 1024


def uv_to_direction(u: np.ndarray, v: np.ndarray) -> np.ndarray:
  """World direction of equirect coordinates: row 0 = +z, centre column = +x (envmap.py:163-169)."""
  phi = (u - 0.5) * (2.0 * math.pi)
  theta = v * math.pi
  s = np.sin(theta)
  return np.stack([s * np.cos(phi), s * np.sin(phi), np.cos(theta)], axis=-1)


def _skybox_faces(mjm, tex_id: int) -> np.ndarray:
  """(6, W, W, 3) faces of a skybox texture in [0, 1], order +X, -X, +Y, -Y, +Z, -Z.

  MuJoCo stores a cube texture as its six square faces stacked along the rows; a texture that is
  a single square image is used for every face.
  """
  w, h = int(mjm.tex_width[tex_id]), int(mjm.tex_height[tex_id])
  nc = int(mjm.tex_nchannel[tex_id])
  adr = int(mjm.tex_adr[tex_id])
  img = np.asarray(mjm.tex_data[adr : adr + w * h * nc], dtype=np.float32).reshape(h, w, nc) / 255.0
  img = np.repeat(img, 3, axis=2) if nc == 1 else img[..., :3]
  if h == 6 * w:
    return img.reshape(6, w, w, 3)
  if h != w:
    raise ValueError(f"skybox texture {tex_id} must be square or six stacked square faces, got {w}x{h}")
  return np.broadcast_to(img, (6, w, w, 3))


def _sample_cube(faces: np.ndarray, d: np.ndarray) -> np.ndarray:
  """Bilinear lookup of cube-map faces along directions `d` (..., 3) in the cube's own frame.

  Face selection and face coordinates follow the OpenGL cube-map convention (the major axis picks
  the face; s, t in [0, 1] with t = 0 on the face's first row), clamped at the face edges.
  """
  x, y, z = d[..., 0], d[..., 1], d[..., 2]
  ax, ay, az = np.abs(x), np.abs(y), np.abs(z)
  face = np.where(
    (ax >= ay) & (ax >= az), np.where(x >= 0, 0, 1), np.where(ay >= az, np.where(y >= 0, 2, 3), np.where(z >= 0, 4, 5))
  )
  # (sc, tc, ma) per face: +X (-z, -y), -X (z, -y), +Y (x, z), -Y (x, -z), +Z (x, -y), -Z (-x, -y).
  sc = np.choose(face, [-z, z, x, x, x, -x])
  tc = np.choose(face, [-y, -y, z, -z, -y, -y])
  ma = np.maximum(np.choose(face, [ax, ax, ay, ay, az, az]), 1e-30)
  n = faces.shape[1]
  s = np.clip((sc / ma + 1.0) * 0.5 * n - 0.5, 0.0, n - 1.0)
  t = np.clip((tc / ma + 1.0) * 0.5 * n - 0.5, 0.0, n - 1.0)
  s0, t0 = np.floor(s).astype(np.int64), np.floor(t).astype(np.int64)
  s1, t1 = np.minimum(s0 + 1, n - 1), np.minimum(t0 + 1, n - 1)
  fs, ft = (s - s0)[..., None], (t - t0)[..., None]
  top = faces[face, t0, s0] * (1.0 - fs) + faces[face, t0, s1] * fs
  bottom = faces[face, t1, s0] * (1.0 - fs) + faces[face, t1, s1] * fs
  return top * (1.0 - ft) + bottom * ft


def skybox_image(mjm, width: int = SKYBOX_WIDTH) -> np.ndarray | None:
  """The model's first skybox texture as a linear (width // 2, width, 3) equirect image, or None
  if it has none (envmap.py:307-332).

  Texels are display (sRGB) colours and are linearised after the lookup. MuJoCo draws its skybox
  cube turned so that the cube's +y axis points along world +z: world direction (x, y, z) looks
  up cube direction (x, z, -y).
  """
  ids = np.nonzero(np.asarray(mjm.tex_type) == mdl.TEXTURE_SKYBOX)[0] if mjm.ntex else []
  if len(ids) == 0:
    return None
  from mujoco_warp_pbr_amd.material import srgb_to_linear

  faces = _skybox_faces(mjm, int(ids[0]))
  h = width // 2
  v, u = np.meshgrid((np.arange(h) + 0.5) / h, (np.arange(width) + 0.5) / width, indexing="ij")
  world = uv_to_direction(u, v)
  cube = np.stack([world[..., 0], world[..., 2], -world[..., 1]], axis=-1)
  return srgb_to_linear(_sample_cube(faces, cube)).astype(np.float32)


def texel_angle(width: float) -> float:
  """Angular size of one radiance texel at mip 0."""
  return 2.0 * math.pi / width


def _stream(device) -> int:
  dev = torch.device(device)
  if dev.type == "cuda":
    return int(torch.cuda.current_stream(dev).cuda_stream)
  return 0
