"""Environment maps (reference envmap.py:57-576): Radiance RGBE reading, and the device tables the
renderer samples, all built on the GPU from the uploaded radiance image:

- an RGBA16F radiance mip chain (background, sharp reflections);
- an 8-level GGX-prefiltered stack whose depth axis is perceptual roughness (split-sum IBL);
- nine SH coefficients of the Lambertian-convolved radiance (diffuse IBL);
- piecewise-constant importance-sampling tables (path tracer).

All maps of a context share the prefiltered-stack and sampling-table shapes. A context created
with precompute=False (host-side harnesses on device="cpu") gets the buffers zero-filled.
"""

from __future__ import annotations

import ctypes
import dataclasses
import math
from pathlib import Path

import numpy as np
import torch

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd import model as mdl

PREFILTER_WIDTH = 256
PREFILTER_LEVELS = 8
PREFILTER_SAMPLES = 512
MAX_RADIANCE_WIDTH = 2048
CDF_WIDTH = 256
MAX_LEVELS = 16


# ----------------------------------------------------------------------------- RGBE files
class _RgbeFile:
  """A Radiance picture held in memory: the '#?' signature, header lines up to a blank line, a
  '-Y H +X W' resolution line, then H scanlines of W RGBE pixels, each either flat (4 bytes a
  pixel) or in the adaptive run-length form (2, 2, W >> 8, W & 255, then the R, G, B and E planes
  as packets: n > 128 repeats the next byte n - 128 times, 0 < n <= 128 copies n literal bytes).
  Error messages are those of the reference's loader (its tests pin them)."""

  def __init__(self, path):
    self.name = str(path)
    self.buf = memoryview(Path(path).read_bytes())
    self.pos = 0

  def _error(self, what: str) -> ValueError:
    return ValueError(f"{self.name}: {what}")

  def resolution(self) -> tuple[int, int]:
    raw = self.buf.tobytes()
    if not raw.startswith(b"#?"):
      raise self._error("not a Radiance .hdr file (missing '#?' signature)")
    blank = raw.find(b"\n\n")
    eol = raw.find(b"\n", blank + 2) if blank >= 0 else -1
    if blank < 0 or eol < 0:
      raise self._error("truncated .hdr header")
    fields = raw[blank + 2 : eol].split()
    if len(fields) != 4 or fields[0] != b"-Y" or fields[2] != b"+X" or not (fields[1].isdigit() and fields[3].isdigit()):
      raise self._error("only '-Y H +X W' oriented .hdr files are supported")
    self.pos = eol + 1
    return int(fields[1]), int(fields[3])

  def _run_length_start(self, width: int) -> bool:
    head = self.buf[self.pos : self.pos + 3]
    return 8 <= width <= 0x7FFF and len(head) == 3 and head[0] == 2 and head[1] == 2 and not head[2] & 0x80

  def _flat(self, y: int, width: int) -> np.ndarray:
    end = self.pos + 4 * width
    if end > len(self.buf):
      raise self._error(f"truncated pixel data in scanline {y}")
    px = np.frombuffer(self.buf[self.pos : end], np.uint8).reshape(width, 4)
    self.pos = end
    return px

  def _plane(self, y: int, width: int) -> np.ndarray:
    plane = np.empty(width, np.uint8)
    filled, buf, size = 0, self.buf, len(self.buf)
    while filled < width:
      if self.pos >= size:
        raise self._error(f"truncated run-length data in scanline {y}")
      code = buf[self.pos]
      self.pos += 1
      run = code - 128 if code > 128 else code
      repeat = code > 128
      if run == 0 or run > width - filled or self.pos + (1 if repeat else run) > size:
        raise self._error(f"bad run-length data in scanline {y}")
      if repeat:
        plane[filled : filled + run] = buf[self.pos]
        self.pos += 1
      else:
        plane[filled : filled + run] = np.frombuffer(buf[self.pos : self.pos + run], np.uint8)
        self.pos += run
      filled += run
    return plane

  def _run_length(self, y: int, width: int) -> np.ndarray:
    stated = (self.buf[self.pos + 2] << 8) | self.buf[self.pos + 3] if self.pos + 3 < len(self.buf) else -1
    if stated != width:
      raise self._error(f"scanline {y} width does not match the header")
    self.pos += 4
    return np.stack([self._plane(y, width) for _ in range(4)], axis=1)

  def pixels(self) -> np.ndarray:
    height, width = self.resolution()
    out = np.empty((height, width, 4), np.uint8)
    for y in range(height):
      out[y] = self._run_length(y, width) if self._run_length_start(width) else self._flat(y, width)
    return out


def load_hdr(path) -> np.ndarray:
  """Reads a Radiance RGBE (.hdr) file as float32 (height, width, 3) radiance (reference
  envmap.py:106-142): channel value = byte * 2^(E - 136), zero where E = 0.

  Raises:
    ValueError: not a '-Y H +X W' RGBE file, or truncated / malformed pixel data.
  """
  rgbe = _RgbeFile(path).pixels()
  e = rgbe[..., 3:4].astype(np.int32)
  scale = np.ldexp(np.float32(1.0), e - 136).astype(np.float32) * (e > 0)
  return rgbe[..., :3].astype(np.float32) * scale


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


# ----------------------------------------------------------------------------- device tables
def halve_to_width(image: np.ndarray, max_width: int) -> np.ndarray:
  """2x2 block means (odd trailing rows / columns dropped) until the width is at most max_width
  (the reference loads maps wider than 2048 texels this way, envmap.py:145-151)."""
  img = np.asarray(image)
  while img.shape[1] > max_width:
    h2, w2 = img.shape[0] // 2, img.shape[1] // 2
    quads = img[: 2 * h2, : 2 * w2].astype(np.float64)
    img = 0.25 * (quads[0::2, 0::2] + quads[0::2, 1::2] + quads[1::2, 0::2] + quads[1::2, 1::2])
  return img


def mip_shapes(width: int, height: int) -> list[tuple[int, int]]:
  """(width, height) of every level of a mip chain down to 1 x 1 (an axis of one texel stays)."""
  shapes = [(width, height)]
  while shapes[-1] != (1, 1) and len(shapes) < MAX_LEVELS:
    w, h = shapes[-1]
    shapes.append((max(w // 2, 1), max(h // 2, 1)))
  return shapes


def table_shape(width: int, height: int, cdf_width: int = CDF_WIDTH) -> tuple[int, int]:
  """(height, width) of the sampling tables of a map: halved while wider than cdf_width."""
  while width > cdf_width:
    width, height = width // 2, height // 2
  return height, width


@dataclasses.dataclass
class Environments:
  """Device resources of the environment maps (reference envmap.py:493-515)."""

  nenv: int
  radiance: torch.Tensor  # uint16 (f16 bits) pool of every map's mip chain
  level_adr: torch.Tensor  # (nenv, MAX_LEVELS) int64, offsets in halves
  desc: torch.Tensor  # (nenv, 4) int32: width, height, nlevels, format (1: RGBA16F)
  prefiltered: torch.Tensor  # (nenv, L, H, W, 4) f16 bits
  sh: torch.Tensor  # (nenv, 9, 3)
  width: float
  prob: torch.Tensor  # (nenv, h, w)
  marginal: torch.Tensor  # (nenv, h + 1)
  conditional: torch.Tensor  # (nenv, h, w + 1)
  images: list  # the float32 radiance images the tables were built from
  table_images: list  # the images the sampling tables were built from (resized to one shape)

  @property
  def pf_shape(self):
    return tuple(self.prefiltered.shape[1:4])


def _resample_nearest(image: np.ndarray, height: int, width: int) -> np.ndarray:
  rows = np.minimum(((np.arange(height) + 0.5) * image.shape[0] / height).astype(int), image.shape[0] - 1)
  cols = np.minimum(((np.arange(width) + 0.5) * image.shape[1] / width).astype(int), image.shape[1] - 1)
  return image[rows][:, cols]


def build_environments(images: list[np.ndarray], device, nsample: int = PREFILTER_SAMPLES, precompute: bool = True) -> Environments:
  """Device resources of equirect radiance maps. With precompute (a GPU device) the radiance mip
  chains, SH9 coefficients, sampling tables and prefiltered stacks are built by HIP kernels from
  the uploaded images; otherwise they are left zero for the caller to fill."""
  dev = torch.device(device)
  imgs = [np.ascontiguousarray(halve_to_width(np.asarray(im)[..., :3], MAX_RADIANCE_WIDTH), np.float32) for im in images]
  nenv = len(imgs)
  # sampling tables share the first map's shape; other maps are resampled to it first
  th, tw = table_shape(imgs[0].shape[1], imgs[0].shape[0])
  table_imgs = [im if table_shape(im.shape[1], im.shape[0]) == (th, tw) else _resample_nearest(im, th, tw) for im in imgs]
  # radiance pool layout
  level_adr = np.zeros((nenv, MAX_LEVELS), np.int64)
  desc = np.zeros((nenv, 4), np.int32)
  offset = 0
  for e, im in enumerate(imgs):
    shapes = mip_shapes(im.shape[1], im.shape[0])
    desc[e] = (im.shape[1], im.shape[0], len(shapes), 1)
    for lv, (w, h) in enumerate(shapes):
      level_adr[e, lv] = offset
      offset += 4 * w * h
  pw = max(min(PREFILTER_WIDTH, min(im.shape[1] for im in imgs)), 2)
  env = Environments(
    nenv=nenv,
    radiance=torch.zeros(offset, dtype=torch.int16, device=dev),
    level_adr=torch.as_tensor(level_adr).to(dev),
    desc=torch.as_tensor(desc).to(dev),
    prefiltered=torch.zeros((nenv, PREFILTER_LEVELS, pw // 2, pw, 4), dtype=torch.int16, device=dev),
    sh=torch.zeros((nenv, 9, 3), dtype=torch.float32, device=dev),
    width=float(max(im.shape[1] for im in imgs)),
    prob=torch.zeros((nenv, th, tw), dtype=torch.float32, device=dev),
    marginal=torch.zeros((nenv, th + 1), dtype=torch.float32, device=dev),
    conditional=torch.zeros((nenv, th, tw + 1), dtype=torch.float32, device=dev),
    images=imgs,
    table_images=table_imgs,
  )
  if precompute:
    _build_on_device(env, level_adr, nsample)
  return env


def _build_on_device(env: Environments, level_adr: np.ndarray, nsample: int) -> None:
  dev = env.radiance.device
  stream = _stream(dev)
  largest = max(im.shape[0] * im.shape[1] for im in env.images + env.table_images)
  maxh = max(im.shape[0] for im in env.images)
  work = torch.empty(max(8 * largest, 27 * maxh, 1), dtype=torch.float64, device=dev)
  base = env.radiance.data_ptr()
  for e, (im, tim) in enumerate(zip(env.images, env.table_images)):
    h, w = im.shape[:2]
    nlevels = int(env.desc[e, 2])
    src = torch.as_tensor(im).to(dev)
    adr = (ctypes.c_int64 * nlevels)(*(int(x) for x in level_adr[e, :nlevels]))  # read on the host by the launcher
    _abi.call("mjpbr_radiance_mips", _abi.ptr(src), w, h, base, ctypes.cast(adr, ctypes.c_void_p), nlevels, _abi.ptr(work), work.numel(),
              stream)
    _abi.call("mjpbr_env_sh9", _abi.ptr(src), w, h, _abi.ptr(work), work.numel(), _abi.ptr(env.sh[e]), stream)
    tsrc = src if tim is im else torch.as_tensor(np.ascontiguousarray(tim, np.float32)).to(dev)
    _abi.call(
      "mjpbr_env_sampling_tables",
      _abi.ptr(tsrc),
      int(tim.shape[1]),
      int(tim.shape[0]),
      CDF_WIDTH,
      _abi.ptr(work),
      work.numel(),
      _abi.ptr(env.prob[e]),
      _abi.ptr(env.marginal[e]),
      _abi.ptr(env.conditional[e]),
      stream,
    )
  L, H, W = env.pf_shape
  for e in range(env.nenv):
    _abi.call(
      "mjpbr_env_prefilter",
      _abi.ptr(env.radiance),
      _abi.ptr(env.level_adr[e]),
      _abi.ptr(env.desc[e]),
      _abi.ptr(env.prefiltered[e]),
      L,
      H,
      W,
      nsample,
      stream,
    )


def empty_environments(device) -> Environments:
  """Placeholder resources of a context without environment map (constant ambient)."""
  dev = torch.device(device)
  one = np.zeros((1, 1, 3), np.float32)
  return Environments(
    nenv=1,
    radiance=torch.tensor([0, 0, 0, 15360], dtype=torch.int16, device=dev),  # one texel (0, 0, 0, 1)
    level_adr=torch.zeros((1, MAX_LEVELS), dtype=torch.int64, device=dev),
    desc=torch.tensor([[1, 1, 1, 1]], dtype=torch.int32, device=dev),
    prefiltered=torch.zeros((1, 2, 1, 2, 4), dtype=torch.int16, device=dev),
    sh=torch.zeros((1, 9, 3), dtype=torch.float32, device=dev),
    width=1.0,
    prob=torch.ones((1, 1, 1), dtype=torch.float32, device=dev),
    marginal=torch.tensor([[0.0, 1.0]], dtype=torch.float32, device=dev),
    conditional=torch.tensor([[[0.0, 1.0]]], dtype=torch.float32, device=dev),
    images=[],
    table_images=[],
  )


# ----------------------------------------------------------------------------- skybox
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
