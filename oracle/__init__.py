"""CPU oracle of the renderer — TEST INFRASTRUCTURE ONLY.

A plain-C restatement of the reference renderer (julien-blanchon/mujoco_warp_pbr,
mujoco_warp_pbr/_src/render.py, pathtrace.py, denoise.py and the modules they call), compiled
with gcc from the sources in this folder (`make -C oracle`). Only tests/, bench.py's
cpu_baseline leg and __graft_entry__.smoke() may use it, and only as the checker: the product
(mujoco_warp_pbr_amd) never imports it.

`rt_render(m, d, rc, seed)` renders the same scene as `mujoco_warp_pbr_amd.render` with every
ray query brute force (no BVH) on the host and returns the outputs as NumPy arrays.
"""

from __future__ import annotations

import copy
import ctypes
import hashlib
import dataclasses
import subprocess
import threading
from pathlib import Path

import numpy as np
import torch

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libmjpbr_oracle.so"
_lib = None
_lock = threading.Lock()


def build() -> Path:
  """Compiles the oracle library with gcc (make)."""
  subprocess.run(["make", "-C", str(HERE), "-s"], check=True)
  return LIB


def library() -> ctypes.CDLL:
  global _lib
  with _lock:
    if _lib is None:
      if not LIB.exists():
        build()
      lib = ctypes.CDLL(str(LIB))
      lib.oracle_header_hash.restype = ctypes.c_char_p
      want = hashlib.sha256((HERE.parent / "include" / "mjpbr.h").read_bytes()).hexdigest()[:16]
      if lib.oracle_header_hash().decode() != want:
        raise RuntimeError(f"{LIB} was built against another include/mjpbr.h: rebuild it (make -C oracle)")
      vp = ctypes.c_void_p
      lib.oracle_rt_render.argtypes = [vp, vp, vp, ctypes.c_int32, ctypes.c_int, ctypes.c_int]
      lib.oracle_rt_render.restype = ctypes.c_int
      lib.oracle_dfg_lut.argtypes = [vp, ctypes.c_int, ctypes.c_int]
      lib.oracle_prefilter_texel.argtypes = [vp, vp, vp] + [ctypes.c_int] * 7 + [vp]
      lib.oracle_closest.argtypes = [vp, vp, ctypes.c_int, vp, vp, ctypes.c_float, vp, vp]
      for name in ("oracle_pt_render", "oracle_atrous"):
        if hasattr(lib, name):
          getattr(lib, name).restype = ctypes.c_int
      f = ctypes.c_float
      lib.oracle_ray_primitive.argtypes = [ctypes.c_int] + [vp] * 6
      lib.oracle_ray_primitive.restype = f
      lib.oracle_ray_triangle.argtypes = [vp] * 5 + [f, vp]
      lib.oracle_ray_triangle.restype = ctypes.c_int
      lib.oracle_display.argtypes = [vp, ctypes.c_int, vp]
      lib.oracle_encode.argtypes = [vp]
      lib.oracle_encode.restype = ctypes.c_uint32
      lib.oracle_spec_ao.argtypes = [f, f, f]
      lib.oracle_spec_ao.restype = f
      lib.oracle_eval_light.argtypes = [ctypes.c_int] * 3 + [vp, vp, f, f, vp, f, f, vp, vp]
      lib.oracle_atrous_pass.argtypes = [vp, vp, vp, vp, ctypes.c_int]
      lib.oracle_spec_lobe.argtypes = [vp, f, f, f, f, f, vp]
      lib.oracle_dfg_lookup.argtypes = [vp, ctypes.c_int, f, f, vp]
      lib.oracle_lobes.argtypes = [vp, f, f, f, f, vp, ctypes.c_int, vp]
      lib.oracle_atrous_pass.restype = ctypes.c_int
      _lib = lib
    return _lib


def _move(obj, device, memo):
  if isinstance(obj, torch.Tensor):
    return obj.to(device)
  if id(obj) in memo:
    return memo[id(obj)]
  if dataclasses.is_dataclass(obj) and not isinstance(obj, type):
    new = copy.copy(obj)
    memo[id(obj)] = new
    for f in dataclasses.fields(obj):
      setattr(new, f.name, _move(getattr(obj, f.name), device, memo))
    return new
  if isinstance(obj, tuple):
    return tuple(_move(x, device, memo) for x in obj)
  if isinstance(obj, dict):
    return {k: _move(v, device, memo) for k, v in obj.items()}
  if hasattr(obj, "__dict__") and type(obj).__module__.startswith("mujoco_warp_pbr_amd"):
    new = copy.copy(obj)
    memo[id(obj)] = new
    for k, v in vars(obj).items():
      setattr(new, k, _move(v, device, memo))
    return new
  return obj


def to_host(*objs):
  """CPU copies of model / data / context objects (their tensors moved to the host)."""
  memo: dict = {}
  return tuple(_move(o, "cpu", memo) for o in objs)


def _outputs(rc) -> dict[str, np.ndarray]:
  base = rc.base
  return {
    "rgb": base.rgb_data.numpy().view(np.uint32).copy(),
    "depth": base.depth_data.numpy().copy(),
    "seg": base.seg_data.numpy().copy(),
    "hdr": rc.hdr_data.numpy().copy(),
    "albedo": rc.guide_albedo.numpy().copy(),
    "normal": rc.guide_normal.numpy().copy(),
  }


def rt_render(m, d, rc, seed: int = 0, threads: int = 1, worlds: tuple[int, int] | None = None) -> dict[str, np.ndarray]:
  """Real-time tier on the CPU oracle. The context's BVH must have been built (it is not read,
  but the enabled-geom list and mesh tables are shared). Returns numpy copies of the outputs."""
  from mujoco_warp_pbr_amd.render import outputs_struct
  from mujoco_warp_pbr_amd.render import scene_struct
  from mujoco_warp_pbr_amd.render import settings_struct

  from mujoco_warp_pbr_amd.render import camera_chunks

  hm, hd, hrc = to_host(m, d, rc)
  for t in (hrc.base.rgb_data, hrc.base.depth_data, hrc.base.seg_data, hrc.hdr_data, hrc.guide_albedo, hrc.guide_normal):
    t.zero_()
  s = settings_struct(hrc)
  lib = library()
  w0, w1 = worlds if worlds is not None else (0, hd.nworld)
  for cams in camera_chunks(hrc):  # the renderer's launches of at most MJPBR_MAX_CAMS cameras
    sc = scene_struct(hm, hd, hrc, cams)
    out = outputs_struct(hrc, cams)
    args = (ctypes.c_void_p(ctypes.addressof(sc)), ctypes.c_void_p(ctypes.addressof(s)), ctypes.c_void_p(ctypes.addressof(out)))
    if threads <= 1:
      lib.oracle_rt_render(*args, int(seed), w0, w1)
    else:
      bounds = np.linspace(w0, w1, threads + 1).astype(int)
      pool = [threading.Thread(target=lib.oracle_rt_render, args=(*args, int(seed), int(a), int(b))) for a, b in zip(bounds[:-1], bounds[1:])]
      for th in pool:
        th.start()
      for th in pool:
        th.join()
  return _outputs(hrc)


def _aov(rc) -> dict[str, np.ndarray]:
  return {
    "aov_albedo": rc.aov_albedo.numpy().view(np.float16).astype(np.float32).copy(),
    "aov_normal": rc.aov_normal.numpy().view(np.float16).astype(np.float32).copy(),
    "aov_depth": rc.aov_depth.numpy().copy(),
    "aov_emission": rc.aov_emission.numpy().copy(),
    "aov_variance": rc.aov_variance.numpy().copy(),
  }


def pt_render(m, d, rc, seed: int = 0, threads: int = 1, worlds: tuple[int, int] | None = None, per_sample: bool = False) -> dict:
  """Path-traced tier on the CPU oracle. per_sample=True also returns every sample's radiance,
  (nworld, total_rays, path_samples, 3) in camera-concatenated row-major pixel order."""
  from mujoco_warp_pbr_amd.render import outputs_struct
  from mujoco_warp_pbr_amd.render import scene_struct
  from mujoco_warp_pbr_amd.render import settings_struct

  from mujoco_warp_pbr_amd.render import camera_chunks

  hm, hd, hrc = to_host(m, d, rc)
  for t in (hrc.base.rgb_data, hrc.base.depth_data, hrc.base.seg_data, hrc.hdr_data, hrc.guide_albedo, hrc.guide_normal):
    t.zero_()
  s = settings_struct(hrc)
  lib = library()
  vp = ctypes.c_void_p
  lib.oracle_pt_render.argtypes = [vp, vp, vp, ctypes.c_int32, ctypes.c_int, ctypes.c_int, vp]
  w0, w1 = worlds if worlds is not None else (0, hd.nworld)
  samples = None
  if per_sample:
    samples = np.zeros((w1 - w0, hrc.base.total_rays, hrc.path_samples, 3), np.float32)
  for cams in camera_chunks(hrc):  # the renderer's launches of at most MJPBR_MAX_CAMS cameras
    sc = scene_struct(hm, hd, hrc, cams)
    out = outputs_struct(hrc, cams)
    args = (vp(ctypes.addressof(sc)), vp(ctypes.addressof(s)), vp(ctypes.addressof(out)))

    def run(a, b):
      ptr = None if samples is None else vp(samples.ctypes.data + (a - w0) * samples[0].nbytes)
      lib.oracle_pt_render(*args, int(seed), int(a), int(b), ptr)

    if threads <= 1:
      run(w0, w1)
    else:
      bounds = np.linspace(w0, w1, threads + 1).astype(int)
      pool = [threading.Thread(target=run, args=(int(a), int(b))) for a, b in zip(bounds[:-1], bounds[1:]) if b > a]
      for th in pool:
        th.start()
      for th in pool:
        th.join()
  result = _aov(hrc)
  result["noisy_hdr"] = hrc.hdr_data.numpy().copy()
  if hrc.denoise:
    _run_atrous(hrc)
  result.update(_outputs(hrc))
  if samples is not None:
    result["samples"] = samples
  return result


def _run_atrous(hrc) -> None:
  from mujoco_warp_pbr_amd.denoise import atrous_params
  from mujoco_warp_pbr_amd.render import camera_chunks
  from mujoco_warp_pbr_amd.render import outputs_struct

  lib = library()
  vp = ctypes.c_void_p
  lib.oracle_atrous.argtypes = [vp, vp, vp, vp]
  n = hrc.base.nworld * int(hrc.base.rgb_data.shape[1])
  w0 = np.zeros((n, 4), np.float32)
  w1 = np.zeros((n, 4), np.float32)
  for cams in camera_chunks(hrc):
    p = atrous_params(hrc, hrc.base.nworld, cams)
    out = outputs_struct(hrc, cams)
    lib.oracle_atrous(vp(ctypes.addressof(p)), vp(ctypes.addressof(out)), w0.ctypes.data, w1.ctypes.data)


def atrous(rc, hdr: np.ndarray, aov: dict, render_hdr: bool = True) -> dict:
  """The oracle's a-trous filter of a context's layout on the given noisy HDR and features
  (host arrays as returned by the renderers); returns rgb and (filtered) hdr."""
  (hrc,) = to_host(rc)
  hrc.render_hdr = render_hdr
  hrc.hdr_data.copy_(torch.from_numpy(np.ascontiguousarray(hdr, np.float32)))
  hrc.aov_albedo.copy_(torch.from_numpy(np.ascontiguousarray(aov["aov_albedo"], np.float16).view(np.int16)))
  hrc.aov_normal.copy_(torch.from_numpy(np.ascontiguousarray(aov["aov_normal"], np.float16).view(np.int16)))
  for k in ("aov_depth", "aov_emission", "aov_variance"):
    getattr(hrc, k).copy_(torch.from_numpy(np.ascontiguousarray(aov[k], np.float32)))
  _run_atrous(hrc)
  return {"rgb": hrc.base.rgb_data.numpy().view(np.uint32).copy(), "hdr": hrc.hdr_data.numpy().copy()}


def dfg_lut(size: int = 128, nsample: int = 1024) -> np.ndarray:
  lut = np.zeros((size, size, 2), np.float32)
  library().oracle_dfg_lut(lut.ctypes.data, size, nsample)
  return lut


def fill_tables(rc, dfg_samples: int = 1024, prefilter_samples: int = 512) -> None:
  """Fills a host context made with precompute=False: DFG LUT, texture mip chains, radiance mip
  chains, SH9 coefficients, sampling tables and prefiltered stacks (oracle)."""
  from oracle import blas
  from oracle import tables

  mb = rc.base.mesh_bvhs
  if mb is not None and mb.faces:
    blas.fill_mesh_bvhs(rc.base, mb.host_vert, mb.faces)
  n = int(rc.dfg_lut.shape[0])
  rc.dfg_lut.copy_(torch.from_numpy(dfg_lut(n, dfg_samples)))
  tables.fill_textures(rc.textures)
  env = rc.env
  if not env.images:
    return
  tables.fill_environments(env)
  L, H, W = (int(v) for v in env.prefiltered.shape[1:4])
  stack = np.zeros((env.nenv, L, H, W, 4), np.float16)
  for e in range(env.nenv):
    for level in range(L):
      for y in range(H):
        for x in range(W):
          stack[e, level, y, x, :3] = prefilter_texel(env, e, level, y, x, prefilter_samples)
          stack[e, level, y, x, 3] = 1.0
  env.prefiltered.copy_(torch.from_numpy(stack.view(np.int16)))


def prefilter_texel(env, e: int, level: int, y: int, x: int, nsample: int) -> np.ndarray:
  """One texel of the GGX-prefiltered stack of environment map e (host copies of the tables)."""
  rad = env.radiance.cpu().contiguous()
  adr = env.level_adr[e].cpu().contiguous()
  desc = env.desc[e].cpu().contiguous()
  L, H, W = (int(v) for v in env.prefiltered.shape[1:4])
  out = np.zeros(3, np.float32)
  library().oracle_prefilter_texel(rad.data_ptr(), adr.data_ptr(), desc.data_ptr(), L, H, W, nsample, level, y, x, out.ctypes.data)
  return out


# ----------------------------------------------------------------------------- scalar building blocks
def _f32(x, n: int | None = None) -> np.ndarray:
  a = np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1))
  if n is not None:
    assert a.size == n, a.size
  return a


def ray_primitive(geom_type: int, pos, mat, size, origin, direction) -> tuple[float, np.ndarray]:
  """Distance (-1: miss) and world normal of one ray against one MuJoCo primitive."""
  n = np.zeros(3, np.float32)
  args = [_f32(pos, 3), _f32(mat, 9), _f32(size, 3), _f32(origin, 3), _f32(direction, 3)]
  x = library().oracle_ray_primitive(int(geom_type), *(a.ctypes.data for a in args), n.ctypes.data)
  return float(x), n


def ray_triangle(origin, direction, v0, v1, v2, tmax: float) -> tuple[bool, float, float, float]:
  """(hit, t, b1, b2) of the oracle's ray-triangle test (b1, b2: weights of v1, v2)."""
  out = np.zeros(3, np.float32)
  args = [_f32(origin, 3), _f32(direction, 3), _f32(v0, 3), _f32(v1, 3), _f32(v2, 3)]
  hit = library().oracle_ray_triangle(*(a.ctypes.data for a in args), float(tmax), out.ctypes.data)
  return bool(hit), float(out[0]), float(out[1]), float(out[2])


def display(color, tone_mapper: int) -> np.ndarray:
  """Exposed linear HDR -> tone-mapped display-linear colour in [0, 1]."""
  c, out = _f32(color, 3), np.zeros(3, np.float32)
  library().oracle_display(c.ctypes.data, int(tone_mapper), out.ctypes.data)
  return out


def encode(color) -> int:
  """Display-linear colour -> packed sRGB8 word."""
  c = _f32(color, 3)
  return int(library().oracle_encode(c.ctypes.data))


def spec_ao(nov: float, ao: float, alpha: float) -> float:
  return float(library().oracle_spec_ao(float(nov), float(ao), float(alpha)))


def eval_light(light_type: int, model: int, has_spot: bool, pos, direction, cutoff: float, exponent: float, attenuation,
               intensity: float, light_range: float, point) -> dict:
  """One light evaluated at `point` (white light colour): direction to the light, distance and
  the diffuse / specular scale."""
  out = np.zeros(10, np.float32)
  a = [_f32(pos, 3), _f32(direction, 3), _f32(attenuation, 3), _f32(point, 3)]
  library().oracle_eval_light(int(light_type), int(model), int(bool(has_spot)), a[0].ctypes.data, a[1].ctypes.data, float(cutoff),
                              float(exponent), a[2].ctypes.data, float(intensity), float(light_range), a[3].ctypes.data,
                              out.ctypes.data)
  return {"l": out[0:3].copy(), "dist": float(out[3]), "diffuse": out[4:7].copy(), "specular": out[7:10].copy()}


def atrous_pass(src: np.ndarray, normal: np.ndarray, depth: np.ndarray, albedo: np.ndarray, step: int,
                sigma=(4.0, 128.0, 0.02, 0.05)) -> np.ndarray:
  """One a-trous pass over one (h, w) image: src (h, w, 4) demodulated colour + variance,
  normal / albedo (h, w, 3) (stored as f16 like the renderer's features), depth (h, w)."""
  from mujoco_warp_pbr_amd import _abi

  h, w = depth.shape
  S = _abi.structs()
  p = S["MjpbrAtrousParams"]()
  p.nworld, p.nrender, p.stride, p.iterations = 1, 1, h * w, 1
  p.cam_res[0][0], p.cam_res[0][1] = w, h
  p.rgb_adr[0] = 0
  p.sigma_color, p.sigma_normal, p.sigma_depth, p.sigma_albedo = (float(x) for x in sigma)
  o = S["MjpbrOutputs"]()
  o.aov_stride = h * w
  nrm = np.ascontiguousarray(np.asarray(normal, np.float16).reshape(-1, 3))
  alb = np.ascontiguousarray(np.asarray(albedo, np.float16).reshape(-1, 3))
  dep = np.ascontiguousarray(np.asarray(depth, np.float32).reshape(-1))
  o.aov_normal = nrm.ctypes.data
  o.aov_albedo = alb.ctypes.data
  o.aov_depth = dep.ctypes.data
  s = np.ascontiguousarray(np.asarray(src, np.float32).reshape(-1, 4))
  dst = np.zeros_like(s)
  vp = ctypes.c_void_p
  library().oracle_atrous_pass(vp(ctypes.addressof(p)), vp(ctypes.addressof(o)), s.ctypes.data, dst.ctypes.data, int(step))
  return dst.reshape(h, w, 4)


def spec_lobe(f0, alpha: float, noh: float, nov: float, nol: float, loh: float) -> np.ndarray:
  """F D V of the specular lobe (no NoL cosine)."""
  a, out = _f32(f0, 3), np.zeros(3, np.float32)
  library().oracle_spec_lobe(a.ctypes.data, float(alpha), float(noh), float(nov), float(nol), float(loh), out.ctypes.data)
  return out


def dfg_lookup(lut: np.ndarray, nov: float, pr: float) -> np.ndarray:
  t = np.ascontiguousarray(lut, np.float32)
  out = np.zeros(2, np.float32)
  library().oracle_dfg_lookup(t.ctypes.data, int(t.shape[0]), float(nov), float(pr), out.ctypes.data)
  return out


def lobes(base_color, metallic: float, reflectance: float, pr: float, nov: float, lut: np.ndarray) -> dict:
  t = np.ascontiguousarray(lut, np.float32)
  b, out = _f32(base_color, 3), np.zeros(14, np.float32)
  library().oracle_lobes(b.ctypes.data, float(metallic), float(reflectance), float(pr), float(nov), t.ctypes.data, int(t.shape[0]),
                         out.ctypes.data)
  return {"pr": out[0], "alpha": out[1], "f0": out[2:5], "energy_comp": out[5:8], "spec_albedo": out[8:11], "diffuse": out[11:14]}
