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
      vp = ctypes.c_void_p
      lib.oracle_rt_render.argtypes = [vp, vp, vp, ctypes.c_int32, ctypes.c_int, ctypes.c_int]
      lib.oracle_rt_render.restype = ctypes.c_int
      lib.oracle_dfg_lut.argtypes = [vp, ctypes.c_int, ctypes.c_int]
      lib.oracle_prefilter_texel.argtypes = [vp, vp, vp] + [ctypes.c_int] * 7 + [vp]
      lib.oracle_closest.argtypes = [vp, vp, ctypes.c_int, vp, vp, ctypes.c_float, vp, vp]
      for name in ("oracle_pt_render", "oracle_atrous"):
        if hasattr(lib, name):
          getattr(lib, name).restype = ctypes.c_int
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

  hm, hd, hrc = to_host(m, d, rc)
  for t in (hrc.base.rgb_data, hrc.base.depth_data, hrc.base.seg_data, hrc.hdr_data, hrc.guide_albedo, hrc.guide_normal):
    t.zero_()
  sc = scene_struct(hm, hd, hrc)
  s = settings_struct(hrc)
  out = outputs_struct(hrc)
  lib = library()
  w0, w1 = worlds if worlds is not None else (0, hd.nworld)
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

  hm, hd, hrc = to_host(m, d, rc)
  for t in (hrc.base.rgb_data, hrc.base.depth_data, hrc.base.seg_data, hrc.hdr_data, hrc.guide_albedo, hrc.guide_normal):
    t.zero_()
  sc = scene_struct(hm, hd, hrc)
  s = settings_struct(hrc)
  out = outputs_struct(hrc)
  lib = library()
  vp = ctypes.c_void_p
  lib.oracle_pt_render.argtypes = [vp, vp, vp, ctypes.c_int32, ctypes.c_int, ctypes.c_int, vp]
  w0, w1 = worlds if worlds is not None else (0, hd.nworld)
  samples = None
  if per_sample:
    samples = np.zeros((w1 - w0, hrc.base.total_rays, hrc.path_samples, 3), np.float32)
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
    _run_atrous(hrc, out)
  result.update(_outputs(hrc))
  if samples is not None:
    result["samples"] = samples
  return result


def _run_atrous(hrc, out) -> None:
  from mujoco_warp_pbr_amd.denoise import atrous_params

  lib = library()
  vp = ctypes.c_void_p
  lib.oracle_atrous.argtypes = [vp, vp, vp, vp]
  p = atrous_params(hrc, hrc.base.nworld)
  n = hrc.base.nworld * int(hrc.base.rgb_data.shape[1])
  w0 = np.zeros((n, 4), np.float32)
  w1 = np.zeros((n, 4), np.float32)
  lib.oracle_atrous(vp(ctypes.addressof(p)), vp(ctypes.addressof(out)), w0.ctypes.data, w1.ctypes.data)


def atrous(rc, hdr: np.ndarray, aov: dict, render_hdr: bool = True) -> dict:
  """The oracle's a-trous filter of a context's layout on the given noisy HDR and features
  (host arrays as returned by the renderers); returns rgb and (filtered) hdr."""
  from mujoco_warp_pbr_amd.render import outputs_struct

  (hrc,) = to_host(rc)
  hrc.render_hdr = render_hdr
  hrc.hdr_data.copy_(torch.from_numpy(np.ascontiguousarray(hdr, np.float32)))
  hrc.aov_albedo.copy_(torch.from_numpy(np.ascontiguousarray(aov["aov_albedo"], np.float16).view(np.int16)))
  hrc.aov_normal.copy_(torch.from_numpy(np.ascontiguousarray(aov["aov_normal"], np.float16).view(np.int16)))
  for k in ("aov_depth", "aov_emission", "aov_variance"):
    getattr(hrc, k).copy_(torch.from_numpy(np.ascontiguousarray(aov[k], np.float32)))
  out = outputs_struct(hrc)
  _run_atrous(hrc, out)
  return {"rgb": hrc.base.rgb_data.numpy().view(np.uint32).copy(), "hdr": hrc.hdr_data.numpy().copy()}


def dfg_lut(size: int = 128, nsample: int = 1024) -> np.ndarray:
  lut = np.zeros((size, size, 2), np.float32)
  library().oracle_dfg_lut(lut.ctypes.data, size, nsample)
  return lut


def fill_tables(rc, dfg_samples: int = 1024, prefilter_samples: int = 512) -> None:
  """Fills a host context made with precompute=False: DFG LUT and prefiltered stacks (oracle)."""
  n = int(rc.dfg_lut.shape[0])
  rc.dfg_lut.copy_(torch.from_numpy(dfg_lut(n, dfg_samples)))
  env = rc.env
  if not env.images:
    return
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
