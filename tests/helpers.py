"""Scene / context builders and comparisons shared by the tests."""

from __future__ import annotations

import math

import numpy as np
import torch

import mujoco_warp_pbr_amd as mj
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd import scenes
from mujoco_warp_pbr_amd.scenes import _Builder

EV100_UNIT = math.log2(1.0 / 1.2)


def make(mjm, nworld=1, device="cpu", cam_res=(32, 32), **kw):
  """(m, d, rc) on `device`; on the CPU the GPU-built tables come from the oracle."""
  import oracle

  kw.setdefault("ev100", EV100_UNIT)
  kw.setdefault("render_hdr", True)
  m = mj.put_model(mjm, device)
  d = mj.put_data(mjm, nworld, device)
  cpu = torch.device(device).type != "cuda"
  rc = mj.create_render_context(mjm, nworld=nworld, cam_res=cam_res, device=device, precompute=not cpu, **kw)
  if cpu:
    oracle.fill_tables(rc, kw.pop("_dfg_samples", 256), 64)
  return m, d, rc


def plane_and_light(light_type, pos=(0, 0, 2.0), direction=(0, 0, -1), rgba=(0.8, 0.8, 0.8, 1.0), intensity=1.0, res=8,
                    cam_height=1.0, fovy=10.0, range_=100.0):
  """Lambertian plane (roughness 1, reflectance 0) lit by one light, camera looking straight down."""
  b = _Builder()
  mat = b.material(rgba=rgba, metallic=0.0, roughness=1.0)
  b.geom(mdl.GEOM_PLANE, (5, 5, 0.1), (0, 0, 0), matid=mat)
  b.light(light_type, pos, direction, intensity=intensity, castshadow=0, range_=range_)
  b.camera((0, 0, cam_height), (0, 0.0001, 0), fovy, (res, res))
  mjm = b.build()
  mjm.cam_xmat[0] = np.eye(3, dtype=np.float32).reshape(9)  # exactly straight down
  mjm.mat_specular[:] = 0.0
  return mjm


def hdr_image(rc, cam=0):
  w, h = (int(x) for x in rc.base.cam_res[cam])
  adr = int(rc.base.rgb_adr[cam])
  return rc.hdr_data[:, adr : adr + w * h].reshape(-1, h, w, 3).cpu().numpy()


def compare_rt(gpu: dict, ref: dict, hdr_atol=1e-3, depth_rtol=1e-5):
  """Parity of a GPU render with the oracle: seg bit-exact, depth rtol, HDR atol, RGB8 +-1."""
  np.testing.assert_array_equal(gpu["seg"], ref["seg"])
  dg, dr = gpu["depth"], ref["depth"]
  np.testing.assert_allclose(dg, dr, rtol=depth_rtol, atol=1e-7)
  np.testing.assert_allclose(gpu["hdr"], ref["hdr"], atol=hdr_atol, rtol=0)
  a = gpu["rgb"].view(np.uint32)
  b = ref["rgb"].view(np.uint32)
  for sh in (16, 8, 0):
    ca = ((a >> sh) & 0xFF).astype(np.int32)
    cb = ((b >> sh) & 0xFF).astype(np.int32)
    assert np.abs(ca - cb).max() <= 1


def gpu_outputs(rc) -> dict:
  torch.cuda.synchronize()
  base = rc.base
  return {
    "rgb": base.rgb_data.cpu().numpy().view(np.uint32),
    "depth": base.depth_data.cpu().numpy(),
    "seg": base.seg_data.cpu().numpy(),
    "hdr": rc.hdr_data.cpu().numpy(),
    "albedo": rc.guide_albedo.cpu().numpy(),
    "normal": rc.guide_normal.cpu().numpy(),
  }


def xyaxes(x, y) -> np.ndarray:
  """Row-major camera frame of MJCF `xyaxes`: columns x, y (orthogonalised), z = x cross y."""
  x = np.asarray(x, np.float64) / np.linalg.norm(x)
  y = np.asarray(y, np.float64)
  y = y - x * np.dot(x, y)
  y /= np.linalg.norm(y)
  return np.stack([x, y, np.cross(x, y)], axis=1).astype(np.float32).reshape(9)


def channels(rgb: np.ndarray) -> np.ndarray:
  """Packed 0xAARRGGBB (uint32) -> (..., 3) floats in [0, 1], like the reference's get_rgb."""
  rgb = np.asarray(rgb).view(np.uint32)
  return np.stack([(rgb >> s) & 0xFF for s in (16, 8, 0)], axis=-1).astype(np.float32) / 255.0


class Shot:
  """One rendered frame, like the reference test fixture's Scene (tests/conftest.py:35-58).

  On the CPU the oracle renders; on a GPU the HIP renderer does.
  """

  def __init__(self, mjm, device="cpu", nworld=1, cam_res=(32, 32), **kw):
    self.mjm = mjm
    self.m, self.d, self.rc = make(mjm, nworld, device=device, cam_res=cam_res, **kw)
    self.out = None

  @property
  def on_gpu(self) -> bool:
    return bool(self.d.geom_xpos.is_cuda)

  def render(self, seed: int = 0) -> "Shot":
    import oracle

    if self.on_gpu:
      mj.refit_bvh(self.m, self.d, self.rc)
      mj.render(self.m, self.d, self.rc, seed=seed)
      self.out = gpu_outputs(self.rc)
    else:
      from mujoco_warp_pbr_amd.render import check_env_id
      from mujoco_warp_pbr_amd.render import warn_ignored_overrides

      check_env_id(self.rc)
      warn_ignored_overrides(self.m, self.rc)
      fn = oracle.pt_render if self.rc.tier == mj.Tier.PATH else oracle.rt_render
      self.out = fn(self.m, self.d, self.rc, seed=seed)
    return self

  def _image(self, key, cam):
    w, h = (int(x) for x in self.rc.base.cam_res[cam])
    adr = int(self.rc.base.rgb_adr[cam])
    return self.out[key][:, adr : adr + w * h].reshape(-1, h, w, *self.out[key].shape[2:])

  def rgb(self, cam: int = 0) -> np.ndarray:
    return channels(self._image("rgb", cam))

  def hdr(self, cam: int = 0) -> np.ndarray:
    return self._image("hdr", cam)

  @property
  def exposure(self) -> float:
    return float(self.rc.exposure.reshape(-1)[0])


def many_cameras_scene(ncam: int):
  """Plane, a sphere, a box and a capsule watched by `ncam` cameras on a circle (their sizes vary:
  8x8 to 16x12), a sun and a point light."""
  b = _Builder()
  b.geom(mdl.GEOM_PLANE, (2, 2, 0.1), (0, 0, 0), matid=b.material(rgba=(0.8, 0.8, 0.8, 1), roughness=0.7))
  b.geom(mdl.GEOM_SPHERE, (0.2, 0, 0), (0.0, 0.0, 0.2), matid=b.material(rgba=(0.9, 0.3, 0.2, 1), roughness=0.4))
  b.geom(mdl.GEOM_BOX, (0.15, 0.1, 0.1), (0.4, 0.1, 0.1), matid=b.material(rgba=(0.2, 0.6, 0.9, 1), roughness=0.5, metallic=1.0))
  b.geom(mdl.GEOM_CAPSULE, (0.08, 0.15, 0), (-0.3, -0.2, 0.25), matid=b.material(rgba=(0.3, 0.9, 0.3, 1), roughness=0.3))
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (-0.4, 0.3, -0.86))
  b.light(mdl.LIGHT_POINT, (0.6, -0.5, 1.5), (0, 0, -1))
  for i in range(ncam):
    a = 2 * np.pi * i / ncam
    b.camera((2.0 * np.cos(a), 2.0 * np.sin(a), 1.2), (0, 0, 0.15), 40.0 + i, (8 + 8 * (i % 2), 8 + 4 * (i % 3 == 0)))
  return b.build()
