"""Pins the CPU oracle of the real-time tier against closed forms and invariants.

The reference's own analytic tests (tests/test_render.py:50-146, tests/test_ibl.py,
tests/test_camera_rays.py:47-54) are restated against the oracle here; the GPU renderer is
then compared to the oracle in test_rt_gpu.py.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
import oracle
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd import scenes
from mujoco_warp_pbr_amd.scenes import _Builder
from tests.helpers import EV100_UNIT
from tests.helpers import make
from tests.helpers import plane_and_light


def _centre_hdr(mjm, **kw):
  m, d, rc = make(mjm, cam_res=(8, 8), **kw)
  out = oracle.rt_render(m, d, rc)
  return out["hdr"][0].reshape(8, 8, 3)[4, 4], out, rc


def _srgb(x):
  return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1 / 2.4) - 0.055)


def test_lambertian_plane_point_light():
  """tests/test_render.py:50-61: radiance (albedo / pi) I / d^2 (rtol 2e-3)."""
  intensity, height, albedo = 3.0, 1.5, 0.8
  mjm = plane_and_light(mdl.LIGHT_POINT, pos=(0, 0, height), rgba=(albedo, albedo, albedo, 1), intensity=intensity, range_=1000.0)
  c, _, _ = _centre_hdr(mjm, tone_mapper=mj.ToneMapper.CLAMP)
  np.testing.assert_allclose(c, albedo / np.pi * intensity / height**2, rtol=2e-3)


def test_hdr_holds_physical_radiance():
  """tests/test_render.py:64-72: a 1e6 lux directional light gives E / pi exactly in HDR."""
  mjm = plane_and_light(mdl.LIGHT_DIRECTIONAL, rgba=(1, 1, 1, 1), intensity=1.0e6)
  c, _, _ = _centre_hdr(mjm, ev100=14.96578428466209)
  np.testing.assert_allclose(c, 1.0e6 / np.pi, rtol=2e-3)


def test_fallback_directional_matches_diffuse_colour():
  """tests/test_render.py:75-83: unset intensity shows the light's diffuse colour."""
  mjm = plane_and_light(mdl.LIGHT_DIRECTIONAL, rgba=(1, 1, 1, 1), intensity=0.0)
  mjm.light_diffuse[0] = (0.5, 0.25, 0.125)
  m, d, rc = make(mjm, cam_res=(8, 8), tone_mapper=mj.ToneMapper.CLAMP, ev100=14.96578428466209)
  out = oracle.rt_render(m, d, rc)
  px = out["rgb"][0].reshape(8, 8)[4, 4]
  rgb = np.array([(px >> s) & 0xFF for s in (16, 8, 0)]) / 255.0
  np.testing.assert_allclose(rgb, _srgb(np.array([0.5, 0.25, 0.125])), atol=2.5 / 255)


@pytest.mark.parametrize("light_model", [mj.LightModel.PHYSICAL, mj.LightModel.MUJOCO])
def test_light_models_agree_for_default_lights(light_model):
  """tests/test_render.py:86-92."""
  mjm = plane_and_light(mdl.LIGHT_DIRECTIONAL, rgba=(0.7, 0.7, 0.7, 1), intensity=0.0)
  mjm.light_diffuse[0] = (0.6, 0.6, 0.6)
  m, d, rc = make(mjm, cam_res=(8, 8), tone_mapper=mj.ToneMapper.CLAMP, light_model=light_model, ev100=14.96578428466209)
  out = oracle.rt_render(m, d, rc)
  px = out["rgb"][0].reshape(8, 8)[4, 4]
  rgb = np.array([(px >> s) & 0xFF for s in (16, 8, 0)]) / 255.0
  np.testing.assert_allclose(rgb, _srgb(np.full(3, 0.7 * 0.6)), atol=2.5 / 255)


def _metal_sphere(roughness):
  b = _Builder()
  mat = b.material(rgba=(1, 1, 1, 1), metallic=1.0, roughness=roughness)
  b.geom(mdl.GEOM_SPHERE, (0.5, 0, 0), (0, 0, 0), matid=mat)
  b.camera((0, -3, 0), (0, 0, 0), 5.0, (8, 8))
  return b.build(headlight_active=True, headlight_ambient=np.full(3, 0.5, np.float32), headlight_diffuse=np.zeros(3, np.float32),
                 headlight_specular=np.zeros(3, np.float32))


@pytest.mark.parametrize("roughness,rtol", [(0.0, 0.01), (1.0, 0.02)])
def test_metal_under_ambient_conserves_energy(roughness, rtol):
  """tests/test_render.py:95-126: smooth / rough white metal reflects the ambient radiance."""
  c, _, rc = _centre_hdr(_metal_sphere(roughness), tone_mapper=mj.ToneMapper.CLAMP, ev100=14.96578428466209)
  exposure = float(rc.exposure[0])
  np.testing.assert_allclose(c * exposure, 0.5, rtol=rtol)


def test_shadow_blocks_direct_light():
  """tests/test_render.py:129-146: an occluder only darkens."""
  b = _Builder()
  mat = b.material(rgba=(1, 1, 1, 1), roughness=1.0)
  b.lists["mat_specular"][-1] = np.asarray(0.0)
  b.geom(mdl.GEOM_PLANE, (5, 5, 0.1), (0, 0, 0), matid=mat)
  b.geom(mdl.GEOM_BOX, (0.3, 0.3, 0.01), (0, 0, 1), matid=mat)
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (0, 0, -1), castshadow=1)
  b.camera((0, 0, 0.5), (0, 0.0001, 0), 30.0, (16, 16))
  mjm = b.build()
  m, d, rc = make(mjm, cam_res=(16, 16))
  lit = oracle.rt_render(m, d, rc)["hdr"][0]
  rc.use_shadows = False
  unlit = oracle.rt_render(m, d, rc)["hdr"][0]
  assert np.all(lit <= unlit + 1e-6)
  assert lit.mean() < 0.5 * unlit.mean()


def test_pinhole_rays_give_planar_depth():
  """tests/test_camera_rays.py:47-54 closed form: a camera at height h over a plane sees planar
  depth h in every pixel and distance h / -dir.z along pinhole_rays."""
  h = 1.7
  mjm = plane_and_light(mdl.LIGHT_DIRECTIONAL, cam_height=h, fovy=50.0, res=12)
  m, d, rc = make(mjm, cam_res=(12, 9))
  out = oracle.rt_render(m, d, rc)
  np.testing.assert_allclose(out["depth"][0], h, rtol=1e-5)
  assert np.all(out["seg"][0, :, 0] == 0) and np.all(out["seg"][0, :, 1] == 5)


def test_background_pixels_are_unsegmented():
  b = _Builder()
  b.geom(mdl.GEOM_SPHERE, (0.2, 0, 0), (0, 0, 0))
  b.camera((0, -3, 0), (0, 0, 0), 30.0, (16, 16))
  m, d, rc = make(b.build(), cam_res=(16, 16))
  out = oracle.rt_render(m, d, rc)
  seg = out["seg"][0].reshape(16, 16, 2)
  assert tuple(seg[0, 0]) == (-1, -1) and tuple(seg[8, 8]) == (0, 5)
  assert out["depth"][0].reshape(16, 16)[0, 0] == 0.0
  # sphere silhouette: radius 0.2 at distance 3 -> half-angle 3.82 deg of a 30 deg view
  frac = np.mean(seg[..., 0] == 0)
  expected = math.pi * math.tan(math.asin(0.2 / 3)) ** 2 / (2 * math.tan(math.radians(15))) ** 2
  assert abs(frac - expected) < 0.02


def test_tie_goes_to_higher_geom_id():
  """traverse.py:311-313: coincident surfaces resolve to the higher geom index."""
  b = _Builder()
  b.geom(mdl.GEOM_BOX, (0.5, 0.5, 0.1), (0, 0, 0))
  b.geom(mdl.GEOM_BOX, (0.5, 0.5, 0.1), (0, 0, 0))
  b.camera((0, 0, 2), (0, 0.0001, 0), 10.0, (4, 4))
  m, d, rc = make(b.build(), cam_res=(4, 4))
  out = oracle.rt_render(m, d, rc)
  assert np.all(out["seg"][0, :, 0] == 1)


def test_alpha_zero_geoms_are_invisible():
  b = _Builder()
  b.geom(mdl.GEOM_SPHERE, (0.5, 0, 0), (0, 0, 0), rgba=(1, 0, 0, 0))
  b.geom(mdl.GEOM_PLANE, (5, 5, 0.1), (0, 0, -1))
  b.camera((0, 0, 2), (0, 0.0001, 0), 10.0, (4, 4))
  m, d, rc = make(b.build(), cam_res=(4, 4))
  out = oracle.rt_render(m, d, rc)
  assert np.all(out["seg"][0, :, 0] == 1)


def test_dfg_lut_invariants():
  """ibl.py:62-93: 0 <= int V Fc <= int V <= 1; smooth grazing-free texels near 1."""
  lut = oracle.dfg_lut(32, 256)
  x, y = lut[..., 0], lut[..., 1]
  assert np.all(x >= 0) and np.all(x <= y + 1e-6) and np.all(y <= 1.0 + 1e-3)
  # smooth surface seen head-on: the GGX lobe is a delta, int V -> 1, int V Fc -> 0
  assert y[0, -1] == pytest.approx(1.0, abs=0.02) and x[0, -1] < 0.01


def test_prefiltered_constant_environment_is_constant():
  mjm = scenes.primitives_scene(0, 8)
  _, _, rc = make(mjm, cam_res=(8, 8), envmap=scenes.constant_sky(0.3, 8, 16))
  stack = rc.env.prefiltered.numpy().view(np.float16).astype(np.float32)
  np.testing.assert_allclose(stack[..., :3], 0.3, rtol=2e-3)


def test_ibl_white_furnace():
  """tests/test_ibl.py:39-45: a white dielectric / metal sphere in a constant environment
  reflects it (rtol 0.03)."""
  for metallic in (0.0, 1.0):
    for rough in (0.0, 0.5, 1.0):
      b = _Builder()
      mat = b.material(rgba=(1, 1, 1, 1), metallic=metallic, roughness=rough)
      b.geom(mdl.GEOM_SPHERE, (0.5, 0, 0), (0, 0, 0), matid=mat)
      b.camera((0, -3, 0), (0, 0, 0), 5.0, (8, 8))
      c, _, _ = _centre_hdr(b.build(), envmap=scenes.constant_sky(1.0, 16, 32), tone_mapper=mj.ToneMapper.CLAMP)
      np.testing.assert_allclose(c, 1.0, rtol=0.03, err_msg=f"metallic {metallic} roughness {rough}")


def test_sincos_turns_accuracy():
  """sin/cos of 2 pi t used for every sampled direction: within 1e-6 of libm over many turns."""
  import ctypes

  lib = oracle.library()
  lib.oracle_sincos_turns.argtypes = [ctypes.c_float, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
  s, c = ctypes.c_float(), ctypes.c_float()
  for t in np.linspace(-1.0, 2.0, 3001):
    lib.oracle_sincos_turns(float(t), ctypes.byref(s), ctypes.byref(c))
    t32 = float(np.float32(t))
    assert abs(s.value - math.sin(2 * math.pi * t32)) < 1e-6
    assert abs(c.value - math.cos(2 * math.pi * t32)) < 1e-6


def test_rng_is_uniform_and_decorrelated():
  """The counter-based PCG stream: uniform on [0, 1), adjacent counters and seeds uncorrelated."""
  import ctypes

  lib = oracle.library()
  lib.oracle_rand.argtypes = [ctypes.c_int32, ctypes.c_int32]
  lib.oracle_rand.restype = ctypes.c_float
  vals = np.array([lib.oracle_rand(7, k) for k in range(20000)])
  assert vals.min() >= 0.0 and vals.max() < 1.0
  hist, _ = np.histogram(vals, bins=20, range=(0, 1))
  chi2 = np.sum((hist - 1000.0) ** 2 / 1000.0)
  assert chi2 < 45.0  # 19 degrees of freedom, p ~ 1e-3
  assert abs(np.corrcoef(vals[:-1], vals[1:])[0, 1]) < 0.03
  seeds = np.array([lib.oracle_rand(s, 5) for s in range(20000)])
  assert abs(np.corrcoef(seeds[:-1], seeds[1:])[0, 1]) < 0.03
  # 24-bit resolution: values are multiples of 2^-24
  assert np.all(np.abs(vals * 16777216.0 - np.round(vals * 16777216.0)) == 0)


def test_supersampling_and_soft_shadows_are_deterministic():
  mjm = scenes.primitives_scene(1, 16)
  m, d, rc = make(mjm, nworld=2, cam_res=(16, 16), samples_per_pixel=2, shadow_samples=4, envmap=scenes.constant_sky())
  rc.light_bulbradius = torch.full_like(rc.light_bulbradius, 0.1)
  a = oracle.rt_render(m, d, rc, seed=3)
  b = oracle.rt_render(m, d, rc, seed=3)
  c = oracle.rt_render(m, d, rc, seed=4)
  np.testing.assert_array_equal(a["hdr"], b["hdr"])
  assert np.any(a["hdr"] != c["hdr"])


def per_world_metallic_scene():
  """tests/test_render.py:188-199: a red sphere under the headlight."""
  from tests.helpers import xyaxes

  b = _Builder()
  mat = b.material(rgba=(0.8, 0.2, 0.2, 1), roughness=0.4)
  b.geom(mdl.GEOM_SPHERE, (0.5, 0, 0), (0, 0, 0), matid=mat)
  b.camera((0, -3, 0), (0, 0, 0), 25.0, (32, 32))
  mjm = b.build(headlight_active=True, headlight_ambient=np.full(3, 0.3, np.float32), headlight_diffuse=np.full(3, 0.5, np.float32))
  mjm.cam_xmat[0] = xyaxes((1, 0, 0), (0, 0, 1))
  return mjm


@pytest.mark.parametrize("nworld", [1, 2])
def test_per_world_metallic(nworld):
  """tests/test_render.py:186-208: the last world's sphere is metal, the first renders as alone."""
  from tests.helpers import Shot

  s = Shot(per_world_metallic_scene(), nworld=nworld)
  per_world = np.zeros((nworld, s.mjm.ngeom), dtype=np.float32)
  per_world[nworld - 1] = 1.0
  s.rc.geom_metallic = torch.tensor(per_world)
  rgb = s.render().rgb()
  single = Shot(per_world_metallic_scene()).render().rgb()[0]
  np.testing.assert_allclose(rgb[0], single if nworld == 2 else rgb[0])
  if nworld == 2:
    assert not np.allclose(rgb[0], rgb[1])


def display_scene():
  """tests/test_render.py:211-215: a rough coloured plane under a slanted sun."""
  b = _Builder()
  mat = b.material(rgba=(0.5, 0.6, 0.7, 1), roughness=0.5)
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (0.3, 0, -1), diffuse=(0.7, 0.7, 0.7), castshadow=0)
  b.geom(mdl.GEOM_PLANE, (5, 5, 0.1), (0, 0, 0), matid=mat)
  b.camera((0, 0, 2), (0, 0.0001, 0), 10.0, (32, 32))
  mjm = b.build()
  mjm.cam_xmat[0] = np.eye(3, dtype=np.float32).reshape(9)
  return mjm


def test_hdr_is_consistent_with_display_output():
  """tests/test_render.py:211-219: rgb = sRGB(clip(hdr * exposure)) under the clamp mapper."""
  from mujoco_warp_pbr_amd import io
  from tests.helpers import Shot

  s = Shot(display_scene(), render_hdr=True, tone_mapper=mj.ToneMapper.CLAMP, ev100=io.DEFAULT_EV100).render()
  expected = _srgb(np.clip(s.hdr()[0] * s.exposure, 0.0, 1.0))
  np.testing.assert_allclose(s.rgb()[0], expected, atol=1.5 / 255)


def test_ground_planes_are_tested_outside_the_scene_bvh():
  """Planes go to MjpbrGeometry.direct_geom (tested by every ray before the scene BVH) when the
  model has at most io.MAX_DIRECT_PLANES of them, otherwise only the unbounded ones; other geoms
  stay in the BVH. The oracle's brute force walks the two lists merged in index order, so the
  image equals the one over the single enabled list."""
  b = _Builder()
  b.geom(mdl.GEOM_SPHERE, (0.2, 0, 0), (0, 0, 0.2))
  b.geom(mdl.GEOM_PLANE, (0, 0, 0.1), (0, 0, 0))
  b.geom(mdl.GEOM_PLANE, (0.3, 0.3, 0.1), (0.4, 0.4, 0.3))
  b.geom(mdl.GEOM_BOX, (0.1, 0.1, 0.1), (-0.4, 0.1, 0.1))
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (0.3, 0.2, -1), intensity=2.0, castshadow=1)
  b.camera((0, -2.0, 1.2), (0, 0, 0.1), 45, (24, 24))
  mjm = b.build()
  m, d, rc = make(mjm, cam_res=(24, 24))
  assert rc.base.direct_geom_ids.tolist() == [1, 2]
  assert rc.base.bvh_geom_ids.tolist() == [0, 3]
  assert rc.base.tlas_depth == mj.bvh.scene_depth(2)
  split = oracle.rt_render(m, d, rc)
  rc.base.bvh_geom_ids, rc.base.direct_geom_ids = None, None
  single = oracle.rt_render(m, d, rc)
  for k in ("seg", "depth", "hdr", "rgb"):
    np.testing.assert_array_equal(split[k], single[k])
  seg = split["seg"][..., 0]
  assert {0, 1, 2, 3} <= set(np.unique(seg).tolist())
  # with more planes than the limit, only the unbounded one is tested directly
  b = _Builder()
  for k in range(mj.io.MAX_DIRECT_PLANES + 1):
    b.geom(mdl.GEOM_PLANE, (0.1, 0.1, 0.1) if k else (0, 0, 0.1), (0.3 * k, 0, 0.01 * k))
  b.geom(mdl.GEOM_SPHERE, (0.2, 0, 0), (0, 0, 0.2))
  b.camera((0, -2.0, 1.2), (0, 0, 0.1), 45, (8, 8))
  _, _, rc2 = make(b.build(), cam_res=(8, 8))
  assert rc2.base.direct_geom_ids.tolist() == [0]
