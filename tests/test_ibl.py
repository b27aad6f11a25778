"""Image-based lighting: the reference's tests/test_ibl.py (orientation, background and yaw,
per-world maps, the MJCF skybox fallback) restated on the CPU oracle, and the HIP renderer
compared with the oracle on the same scenes. The white-furnace test is in test_oracle_rt.py.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
import oracle
from mujoco_warp_pbr_amd import envmap
from mujoco_warp_pbr_amd import io
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd.scenes import _Builder
from tests.helpers import Shot
from tests.helpers import compare_rt
from tests.helpers import xyaxes

_FRONT = xyaxes((1, 0, 0), (0, 0, 1))  # MJCF xyaxes="1 0 0 0 0 1": looks along +y


def _build(b, frames, **extra):
  mjm = b.build(**extra)
  mjm.cam_xmat[:] = np.stack(frames)
  return mjm


def _sphere(metallic=0.0, roughness=1.0, material=True):
  """tests/test_ibl.py:26-36: a white sphere seen head-on through a 5 degree camera at -3 y."""
  b = _Builder()
  if material:
    mat = b.material(rgba=(1, 1, 1, 1), metallic=metallic, roughness=roughness)
    b.geom(mdl.GEOM_SPHERE, (0.5, 0, 0), (0, 0, 0), matid=mat)
  else:
    b.geom(mdl.GEOM_SPHERE, (0.5, 0, 0), (0, 0, 0), rgba=(1, 1, 1, 1))
  b.camera((0, -3, 0), (0, 0, 0), 5.0, (32, 32))
  return b, [_FRONT]


def sky_red_ground_blue():
  image = np.zeros((64, 128, 3), dtype=np.float32)
  image[:32, :, 0] = 1.0
  image[32:, :, 2] = 1.0
  return image


def orientation_scene():
  """tests/test_ibl.py:48-63: a matte upward plane seen from above."""
  b = _Builder()
  mat = b.material(rgba=(1, 1, 1, 1), roughness=1.0)
  b.lists["mat_specular"][-1] = np.asarray(0.0)
  b.geom(mdl.GEOM_PLANE, (5, 5, 0.1), (0, 0, 0), matid=mat)
  b.camera((0, 0, 2), (0, 0.0001, 0), 10.0, (32, 32))
  return _build(b, [xyaxes((1, 0, 0), (0, 1, 0))])


def green_half():
  image = np.zeros((64, 128, 3), dtype=np.float32)
  image[:, 64:, 1] = 1.0  # +y half of the horizon (phi > 0)
  return image


def background_scene():
  """tests/test_ibl.py:66-80: a camera at the origin looking along +y into the background."""
  b = _Builder()
  b.geom(mdl.GEOM_SPHERE, (0.1, 0, 0), (0, -5, 0))
  b.camera((0, 0, 0), (0, 1, 0), 10.0, (32, 32))
  return _build(b, [_FRONT])


def skybox_scene(size=16, value=255, fovy=5.0):
  """tests/test_ibl.py:124-135: a flat white skybox (six stacked faces) around a white sphere."""
  b, frames = _sphere(material=False)
  b.lists["cam_fovy"][-1] = np.asarray(fovy)
  b.texture(np.full((6 * size, size, 3), value, np.uint8), mdl.COLORSPACE_AUTO, tex_type=mdl.TEXTURE_SKYBOX)
  return _build(b, frames)


def test_environment_orientation():
  """tests/test_ibl.py:48-63: row 0 of the map is the zenith."""
  c = Shot(orientation_scene(), envmap=sky_red_ground_blue(), envmap_intensity=1.0, render_hdr=True).render().hdr()[0, 16, 16]
  assert c[0] > 0.9 and c[2] < 0.05


def test_background_shows_environment_and_yaw_rotates_it():
  """tests/test_ibl.py:66-85."""
  kw = {"envmap": green_half(), "envmap_intensity": 1.0, "render_hdr": True}
  green = Shot(background_scene(), **kw).render().hdr()[0, 16, 16]
  rotated = Shot(background_scene(), envmap_rotation=180.0, **kw).render().hdr()[0, 16, 16]
  assert green[1] > 0.9
  assert rotated[1] < 0.05


@pytest.mark.parametrize("nworld", [1, 2])
def test_per_world_environment(nworld):
  """tests/test_ibl.py:88-97: env_id picks each world's map."""
  maps = [np.full((32, 64, 3), 0.25, np.float32), np.full((32, 64, 3), 1.0, np.float32)]
  b, frames = _sphere(0.0, 1.0)
  s = Shot(_build(b, frames), nworld=nworld, envmap=maps, envmap_intensity=1.0, render_hdr=True)
  s.rc.env_id = torch.tensor(np.arange(nworld) % 2, dtype=torch.int32)
  hdr = s.render().hdr()
  np.testing.assert_allclose(hdr[0, 16, 16], 0.25, rtol=0.03)
  if nworld == 2:
    np.testing.assert_allclose(hdr[1, 16, 16], 1.0, rtol=0.03)


def test_skybox_fallback_lights_physical_scene():
  """tests/test_ibl.py:124-140: the MJCF skybox lights a physical scene when no map is given."""
  kw = {"tone_mapper": mj.ToneMapper.CLAMP, "ev100": io.DEFAULT_EV100}
  lit = Shot(skybox_scene(), **kw).render().rgb()[0, 16, 16]
  dark = Shot(skybox_scene(), use_skybox=False, **kw).render().rgb()[0, 16, 16]
  assert lit.min() > 0.8
  assert dark.max() < 0.05


def test_skybox_is_background_only_for_mujoco_lights():
  """io.py:267-285: outside the physical light model the skybox shows but lights nothing."""
  s = Shot(skybox_scene(fovy=40.0), light_model=mj.LightModel.MUJOCO, tone_mapper=mj.ToneMapper.CLAMP, ev100=io.DEFAULT_EV100)
  assert not s.rc.has_env and s.rc.env_background
  rgb = s.render().rgb()[0]
  assert rgb[16, 16].max() < 0.05  # the unlit sphere
  assert rgb[0, 0].min() > 0.95  # the white sky behind it


def test_skybox_image_conversion():
  """envmap.py:307-332: cube faces -> linear equirect; each face lands on its own direction."""
  size = 8
  faces = np.zeros((6, size, size, 3), np.uint8)
  for f in range(6):
    faces[f] = 40 * (f + 1)
  b = _Builder()
  b.geom(mdl.GEOM_SPHERE, (0.5, 0, 0), (0, 0, 0))
  b.texture(faces.reshape(6 * size, size, 3), mdl.COLORSPACE_AUTO, tex_type=mdl.TEXTURE_SKYBOX)
  mjm = b.build()
  img = envmap.skybox_image(mjm, 64)
  assert img.shape == (32, 64, 3) and img.dtype == np.float32
  level = lambda f: float(mj.material.srgb_to_linear(40 * (f + 1) / 255.0))  # noqa: E731
  # World +x, -x, +y, -y, +z, -z look up cube +x, -x, -z, +z, +y, -y.
  for world_dir, face in (((1, 0, 0), 0), ((-1, 0, 0), 1), ((0, 1, 0), 5), ((0, -1, 0), 4), ((0, 0, 1), 2), ((0, 0, -1), 3)):
    d = np.asarray(world_dir, np.float64)
    u = (0.5 + np.arctan2(d[1], d[0]) / (2 * np.pi)) % 1.0
    v = np.arccos(np.clip(d[2], -1, 1)) / np.pi
    px = img[min(int(v * 32), 31), min(int(u * 64), 63)]
    np.testing.assert_allclose(px, level(face), rtol=1e-5, err_msg=f"direction {world_dir}")
  assert envmap.skybox_image(_Builder().build()) is None


# The same scenes on the GPU: every output against the oracle.
GPU_CASES = [
  ("orientation", orientation_scene, {"envmap": sky_red_ground_blue(), "envmap_intensity": 1.0}),
  ("background_yaw", background_scene, {"envmap": green_half(), "envmap_intensity": 1.0, "envmap_rotation": 180.0}),
  ("skybox", skybox_scene, {"tone_mapper": mj.ToneMapper.CLAMP}),
  ("skybox_mujoco", lambda: skybox_scene(fovy=40.0), {"light_model": mj.LightModel.MUJOCO}),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,scene,kw", GPU_CASES, ids=[c[0] for c in GPU_CASES])
def test_gpu_ibl_matches_oracle(cuda, name, scene, kw):
  s = Shot(scene(), cuda, nworld=2, **kw).render(seed=2)
  compare_rt(s.out, oracle.rt_render(s.m, s.d, s.rc, seed=2))


@pytest.mark.gpu
def test_gpu_per_world_environment(cuda):
  maps = [np.full((32, 64, 3), 0.25, np.float32), np.full((32, 64, 3), 1.0, np.float32)]
  b, frames = _sphere(0.0, 1.0)
  s = Shot(_build(b, frames), cuda, nworld=3, envmap=maps, envmap_intensity=1.0)
  s.rc.env_id = torch.tensor([1, 0, 1], dtype=torch.int32, device=cuda)
  s.render()
  np.testing.assert_allclose(s.hdr()[:, 16, 16], [[1.0] * 3, [0.25] * 3, [1.0] * 3], rtol=0.03)
  compare_rt(s.out, oracle.rt_render(s.m, s.d, s.rc))
