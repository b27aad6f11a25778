"""Textures, normal maps, anti-aliasing, translucency, reflections and soft shadows: the
reference's tests/test_effects.py restated on the CPU oracle, and the HIP renderer compared with
the oracle on the same scenes.

Each scene is the reference test's MJCF rebuilt with the scene builder (MuJoCo defaults where
the MJCF leaves a field unset: headlight ambient 0.1 / diffuse 0.4 / specular 0.5, material
specular 0.5).
"""

from __future__ import annotations

import numpy as np
import pytest

import mujoco_warp_pbr_amd as mj
import oracle
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd.scenes import _Builder
from tests.helpers import Shot
from tests.helpers import compare_rt
from tests.helpers import xyaxes

_AMBIENT_ONLY = {
  "headlight_active": True,
  "headlight_ambient": np.ones(3, np.float32),
  "headlight_diffuse": np.zeros(3, np.float32),
  "headlight_specular": np.zeros(3, np.float32),
}


def _srgb(x):
  return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1 / 2.4) - 0.055)


def _top_camera(b, fovy=10.0, pos=(0, 0, 2)):
  """MJCF `<camera pos="0 0 2" xyaxes="1 0 0 0 1 0"/>`: straight down, image +x = world +x."""
  b.camera(pos, (0, 0.0001, 0), fovy, (32, 32))


def _build(b, frames, **extra):
  mjm = b.build(**extra)
  mjm.cam_xmat[:] = np.stack(frames)
  return mjm


def _no_specular(b):
  b.lists["mat_specular"][-1] = np.asarray(0.0)


def normal_map_scene(tilt_deg: float, light_deg: float):
  """tests/test_effects.py:28-50: flat normal-map texel tilted towards +u, tilted sun."""
  t = np.radians(tilt_deg)
  rgb = 0.5 * (np.array([np.sin(t), 0.0, np.cos(t)]) + 1.0)
  b = _Builder()
  texel = np.round(np.broadcast_to(rgb, (8, 8, 3)) * 255.0).astype(np.uint8)
  tex = b.texture(texel, mdl.COLORSPACE_LINEAR)
  mat = b.material(rgba=(1, 1, 1, 1), metallic=0.0, roughness=1.0, texid={mdl.ROLE_NORMAL: tex})
  _no_specular(b)
  ld = np.radians(light_deg)
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (-np.sin(ld), 0.0, -np.cos(ld)), diffuse=(0.7, 0.7, 0.7), castshadow=0)
  b.geom(mdl.GEOM_PLANE, (5, 5, 0.1), (0, 0, 0), matid=mat)
  _top_camera(b)
  return _build(b, [xyaxes((1, 0, 0), (0, 1, 0))])


def checker_scene():
  """tests/test_effects.py:63-76: a 256^2 two-colour checker repeated 4000 times."""
  b = _Builder()
  y, x = np.mgrid[0:256, 0:256]
  white = ((x // 128 + y // 128) % 2 == 0)[..., None]
  tex = b.texture(np.where(white, 255, 0).repeat(3, axis=2).astype(np.uint8), mdl.COLORSPACE_AUTO)
  mat = b.material(rgba=(1, 1, 1, 1), roughness=1.0, texid={mdl.ROLE_RGB: tex}, repeat=(4000, 4000))
  _no_specular(b)
  b.geom(mdl.GEOM_PLANE, (5, 5, 0.1), (0, 0, 0), matid=mat)
  _top_camera(b)
  return _build(b, [xyaxes((1, 0, 0), (0, 1, 0))], **_AMBIENT_ONLY)


def edge_scene():
  """tests/test_effects.py:84-93: a white box whose edge cuts one pixel column."""
  b = _Builder()
  b.geom(mdl.GEOM_BOX, (1, 1, 0.1), (-0.987, 0, 0), rgba=(1, 1, 1, 1))
  _top_camera(b, fovy=5.0)
  return _build(b, [xyaxes((1, 0, 0), (0, 1, 0))], **_AMBIENT_ONLY)


def translucent_scene():
  """tests/test_effects.py:102-113: a 25%-opaque blue layer over a red one."""
  b = _Builder()
  b.geom(mdl.GEOM_BOX, (1, 1, 0.01), (0, 0, 0), rgba=(1, 0, 0, 1))
  b.geom(mdl.GEOM_BOX, (1, 1, 0.01), (0, 0, 0.5), rgba=(0, 0, 1, 0.25))
  _top_camera(b)
  return _build(b, [xyaxes((1, 0, 0), (0, 1, 0))], **_AMBIENT_ONLY)


def mirror_scene():
  """tests/test_effects.py:122-133: black mirror plane under a green box above the camera."""
  b = _Builder()
  mat = b.material(rgba=(0, 0, 0, 1), reflectance=1.0, roughness=1.0)
  _no_specular(b)
  b.geom(mdl.GEOM_PLANE, (5, 5, 0.1), (0, 0, 0), matid=mat)
  b.geom(mdl.GEOM_BOX, (2, 2, 0.01), (0, 0, 3), rgba=(0, 1, 0, 1))
  _top_camera(b)
  return _build(b, [xyaxes((1, 0, 0), (0, 1, 0))], **_AMBIENT_ONLY)


def penumbra_scene():
  """tests/test_effects.py:142-153: a 0.4 m bulb over a small occluder, wide camera."""
  b = _Builder()
  b.light(mdl.LIGHT_POINT, (0, 0, 2), (0, 0, -1), diffuse=(0.7, 0.7, 0.7), castshadow=1, intensity=50.0, range_=100.0, bulbradius=0.4)
  b.geom(mdl.GEOM_PLANE, (5, 5, 0.1), (0, 0, 0), rgba=(1, 1, 1, 1))
  b.geom(mdl.GEOM_BOX, (0.25, 0.25, 0.01), (0, 0, 1), rgba=(1, 1, 1, 1))
  b.camera((0, 0.001, 0.6), (0, 0.0011, 0), 120.0, (64, 64))
  return _build(b, [xyaxes((1, 0, 0), (0, 1, 0))])


def _hdr16(mjm, device="cpu", **kw):
  return Shot(mjm, device, render_hdr=True, **kw).render().hdr()[0, 16, 16]


def test_normal_map_tilts_towards_plus_u():
  """tests/test_effects.py:53-60: Lambert radiance follows the tilted normal's cosine."""
  flat = _hdr16(normal_map_scene(0.0, 30.0))
  toward = _hdr16(normal_map_scene(30.0, 30.0))
  away = _hdr16(normal_map_scene(-30.0, 30.0))
  np.testing.assert_allclose(toward / flat, 1.0 / np.cos(np.radians(30.0)), rtol=0.02)
  np.testing.assert_allclose(away / flat, np.cos(np.radians(60.0)) / np.cos(np.radians(30.0)), rtol=0.02)


def test_texture_lod_filters_minified_checker():
  """tests/test_effects.py:63-81: a far-minified checker averages its linear texels."""
  rgb = Shot(checker_scene(), tone_mapper=mj.ToneMapper.CLAMP, specular_aa=False).render().rgb()[0]
  np.testing.assert_allclose(rgb.mean(), _srgb(0.5), atol=0.02)
  assert rgb.std() < 0.02


def test_supersampling_resolves_edge_coverage():
  """tests/test_effects.py:84-99."""
  kwargs = {"tone_mapper": mj.ToneMapper.CLAMP, "cam_res": (8, 8)}
  hard = Shot(edge_scene(), samples_per_pixel=1, **kwargs).render().rgb()[0]
  soft = Shot(edge_scene(), samples_per_pixel=4, **kwargs).render().rgb()[0]
  assert set(np.unique(hard.round(2))) <= {0.0, 1.0}
  assert np.any((soft > 0.05) & (soft < 0.95))


def test_translucent_layer_blends_over_background():
  """tests/test_effects.py:102-119: about 0.25 blue + 0.75 red under constant ambient."""
  s = Shot(translucent_scene(), render_hdr=True, tone_mapper=mj.ToneMapper.CLAMP)
  assert s.rc.has_transparency
  c = s.render().hdr()[0, 16, 16] * s.exposure
  assert c[0] > 0.6 and c[2] > 0.15 and c[1] < 0.05


def test_mirror_reflects_scene():
  """tests/test_effects.py:122-139."""
  s = Shot(mirror_scene(), tone_mapper=mj.ToneMapper.CLAMP)
  assert s.rc.has_reflection
  c = s.render().rgb()[0, 16, 16]
  assert c[1] > 0.9 and c[1] - c[0] > 0.5


def _penumbra_row(device="cpu", **kw):
  return Shot(penumbra_scene(), device, render_hdr=True, cam_res=(64, 64), **kw).render().hdr()[0, 32, :, 0]


def test_soft_shadows_have_penumbra():
  """tests/test_effects.py:142-163: one shadow ray gives a hard edge, 32 a penumbra."""
  unshadowed = _penumbra_row(use_shadows=False)
  hard = _penumbra_row(shadow_samples=1) / unshadowed
  soft = _penumbra_row(shadow_samples=32) / unshadowed
  penumbra = lambda v: int(np.sum((v > 0.05) & (v < 0.95)))  # noqa: E731
  assert penumbra(hard) == 0
  assert penumbra(soft) > 4


# The same scenes on the GPU: every output against the oracle.
GPU_CASES = [
  ("normal_map", lambda: normal_map_scene(30.0, 30.0), {"render_hdr": True}),
  ("checker_lod", checker_scene, {"tone_mapper": mj.ToneMapper.CLAMP, "specular_aa": False}),
  ("edge_ss4", edge_scene, {"tone_mapper": mj.ToneMapper.CLAMP, "cam_res": (8, 8), "samples_per_pixel": 4}),
  ("translucent", translucent_scene, {"tone_mapper": mj.ToneMapper.CLAMP}),
  ("mirror", mirror_scene, {"tone_mapper": mj.ToneMapper.CLAMP}),
  ("soft_shadow", penumbra_scene, {"cam_res": (64, 64), "shadow_samples": 32}),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,scene,kw", GPU_CASES, ids=[c[0] for c in GPU_CASES])
def test_gpu_effects_match_oracle(cuda, name, scene, kw):
  s = Shot(scene(), cuda, nworld=2, **kw).render(seed=5)
  compare_rt(s.out, oracle.rt_render(s.m, s.d, s.rc, seed=5))
