"""Settings changed between frames: the reference's tests/test_kernel_cache.py restated.

The reference specialises its Warp megakernel on "baked" settings and rebuilds it when one
changes; here one precompiled kernel reads every setting from the MjpbrSettings struct that
render() builds from the context on each call, so there is no kernel to rebuild and the
reference's rebuild assertions have no counterpart. What carries over is the behaviour: a
setting changed after the first render renders exactly like a context created with it, replaced
arrays take effect, and a replaced translucent m.geom_rgba warns until rc.has_transparency is
set. On the CPU the oracle renders through the same host marshalling as the GPU path.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd.scenes import _Builder
from tests.helpers import Shot
from tests.helpers import xyaxes


def kernel_cache_scene():
  """tests/test_kernel_cache.py:25-35: plane, sphere and box under a shadowing spot light."""
  b = _Builder()
  b.light(mdl.LIGHT_SPOT, (0, 0, 3), (0, 0, -1), diffuse=(0.7, 0.7, 0.7), castshadow=1)
  b.camera((0, -3, 1.2), (0, 0, 0), 45.0, (24, 16))
  b.geom(mdl.GEOM_PLANE, (3, 3, 0.1), (0, 0, 0), rgba=(0.7, 0.7, 0.7, 1))
  b.geom(mdl.GEOM_SPHERE, (0.5, 0, 0), (0, 0, 0.5), rgba=(0.8, 0.3, 0.2, 1))
  b.geom(mdl.GEOM_BOX, (0.2, 0.2, 0.3), (0.9, 0.3, 0.3), rgba=(0.2, 0.4, 0.9, 1))
  mjm = b.build(headlight_active=True)
  mjm.cam_xmat[0] = xyaxes((1, 0, 0), (0, 0.35, 1))
  return mjm


CHANGES = [
  ("render_order", mj.RenderOrder.WORLD),
  ("render_order", mj.RenderOrder.ROW),
  ("tone_mapper", mj.ToneMapper.ACES),
  ("use_shadows", False),
  ("samples_per_pixel", 2),
  ("shadow_light_fraction", 0.5),
]


def _shot(device="cpu", **kw):
  kw.setdefault("nworld", 2)
  return Shot(kernel_cache_scene(), device, cam_res=(24, 16), render_seg=True, **kw)


def _seg(s):
  return s._image("seg", 0)


def _check_change(device, field, value):
  s = _shot(device).render()
  setattr(s.rc, field, value)
  s.render()
  expected = _shot(device, **{field: value}).render()
  np.testing.assert_allclose(s.rgb(), expected.rgb(), atol=1.5 / 255)
  np.testing.assert_array_equal(_seg(s), _seg(expected))


@pytest.mark.parametrize("field,value", CHANGES, ids=[f"{f}={v}" for f, v in CHANGES])
def test_change_after_first_render(field, value):
  """tests/test_kernel_cache.py:55-70."""
  _check_change("cpu", field, value)


def test_tier_switch_renders_the_other_tier():
  """tests/test_kernel_cache.py:73-81."""
  s = _shot(nworld=1, path_samples=2).render()
  s.rc.tier = int(mj.Tier.PATH)
  s.render()
  expected = _shot(nworld=1, tier=mj.Tier.PATH, path_samples=2).render()
  np.testing.assert_allclose(s.rgb(), expected.rgb(), atol=1.5 / 255)


def test_segmentation_background_without_fill():
  """tests/test_kernel_cache.py:84-91: background pixels are written (-1, -1) every frame."""
  s = _shot(nworld=1).render()
  seg = _seg(s).copy()
  assert np.all(seg[0, 0] == -1)
  s.rc.base.seg_data.fill_(7)
  s.render()
  np.testing.assert_array_equal(_seg(s), seg)


def test_replaced_array_is_used():
  """tests/test_kernel_cache.py:94-105."""
  s = _shot(nworld=1).render()
  before = s.rgb()
  s.render()
  np.testing.assert_array_equal(s.rgb(), before)
  s.rc.geom_roughness = torch.full((1, s.mjm.ngeom), 0.05)
  s.rc.geom_metallic = torch.ones((1, s.mjm.ngeom))
  s.render()
  assert np.abs(s.rgb() - before).max() > 0.05


def test_translucent_override_warns_until_enabled():
  """tests/test_kernel_cache.py:108-121."""
  s = _shot(nworld=1).render()
  assert not s.rc.has_transparency
  rgba = s.m.geom_rgba.clone()
  rgba[..., 3] = 0.5
  s.m.geom_rgba = rgba
  with pytest.warns(UserWarning, match="has_transparency"):
    s.render()
  opaque = s.rgb()
  s.rc.has_transparency = True
  s.render()
  assert np.abs(s.rgb() - opaque).max() > 0.05


def test_reflectance_override_warns():
  """launch.py:59-87: a replaced reflective rc.mat_reflectance needs rc.has_reflection."""
  b = _Builder()
  mat = b.material(rgba=(0.5, 0.5, 0.5, 1), roughness=0.5)
  b.geom(mdl.GEOM_PLANE, (3, 3, 0.1), (0, 0, 0), matid=mat)
  b.geom(mdl.GEOM_SPHERE, (0.3, 0, 0), (0, 0, 0.3))
  b.camera((0, -2, 1), (0, 0, 0.2), 45.0, (16, 16))
  s = Shot(b.build(), cam_res=(16, 16)).render()
  assert not s.rc.has_reflection
  s.rc.mat_reflectance = torch.ones_like(s.rc.mat_reflectance)
  with pytest.warns(UserWarning, match="has_reflection"):
    s.render()


@pytest.mark.gpu
@pytest.mark.parametrize("field,value", CHANGES, ids=[f"{f}={v}" for f, v in CHANGES])
def test_gpu_change_after_first_render(cuda, field, value):
  _check_change(cuda, field, value)


@pytest.mark.gpu
def test_gpu_translucent_override_warns_until_enabled(cuda):
  s = _shot(cuda, nworld=1).render()
  rgba = s.m.geom_rgba.clone()
  rgba[..., 3] = 0.5
  s.m.geom_rgba = rgba
  with pytest.warns(UserWarning, match="has_transparency"):
    s.render()
  opaque = s.rgb()
  s.rc.has_transparency = True
  s.render()
  assert np.abs(s.rgb() - opaque).max() > 0.05
