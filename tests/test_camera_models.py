"""Orthographic and intrinsic-matrix cameras (reference camera.py:159-191, MuJoCo Warp's
compute_ray branches for cam_projection = orthographic and cam_sensorsize != 0).

The oracle is pinned by what the camera models fix geometrically: an orthographic camera's rays
are parallel and its fovy is the view's full vertical extent in length units, so a sphere under
it shows its exact silhouette at its exact depth; an intrinsic camera with the principal point
at the centre and focal length f = sensor_height / (2 tan(fovy / 2)) is the pinhole camera of
that fovy. The sign convention of a principal-point offset is not fixed by anything here (parity
unpinned); it is compared between the GPU and the oracle only.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import mujoco_warp_pbr_amd as mj
import oracle
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd.scenes import _Builder
from tests.helpers import compare_rt
from tests.helpers import gpu_outputs
from tests.helpers import make

RES = 32
HEIGHT = 2.0
SPHERE_Z, SPHERE_R = 0.3, 0.4


def _scene(projection=0, fovy=1.5, sensor=(0.0, 0.0), intrinsic=(0.0, 0.0, 0.0, 0.0), res=(RES, RES)):
  """Plane at z = 0, sphere at (0, 0, 0.3) radius 0.4, camera at height 2 looking straight down
  (camera x = world x, camera y = world y)."""
  b = _Builder()
  b.geom(mdl.GEOM_PLANE, (0.0, 0.0, 0.1), (0, 0, 0), matid=b.material(rgba=(0.7, 0.7, 0.7, 1), roughness=0.8))
  b.geom(mdl.GEOM_SPHERE, (SPHERE_R, 0, 0), (0, 0, SPHERE_Z), matid=b.material(rgba=(0.8, 0.3, 0.2, 1), roughness=0.4))
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (0.3, 0.2, -1.0))
  b.camera((0, 0, HEIGHT), (0, 0.0001, 0), fovy, res)
  mjm = b.build()
  mjm.cam_xmat[0] = np.eye(3, dtype=np.float32).reshape(9)
  mjm.cam_projection[0] = projection
  mjm.cam_sensorsize[0] = sensor
  mjm.cam_intrinsic[0] = intrinsic
  return mjm


def _depth_seg(mjm, res=(RES, RES), device="cpu", **kw):
  m, d, rc = make(mjm, nworld=1, device=device, cam_res=res, **kw)
  if device == "cpu":
    out = oracle.rt_render(m, d, rc)
  else:
    mj.refit_bvh(m, d, rc)
    mj.render(m, d, rc)
    out = gpu_outputs(rc)
  w, h = res
  return out["depth"][0].reshape(h, w), out["seg"][0].reshape(h, w, 2)[..., 0], out, (m, d, rc)


def test_orthographic_camera_shows_exact_silhouette_and_depth():
  extent = 1.5  # full vertical extent in length units
  depth, seg, _, _ = _depth_seg(_scene(projection=1, fovy=extent))
  px = (np.arange(RES) + 0.5) / RES - 0.5
  x = px[None, :] * extent
  y = -px[:, None] * extent
  r2 = x * x + y * y
  inside = r2 < SPHERE_R**2 - 1e-3
  outside = r2 > SPHERE_R**2 + 1e-3
  assert inside.sum() > 50 and outside.sum() > 50
  assert (seg[inside] == 1).all() and (seg[outside] == 0).all()
  want = HEIGHT - SPHERE_Z - np.sqrt(np.maximum(SPHERE_R**2 - r2, 0.0))
  np.testing.assert_allclose(depth[inside], want[inside], rtol=1e-5)
  np.testing.assert_allclose(depth[outside], HEIGHT, rtol=1e-6)


def test_orthographic_aspect_keeps_vertical_extent():
  """A 2:1 image widens the horizontal extent, not the vertical one."""
  extent = 1.2
  depth, seg, _, _ = _depth_seg(_scene(projection=1, fovy=extent, res=(2 * RES, RES)), res=(2 * RES, RES))
  px = (np.arange(2 * RES) + 0.5) / (2 * RES) - 0.5
  py = (np.arange(RES) + 0.5) / RES - 0.5
  x = px[None, :] * extent * 2.0
  y = -py[:, None] * extent
  r2 = x * x + y * y
  clear = np.abs(r2 - SPHERE_R**2) > 2e-3
  np.testing.assert_array_equal(seg[clear], np.where(r2 < SPHERE_R**2, 1, 0)[clear])


def test_centred_intrinsic_camera_equals_pinhole():
  fovy = 40.0
  sensor = (0.024, 0.024)
  f = sensor[1] / (2.0 * math.tan(math.radians(fovy) / 2.0))
  d_pin, s_pin, _, _ = _depth_seg(_scene(fovy=fovy))
  d_int, s_int, _, _ = _depth_seg(_scene(fovy=fovy, sensor=sensor, intrinsic=(f, f, 0.0, 0.0)))
  np.testing.assert_array_equal(s_int, s_pin)
  np.testing.assert_allclose(d_int, d_pin, rtol=2e-6)
  # a longer focal length along x narrows the horizontal field: the sphere spans more columns
  _, s_x2, _, _ = _depth_seg(_scene(fovy=fovy, sensor=sensor, intrinsic=(2 * f, f, 0.0, 0.0)))
  assert (s_x2 == 1).any(axis=0).sum() > (s_pin == 1).any(axis=0).sum()
  assert (s_x2 == 1).any(axis=1).sum() == (s_pin == 1).any(axis=1).sum()


CASES = [
  ("orthographic", dict(projection=1, fovy=1.5)),
  ("intrinsic_centred", dict(fovy=40.0, sensor=(0.024, 0.024), intrinsic=(0.033, 0.033, 0.0, 0.0))),
  ("intrinsic_offset", dict(fovy=40.0, sensor=(0.024, 0.016), intrinsic=(0.03, 0.028, 0.002, -0.0015))),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,cam", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("spp", [1, 2])
def test_gpu_camera_models_match_oracle(cuda, name, cam, spp):
  mjm = _scene(**cam)
  m, d, rc = make(mjm, nworld=2, device=cuda, cam_res=(RES, RES), samples_per_pixel=spp, render_guides=True)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc)
  gpu = gpu_outputs(rc)
  compare_rt(gpu, oracle.rt_render(m, d, rc))
  assert (gpu["seg"][..., 0] == 1).any()
