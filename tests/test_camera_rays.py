"""Ray-bundle cameras (set_camera_rays): the reference's tests/test_camera_rays.py restated on the
CPU oracle, and the GPU kernels against the oracle with bundles.

Scene: the reference test's MJCF (tests/test_camera_rays.py:25-41) rebuilt with the scene
builder: checkered floor, a metal-ish sphere, a green box, a directional light, two cameras.
"""

from __future__ import annotations

import numpy as np
import pytest

import mujoco_warp_pbr_amd as mj
import oracle
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd.scenes import _Builder
from tests.helpers import compare_rt
from tests.helpers import gpu_outputs
from tests.helpers import make

RES = (24, 16)


def _xyaxes(x, y) -> np.ndarray:
  """MuJoCo camera frame from xyaxes: columns x, y (orthogonalised), z = x cross y."""
  x = np.asarray(x, np.float64) / np.linalg.norm(x)
  y = np.asarray(y, np.float64)
  y = y - x * np.dot(x, y)
  y /= np.linalg.norm(y)
  return np.stack([x, y, np.cross(x, y)], axis=1).astype(np.float32).reshape(9)


def _scene(side_camera: bool = True, top_camera: bool = False):
  b = _Builder()
  y, x = np.mgrid[0:64, 0:64]
  checker = ((x // 32 + y // 32) % 2)[..., None]
  img = np.where(checker == 0, np.array([0.1, 0.2, 0.3]), np.array([0.8, 0.8, 0.8]))
  tex = b.texture(np.round(np.concatenate([img, np.ones((64, 64, 1))], -1) * 255).astype(np.uint8), mdl.COLORSPACE_AUTO)
  floor = b.material(rgba=(1, 1, 1, 1), texid={mdl.ROLE_RGBA: tex}, repeat=(4, 4))
  metal = b.material(rgba=(0.8, 0.3, 0.2, 1), metallic=0.5, roughness=0.4)
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (0.3, 0.2, -1), castshadow=1)
  b.geom(mdl.GEOM_PLANE, (2, 2, 0.1), (0, 0, 0), matid=floor)
  b.geom(mdl.GEOM_SPHERE, (0.3, 0, 0), (0, 0, 0.3), matid=metal)
  b.geom(mdl.GEOM_BOX, (0.15, 0.2, 0.2), (0.6, 0.2, 0.2), rgba=(0.2, 0.8, 0.3, 1))
  frames = []
  if top_camera:
    b.camera((0, 0, 3), (0, 0.0001, 0), 45.0, RES)
    frames.append(_xyaxes((1, 0, 0), (0, 1, 0)))
  else:
    b.camera((0, -2, 1), (0, 0, 0), 50.0, RES)
    frames.append(_xyaxes((1, 0, 0), (0, 0.45, 0.9)))
    if side_camera:
      b.camera((1.5, -1.5, 1.2), (0, 0, 0), 70.0, RES)
      frames.append(_xyaxes((0.7, 0.7, 0), (-0.3, 0.3, 0.9)))
  mjm = b.build()
  mjm.cam_xmat[:] = np.stack(frames)
  return mjm


def pinhole_rays(width, height, fovy_deg, offset=0.5) -> np.ndarray:
  """(height, width, 2, 3) camera-frame rays of a pinhole camera (tests/test_camera_rays.py:47-54)."""
  h = np.tan(np.radians(fovy_deg) / 2.0)
  u = (np.arange(width) + offset) / width - 0.5
  v = (np.arange(height) + offset) / height - 0.5
  d = np.stack(np.broadcast_arrays(u[None, :] * 2.0 * h * width / height, -v[:, None] * 2.0 * h, -np.ones((1, 1))), axis=-1)
  d /= np.linalg.norm(d, axis=-1, keepdims=True)
  return np.stack([np.zeros_like(d), d], axis=2).astype(np.float32)


def _bundle(fovys=(50.0, 70.0)) -> np.ndarray:
  return np.stack([pinhole_rays(*RES, f) for f in fovys])


def _ray_scale(offset, fovys=(50.0, 70.0)) -> np.ndarray:
  """Distance along the ray per unit of planar depth, both cameras concatenated."""
  return np.concatenate([1.0 / -pinhole_rays(*RES, f, offset)[..., 1, 2].ravel() for f in fovys])


def _channels(rgb):
  rgb = rgb.view(np.uint32)
  return np.stack([(rgb >> s) & 0xFF for s in (16, 8, 0)], axis=-1).astype(np.int32)


def _oracle(mjm, rays=None, nworld=2, **kw):
  m, d, rc = make(mjm, nworld, device="cpu", cam_res=[RES, RES][: mjm.ncam], camera_rays=rays, **kw)
  render = oracle.pt_render if kw.get("tier") == mj.Tier.PATH else oracle.rt_render
  return m, d, rc, render(m, d, rc, seed=0)


@pytest.mark.parametrize(
  "options",
  [{}, {"samples_per_pixel": 2}, {"tier": mj.Tier.PATH, "path_samples": 4}],
  ids=["realtime", "supersampled", "path"],
)
def test_pinhole_rays_match_pinhole_cameras(options):
  """tests/test_camera_rays.py:71-91."""
  mjm = _scene()
  *_, a = _oracle(mjm, **options)
  *_, b = _oracle(mjm, rays=_bundle(), **options)
  diff = np.abs(_channels(a["rgb"]) - _channels(b["rgb"])).max(axis=-1)
  assert diff.max() <= 2 and np.mean(diff <= 1) > 0.99
  np.testing.assert_array_equal(a["seg"], b["seg"])
  offset = 0.25 if options.get("samples_per_pixel") == 2 else 0.5
  np.testing.assert_allclose(b["depth"], a["depth"] * _ray_scale(offset), rtol=1e-4, atol=1e-6)


def test_rays_are_updated_mirrored_and_removed():
  """tests/test_camera_rays.py:94-119."""
  w, h = RES
  mjm = _scene()
  m, d, rc, plain = _oracle(mjm, nworld=1)
  rays = _bundle()
  mj.set_camera_rays(rc, rays)
  with_rays = oracle.rt_render(m, d, rc)
  assert np.abs(_channels(with_rays["rgb"]) - _channels(plain["rgb"])).max() <= 1
  flipped = rays.copy()
  flipped[..., 0] *= -1.0
  mj.set_camera_rays(rc, [flipped[0], flipped[1]])
  mirrored = oracle.rt_render(m, d, rc)

  def first(a):
    return a[0, : w * h].reshape(h, w, *a.shape[2:])

  np.testing.assert_array_equal(first(mirrored["seg"])[:, ::-1], first(plain["seg"]))
  np.testing.assert_allclose(first(mirrored["depth"])[:, ::-1], first(plain["depth"]) * first(_ray_scale(0.5)[None]), rtol=1e-4)
  mj.set_camera_rays(rc, None)
  again = oracle.rt_render(m, d, rc)
  np.testing.assert_array_equal(again["rgb"], plain["rgb"])


def test_rays_without_direction_miss():
  """tests/test_camera_rays.py:122-134."""
  rays = _bundle()
  rays[:, :6, :, 1] = 0.0
  mjm = _scene()
  *_, ref = _oracle(mjm, nworld=1)
  *_, out = _oracle(mjm, rays=rays, nworld=1)
  w, h = RES
  for cam in range(2):
    px = slice(cam * w * h, (cam + 1) * w * h)
    seg = out["seg"][0, px].reshape(h, w, 2)
    assert np.all(seg[:6] == -1) and np.all(out["depth"][0, px].reshape(h, w)[:6] == 0.0)
    np.testing.assert_array_equal(seg[6:], ref["seg"][0, px].reshape(h, w, 2)[6:])


def test_parallel_rays_measure_distance_from_their_origin():
  """tests/test_camera_rays.py:137-157: origins on a grid 0.25 m in front of a camera 3 m up."""
  w, h = RES
  x = ((np.arange(w) + 0.5) / w - 0.5) * 1.5
  y = (0.5 - (np.arange(h) + 0.5) / h) * 1.0
  origin = np.stack(np.broadcast_arrays(x[None, :], y[:, None], np.full((1, 1), -0.25)), axis=-1)
  direction = np.broadcast_to([0.0, 0.0, -1.0], origin.shape)
  ortho = np.stack([origin, direction], axis=2).astype(np.float32)
  *_, out = _oracle(_scene(top_camera=True), rays=ortho[None], nworld=1)
  depth = out["depth"][0].reshape(h, w)
  seg = out["seg"][0].reshape(h, w, 2)
  floor = seg[..., 0] == 0
  assert floor.any()
  np.testing.assert_allclose(depth[floor], 2.75, atol=1e-4)
  sphere = seg[..., 0] == 1
  np.testing.assert_allclose(depth[sphere].min(), 3.0 - 0.25 - 0.6, atol=0.02)


def test_ray_validation():
  """tests/test_camera_rays.py:183-195."""
  m, d, rc = make(_scene(), 1, device="cpu", cam_res=[RES, RES])
  with pytest.raises(ValueError, match="each of the 2 rendering cameras"):
    mj.set_camera_rays(rc, _bundle()[:1])
  with pytest.raises(ValueError, match=r"shape \(16, 24, 2, 3\)"):
    mj.set_camera_rays(rc, [pinhole_rays(24, 16, 50.0), pinhole_rays(16, 16, 50.0)])
  with pytest.raises(ValueError, match="floating point"):
    mj.set_camera_rays(rc, np.zeros((2, 16, 24, 2, 3), np.int32))
  with pytest.raises(TypeError):
    mj.set_camera_rays(rc, [[1, 2], [3, 4]])
  assert rc.ray_direction is None


@pytest.mark.gpu
@pytest.mark.parametrize("spp", [1, 2])
def test_gpu_bundles_match_oracle(cuda, spp):
  rays = _bundle()
  rays[:, :, :, 1, 0] += 0.02 * np.sin(np.arange(RES[0]))[None, None, :]  # a non-pinhole bundle
  rays[:, 2:4, :, 1] = 0.0  # two rows without rays
  mjm = _scene()
  m, d, rc = make(mjm, 2, device=cuda, cam_res=[RES, RES], camera_rays=rays, samples_per_pixel=spp)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=0)
  compare_rt(gpu_outputs(rc), oracle.rt_render(m, d, rc, seed=0))


@pytest.mark.gpu
def test_gpu_path_traced_bundles_match_oracle(cuda):
  mjm = _scene()
  m, d, rc = make(mjm, 2, device=cuda, cam_res=[RES, RES], camera_rays=_bundle(), tier=mj.Tier.PATH, path_samples=2)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=3)
  gpu = gpu_outputs(rc)
  ref = oracle.pt_render(m, d, rc, seed=3)
  np.testing.assert_array_equal(gpu["seg"], ref["seg"])
  np.testing.assert_allclose(gpu["depth"], ref["depth"], rtol=1e-5, atol=1e-7)
  diff = np.abs(gpu["hdr"] - ref["hdr"]).max(axis=-1)
  assert np.mean(diff <= 1e-3) >= 0.99
