"""Height fields and SDF geoms (traverse.py:161-165, 185-196).

A height field is traced as the closed triangle surface of its elevation grid (top, walls,
base) and shaded like a primitive (face normal, planar texture coordinates); an SDF geom is
traced as its mesh. MuJoCo Warp's own height-field triangulation is not available here, so
the geometry is pinned by closed forms: flat and sloped fields against analytic depths, the
walls and base against ray-slab distances, and the surface against the divergence theorem.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd.scenes import _Builder
from tests.helpers import Shot
from tests.helpers import compare_rt
from tests.helpers import xyaxes

SIZE = (1.0, 0.8, 0.3, 0.1)  # radius x, radius y, elevation, base
H = 2.0  # camera height


def _field(elevation, cam_pos=(0, 0, H), frame=None, fovy=40.0, extra=None, textured=False):
  b = _Builder()
  hf = b.hfield(np.asarray(elevation, np.float64), SIZE)
  texid = None
  if textured:
    checker = ((np.indices((32, 32)).sum(0) // 4) % 2 * 200 + 40).astype(np.uint8)[..., None].repeat(3, 2)
    texid = {mdl.ROLE_RGB: b.texture(checker, mdl.COLORSPACE_AUTO)}
  mat = b.material(rgba=(0.7, 0.6, 0.5, 1), roughness=0.8, texid=texid, repeat=(3, 2))
  b.geom(mdl.GEOM_HFIELD, (SIZE[0], SIZE[1], 0.5 * (SIZE[2] + SIZE[3])), (0, 0, 0), matid=mat, dataid=hf)
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (0.3, 0.2, -1), castshadow=1)
  if extra:
    extra(b)
  b.camera(cam_pos, (0, 0.0001, 0), fovy, (16, 16))
  mjm = b.build()
  mjm.cam_xmat[0] = np.eye(3, dtype=np.float32).reshape(9) if frame is None else frame
  return mjm


def _pixel_dirs(res, fovy):
  """Camera-frame (x, y) per unit of planar depth for each pixel centre, row-major."""
  w, h = res
  t = np.tan(np.radians(fovy) / 2)
  x = ((np.arange(w) + 0.5) / w - 0.5) * 2 * t * w / h
  y = (0.5 - (np.arange(h) + 0.5) / h) * 2 * t
  return np.broadcast_to(x[None, :], (h, w)), np.broadcast_to(y[:, None], (h, w))


def test_surface_is_closed_and_outward():
  """Every face wound outward: the divergence theorem gives base area x base depth plus the
  integral of the piecewise-linear elevation."""
  rng = np.random.default_rng(0)
  e = rng.uniform(0, 1, (5, 7))
  b = _Builder()
  b.hfield(e, SIZE)
  verts, faces = mdl.hfield_mesh(b.build(), 0)
  v = verts.astype(np.float64)
  volume = np.einsum("ij,ij->i", v[faces[:, 0]], np.cross(v[faces[:, 1]], v[faces[:, 2]])).sum() / 6
  nr, nc = e.shape
  cell = (2 * SIZE[0] / (nc - 1)) * (2 * SIZE[1] / (nr - 1))
  z = e * SIZE[2]
  # two triangles per cell: (a, b, c) and (a, c, d) each of area cell / 2
  a, bb, c, d = z[:-1, :-1], z[:-1, 1:], z[1:, 1:], z[1:, :-1]
  integral = (cell / 2 * ((a + bb + c) / 3 + (a + c + d) / 3)).sum()
  assert volume == pytest.approx(integral + SIZE[3] * 4 * SIZE[0] * SIZE[1], rel=1e-6)
  assert len(faces) == 2 * (nr - 1) * (nc - 1) + 2 * (2 * (nr + nc) - 4) + 2


def test_flat_field_depth_and_normal():
  """A constant elevation e is a plane at z = e * elevation: planar depth H - z everywhere."""
  s = Shot(_field(np.full((6, 9), 0.5)), cam_res=(16, 16)).render()
  seg = s.out["seg"][0]
  assert np.all(seg[:, 0] == 0) and np.all(seg[:, 1] == 5)
  np.testing.assert_allclose(s.out["depth"][0], H - 0.5 * SIZE[2], rtol=1e-6)
  # face normal +z: the same shading as a box top at the same height
  b = _Builder()
  mat = b.material(rgba=(0.7, 0.6, 0.5, 1), roughness=0.8)
  b.geom(mdl.GEOM_BOX, (5, 5, 0.5 * 0.5 * SIZE[2]), (0, 0, 0.25 * SIZE[2]), matid=mat)
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (0.3, 0.2, -1), castshadow=1)
  b.camera((0, 0, H), (0, 0.0001, 0), 40.0, (16, 16))
  box = b.build()
  box.cam_xmat[0] = np.eye(3, dtype=np.float32).reshape(9)
  ref = Shot(box, cam_res=(16, 16)).render()
  np.testing.assert_allclose(s.out["hdr"], ref.out["hdr"], atol=1e-5)


def test_sloped_field_matches_analytic_depth():
  """Elevation linear in x: z = k (x + sx), k = elevation / (2 sx); a ray of planar direction
  (dx, dy, -1) from height H hits at planar depth (H - k sx) / (1 + k dx)."""
  nr, nc = 4, 11
  e = np.tile(np.linspace(0, 1, nc), (nr, 1))
  s = Shot(_field(e), cam_res=(16, 16)).render()
  dx, _ = _pixel_dirs((16, 16), 40.0)
  k = SIZE[2] / (2 * SIZE[0])
  expected = (H - k * SIZE[0]) / (1 + k * dx)
  depth = s.out["depth"][0].reshape(16, 16)
  hit = s.out["seg"][0].reshape(16, 16, 2)[..., 0] == 0
  assert hit.all()
  np.testing.assert_allclose(depth, expected, rtol=1e-5)


def test_walls_and_base_close_the_solid():
  """Looking along +x at the -x wall and up at the base: distances of the slab planes."""
  e = np.full((3, 3), 0.6)
  # camera at x = -3 looking along +x at mid-height of the wall (z in (-base, 0.6 * elevation))
  wall = _field(e, cam_pos=(-3.0, 0, 0.0), frame=xyaxes((0, -1, 0), (0, 0, 1)), fovy=5.0)
  s = Shot(wall, cam_res=(16, 16)).render()
  np.testing.assert_allclose(s.out["depth"][0], 3.0 - SIZE[0], rtol=1e-5)
  # camera below the base looking up (+z): planar depth to z = -base
  below = _field(e, cam_pos=(0, 0, -1.0), frame=xyaxes((1, 0, 0), (0, -1, 0)), fovy=20.0)
  s = Shot(below, cam_res=(16, 16)).render()
  np.testing.assert_allclose(s.out["depth"][0], 1.0 - SIZE[3], rtol=1e-5)
  assert np.all(s.out["seg"][0][:, 0] == 0)


def test_sdf_geom_traces_its_mesh():
  """An SDF geom renders exactly like a mesh geom over the same mesh."""
  from mujoco_warp_pbr_amd import scenes

  def scene(kind):
    b = _Builder()
    rng = np.random.default_rng(3)
    mesh = b.mesh(*scenes.convex_hull_mesh(rng, 200, 0.3))
    b.geom(mdl.GEOM_PLANE, (3, 3, 0.1), (0, 0, 0))
    b.geom(kind, (0, 0, 0), (0, 0, 0.35), dataid=mesh, rgba=(0.8, 0.3, 0.2, 1))
    b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (0.3, 0.2, -1), castshadow=1)
    b.camera((0, -1.5, 1.0), (0, 0, 0.3), 45.0, (16, 16))
    return b.build()

  a = Shot(scene(mdl.GEOM_MESH), cam_res=(16, 16)).render().out
  b = Shot(scene(mdl.GEOM_SDF), cam_res=(16, 16)).render().out
  assert np.sum(a["seg"][0][:, 0] == 1) > 20
  for key in ("seg", "depth", "hdr", "rgb"):
    np.testing.assert_array_equal(a[key], b[key])


def test_hfield_validation():
  b = _Builder()
  b.hfield(np.zeros((1, 5)), SIZE)
  with pytest.raises(ValueError, match="at least 2 x 2"):
    b.build().validate()
  b = _Builder()
  b.geom(mdl.GEOM_HFIELD, (1, 1, 1), (0, 0, 0), dataid=0)
  with pytest.raises(ValueError, match="geom_dataid out of range"):
    b.build().validate()


def _terrain_scene():
  """Rolling terrain under a sphere and a box, textured, with a point light casting shadows."""
  rng = np.random.default_rng(7)
  y, x = np.mgrid[0:24, 0:32] / 8.0
  e = 0.5 + 0.25 * np.sin(x * 1.7) * np.cos(y * 1.3) + 0.05 * rng.uniform(-1, 1, x.shape)

  def extra(b):
    b.geom(mdl.GEOM_SPHERE, (0.2, 0, 0), (0.2, 0.1, 0.45))
    b.geom(mdl.GEOM_BOX, (0.1, 0.15, 0.1), (-0.4, -0.2, 0.35), rgba=(0.2, 0.5, 0.8, 1))
    b.light(mdl.LIGHT_POINT, (0.5, -0.5, 1.5), (0, 0, -1), castshadow=1, bulbradius=0.05)

  return _field(e, cam_pos=(0, -1.8, 1.4), frame=xyaxes((1, 0, 0), (0, 0.6, 0.8)), fovy=50.0, extra=extra, textured=True)


def test_terrain_scene_renders_on_the_oracle():
  """The textured terrain of the GPU parity tests: the field fills most of the view around the
  sphere and the box, and the checker texture (planar coordinates) varies its albedo."""
  s = Shot(_terrain_scene(), cam_res=(32, 32), render_guides=True).render()
  seg = s.out["seg"][0].reshape(32, 32, 2)[..., 0]
  assert np.mean(seg == 0) > 0.5 and np.any(seg == 1) and np.any(seg == 2)
  albedo = s.out["albedo"][0].reshape(32, 32, 3)[seg == 0]
  assert albedo[:, 0].max() - albedo[:, 0].min() > 0.2


@pytest.mark.gpu
@pytest.mark.parametrize("spp", [1, 2])
def test_gpu_hfield_matches_oracle(cuda, spp):
  s = Shot(_terrain_scene(), cuda, nworld=2, cam_res=(32, 32), samples_per_pixel=spp, shadow_samples=4).render(seed=1)
  compare_rt(s.out, oracle.rt_render(s.m, s.d, s.rc, seed=1))


@pytest.mark.gpu
def test_gpu_hfield_path_traced_matches_oracle(cuda):
  import mujoco_warp_pbr_amd as mj

  s = Shot(_terrain_scene(), cuda, nworld=2, cam_res=(32, 32), tier=mj.Tier.PATH, path_samples=2).render(seed=4)
  ref = oracle.pt_render(s.m, s.d, s.rc, seed=4)
  np.testing.assert_array_equal(s.out["seg"], ref["seg"])
  np.testing.assert_allclose(s.out["depth"], ref["depth"], rtol=1e-5, atol=1e-7)
  diff = np.abs(s.out["hdr"] - ref["hdr"]).max(axis=-1)
  assert np.mean(diff <= 1e-3) >= 0.99
