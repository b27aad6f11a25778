"""The MjModel adapter: objects shaped like `mujoco.MjModel` / `mujoco.MjData` go straight into
create_render_context, put_model and put_data (reference io.py:331-509 reads these fields).

MuJoCo is not installed here, so the models are stand-ins carrying MjModel's attribute names,
shapes and nested `vis.headlight` / `stat` records, built from the package's own scenes; the
adapter must give back exactly the HostModel they were built from.
"""

from __future__ import annotations

import sys
import types

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
from mujoco_warp_pbr_amd import mjmodel
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd import scenes

POSES = ("geom_xpos", "geom_xmat", "light_xpos", "light_xdir", "cam_xpos", "cam_xmat")


def _as_mujoco(host: mdl.HostModel, old_release: bool = False):
  """(model, data) stand-ins for host: MjModel fields (flat xmat rows as MuJoCo stores them),
  counts, vis.headlight and stat; MjData with the poses."""
  model = types.SimpleNamespace()
  for f in host.__dataclass_fields__:
    if f in POSES or f.startswith(("headlight_", "stat_")):
      continue
    setattr(model, f, None if getattr(host, f) is None else np.array(getattr(host, f)))
  for k in ("ngeom", "nmat", "nlight", "ncam", "nmesh", "ntex", "nhfield"):
    setattr(model, k, getattr(host, k))
  model.nflex = 0
  model.vis = types.SimpleNamespace(headlight=types.SimpleNamespace(
    active=int(host.headlight_active), ambient=host.headlight_ambient.copy(), diffuse=host.headlight_diffuse.copy(),
    specular=host.headlight_specular.copy()))
  model.stat = types.SimpleNamespace(center=host.stat_center.astype(np.float64), extent=np.float64(host.stat_extent))
  model.geom_size = model.geom_size.astype(np.float64)  # MuJoCo stores mjtNum
  if old_release:
    model.light_directional = (model.light_type == mdl.LIGHT_DIRECTIONAL).astype(np.uint8)
    del model.light_type
    model.cam_orthographic = model.cam_projection
    del model.cam_projection
    del model.mat_metallic, model.mat_roughness, model.light_intensity, model.light_range, model.tex_colorspace
    model.mat_texid = model.mat_texid[:, mdl.ROLE_RGB].copy()
  data = types.SimpleNamespace(**{f: np.array(getattr(host, f), np.float64) for f in POSES})
  return model, data


def _host_equal(a: mdl.HostModel, b: mdl.HostModel, skip=()):
  for f in a.__dataclass_fields__:
    if f in skip:
      continue
    x, y = getattr(a, f), getattr(b, f)
    if x is None or y is None:
      assert x is None and y is None, f
      continue
    np.testing.assert_array_equal(np.asarray(x), np.asarray(y), err_msg=f)
    assert np.asarray(x).dtype == np.asarray(y).dtype or np.isscalar(x), f


@pytest.mark.parametrize("scene", ["full", "mesh", "hfield"])
def test_mjmodel_fields_round_trip(scene):
  host = {"full": lambda: scenes.full_scene(2, 16, 32), "mesh": lambda: scenes.mesh_scene(1, 16, 3, 2, 200),
          "hfield": lambda: _hfield_scene()}[scene]()
  model, data = _as_mujoco(host)
  got = mjmodel.from_mujoco(model, data)
  _host_equal(got, host)


def _hfield_scene():
  b = scenes._Builder()
  h = b.hfield(np.random.default_rng(0).uniform(0, 1, (9, 7)).astype(np.float32), (1.0, 0.8, 0.3, 0.1))
  b.geom(mdl.GEOM_HFIELD, (0, 0, 0), (0, 0, 0), dataid=h)
  b.geom(mdl.GEOM_SPHERE, (0.2, 0, 0), (0, 0, 0.5))
  b.light(mdl.LIGHT_POINT, (0, 0, 2), (0, 0, -1))
  b.camera((0, -2, 1.5), (0, 0, 0), 45.0, (16, 16))
  return b.build()


def test_older_mujoco_fields_are_translated():
  """light_directional -> light_type (non-directional lights were spot lights), cam_orthographic
  -> cam_projection, one texture id a material -> the RGB role, metallic / roughness unset."""
  host = scenes.full_scene(3, 16, 32)
  model, data = _as_mujoco(host, old_release=True)
  got = mjmodel.from_mujoco(model, data)
  want_type = np.where(host.light_type == mdl.LIGHT_DIRECTIONAL, mdl.LIGHT_DIRECTIONAL, mdl.LIGHT_SPOT)
  np.testing.assert_array_equal(got.light_type, want_type)
  np.testing.assert_array_equal(got.cam_projection, host.cam_projection)
  np.testing.assert_array_equal(got.mat_texid[:, mdl.ROLE_RGB], host.mat_texid[:, mdl.ROLE_RGB])
  assert np.all(np.delete(got.mat_texid, mdl.ROLE_RGB, axis=1) == -1)
  assert np.all(got.mat_metallic == -1) and np.all(got.mat_roughness == -1)
  assert np.all(got.light_intensity == 0) and np.all(got.tex_colorspace == mdl.COLORSPACE_AUTO)
  _host_equal(got, host, skip=("light_type", "mat_texid", "mat_metallic", "mat_roughness", "light_intensity", "light_range",
                               "tex_colorspace"))


def test_missing_poses_need_mujoco(monkeypatch):
  model, _ = _as_mujoco(scenes.primitives_scene(0, 16))
  monkeypatch.setitem(sys.modules, "mujoco", None)  # import mujoco fails
  with pytest.raises(ValueError, match="MjData"):
    mjmodel.from_mujoco(model)


def _fake_mujoco(data):
  """A `mujoco` module whose mj_forward leaves the given poses in the MjData."""
  mod = types.ModuleType("mujoco")

  class MjData:
    def __init__(self, m):
      self.model = m

  def mj_forward(m, d):
    for f in POSES:
      setattr(d, f, getattr(data, f).copy())

  mod.MjData, mod.mj_forward = MjData, mj_forward
  return mod


def test_mjmodel_goes_straight_into_the_api(monkeypatch):
  """create_render_context (poses from mj_forward), put_model and MuJoCo Warp's
  put_data(mjm, mjd, nworld) accept the MjModel; the results equal those of the HostModel."""
  host = scenes.primitives_scene(1, 16)
  model, data = _as_mujoco(host)
  monkeypatch.setitem(sys.modules, "mujoco", _fake_mujoco(data))
  rc = mj.create_render_context(model, nworld=2, cam_res=(16, 16), device="cpu", precompute=False)
  rc_host = mj.create_render_context(host, nworld=2, cam_res=(16, 16), device="cpu", precompute=False)
  for name in ("light_intensity", "light_range", "light_bulbradius", "geom_metallic", "geom_roughness", "geom_reflectance", "exposure"):
    torch.testing.assert_close(getattr(rc, name), getattr(rc_host, name))
  torch.testing.assert_close(rc.base.tlas_box, rc_host.base.tlas_box)
  m, m_host = mj.put_model(model, "cpu"), mj.put_model(host, "cpu")
  torch.testing.assert_close(m.geom_size, m_host.geom_size)
  d = mj.put_data(model, data, nworld=3, device="cpu")
  d_host = mj.put_data(host, 3, "cpu")
  assert d.nworld == 3
  torch.testing.assert_close(d.geom_xpos, d_host.geom_xpos)
  torch.testing.assert_close(d.geom_xmat, d_host.geom_xmat)
  with pytest.raises(TypeError):
    mj.put_data(host, 2, "cpu", "extra")
