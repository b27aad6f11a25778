"""`mujoco.MjModel` / `mujoco.MjData` in, `HostModel` out: what a user of the reference passes to
`create_render_context`, `put_model` and `put_data` (reference io.py:331-509, which reads the
model's geom, material, light, camera, mesh, texture, height-field, visual and statistic fields
and runs `mj_forward` for the poses).

MuJoCo is not a dependency of this package: the adapter reads attributes by name from any
object shaped like an MjModel (and an MjData for poses), converting dtypes and shapes. Fields of
older MuJoCo releases are translated: `light_directional` (before `light_type`),
`cam_orthographic` (before `cam_projection`), a single `mat_texid` column (before texture
roles), and materials without `mat_metallic` / `mat_roughness` (unset, -1).
"""

from __future__ import annotations

import numpy as np

from mujoco_warp_pbr_amd import model as mdl

# HostModel field <- MjModel attribute, dtype (shapes follow from the model's counts)
_MODEL_FIELDS = {
  "geom_type": np.int32,
  "geom_size": np.float32,
  "geom_rgba": np.float32,
  "geom_matid": np.int32,
  "geom_dataid": np.int32,
  "geom_group": np.int32,
  "mat_rgba": np.float32,
  "mat_emission": np.float32,
  "mat_reflectance": np.float32,
  "mat_shininess": np.float32,
  "mat_specular": np.float32,
  "mat_texuniform": np.int32,
  "mat_texrepeat": np.float32,
  "light_castshadow": np.int32,
  "light_active": np.int32,
  "light_diffuse": np.float32,
  "light_specular": np.float32,
  "light_ambient": np.float32,
  "light_attenuation": np.float32,
  "light_cutoff": np.float32,
  "light_exponent": np.float32,
  "light_bulbradius": np.float32,
  "cam_fovy": np.float32,
  "cam_sensorsize": np.float32,
  "cam_intrinsic": np.float32,
  "cam_resolution": np.int32,
  "mesh_vertadr": np.int32,
  "mesh_vertnum": np.int32,
  "mesh_faceadr": np.int32,
  "mesh_facenum": np.int32,
  "mesh_normaladr": np.int32,
  "mesh_texcoordadr": np.int32,
  "mesh_vert": np.float32,
  "mesh_normal": np.float32,
  "mesh_face": np.int32,
  "mesh_facenormal": np.int32,
  "mesh_facetexcoord": np.int32,
  "mesh_texcoord": np.float32,
  "tex_type": np.int32,
  "tex_width": np.int32,
  "tex_height": np.int32,
  "tex_nchannel": np.int32,
  "tex_adr": np.int64,
  "tex_data": np.uint8,
  "hfield_size": np.float32,
  "hfield_nrow": np.int32,
  "hfield_ncol": np.int32,
  "hfield_adr": np.int32,
  "hfield_data": np.float32,
}
# the same from MjData (poses after mj_forward)
_DATA_FIELDS = ("geom_xpos", "geom_xmat", "light_xpos", "light_xdir", "cam_xpos", "cam_xmat")
# 2-D fields: their second dimension
_WIDTH = {
  "geom_size": 3, "geom_rgba": 4, "geom_xpos": 3, "geom_xmat": 9, "mat_rgba": 4, "mat_texrepeat": 2, "light_diffuse": 3,
  "light_specular": 3, "light_ambient": 3, "light_attenuation": 3, "light_xpos": 3, "light_xdir": 3, "cam_xpos": 3, "cam_xmat": 9,
  "cam_sensorsize": 2, "cam_intrinsic": 4, "cam_resolution": 2, "mesh_vert": 3, "mesh_normal": 3, "mesh_face": 3,
  "mesh_facenormal": 3, "mesh_facetexcoord": 3, "mesh_texcoord": 2, "hfield_size": 4,
}


def is_host_model(obj) -> bool:
  return isinstance(obj, mdl.HostModel)


def _array(obj, name, dtype, default=None):
  value = getattr(obj, name, None)
  if value is None:
    if default is None:
      raise ValueError(f"the model has no field {name!r}")
    value = default
  a = np.array(value, dtype=dtype, copy=True)
  if name in _WIDTH:
    a = a.reshape(-1, _WIDTH[name])
  else:
    a = a.reshape(-1)
  return a


def _poses(model, data):
  """MjData with the poses of mj_forward: `data` as given, else mj_forward at qpos0 through the
  mujoco package (which a caller holding an MjModel has)."""
  if data is not None:
    return data
  try:
    import mujoco  # noqa: PLC0415
  except ImportError as e:
    raise ValueError("an MjModel needs poses: pass its MjData after mujoco.mj_forward, or install mujoco") from e
  data = mujoco.MjData(model)
  mujoco.mj_forward(model, data)
  return data


def from_mujoco(model, data=None) -> mdl.HostModel:
  """HostModel of an object shaped like `mujoco.MjModel`, with the poses of `data` (an MjData
  after mj_forward; None runs mj_forward at qpos0)."""
  data = _poses(model, data)
  ngeom, nmat = int(model.ngeom), int(model.nmat)
  nlight, ncam = int(model.nlight), int(model.ncam)
  kw = {}
  for name, dtype in _MODEL_FIELDS.items():
    kw[name] = _array(model, name, dtype, default=_defaults(name, model))
  for name in _DATA_FIELDS:
    kw[name] = _array(data, name, np.float32)
  # materials: metallic / roughness (-1: unset) and per-role texture ids
  kw["mat_metallic"] = _array(model, "mat_metallic", np.float32, default=np.full(nmat, -1.0))
  kw["mat_roughness"] = _array(model, "mat_roughness", np.float32, default=np.full(nmat, -1.0))
  texid = np.array(getattr(model, "mat_texid"), dtype=np.int32)
  if texid.size == nmat and (texid.ndim == 1 or texid.shape[-1] == 1):
    roles = np.full((nmat, mdl.NTEXROLE), -1, np.int32)
    roles[:, mdl.ROLE_RGB] = texid.reshape(-1)
    texid = roles
  kw["mat_texid"] = texid.reshape(nmat, mdl.NTEXROLE)
  # lights: type (older releases: directional flag, other lights spot lights), intensity, range
  if getattr(model, "light_type", None) is not None:
    kw["light_type"] = _array(model, "light_type", np.int32)
  else:
    directional = _array(model, "light_directional", np.int32)
    kw["light_type"] = np.where(directional != 0, mdl.LIGHT_DIRECTIONAL, mdl.LIGHT_SPOT).astype(np.int32)
  kw["light_intensity"] = _array(model, "light_intensity", np.float32, default=np.zeros(nlight))
  kw["light_range"] = _array(model, "light_range", np.float32, default=np.full(nlight, 10.0))
  # cameras: projection (older releases: orthographic flag)
  proj = getattr(model, "cam_projection", None)
  kw["cam_projection"] = _array(model, "cam_projection" if proj is not None else "cam_orthographic", np.int32)
  kw["tex_colorspace"] = _array(model, "tex_colorspace", np.int32, default=np.full(int(model.ntex), mdl.COLORSPACE_AUTO))
  mesh_quat = getattr(model, "mesh_quat", None)
  kw["mesh_quat"] = None if mesh_quat is None else np.array(mesh_quat, np.float64).reshape(-1, 4)
  head = model.vis.headlight
  kw["headlight_active"] = bool(head.active)
  for k in ("ambient", "diffuse", "specular"):
    kw[f"headlight_{k}"] = np.array(getattr(head, k), np.float32).reshape(3)
  kw["stat_center"] = np.array(model.stat.center, np.float32).reshape(3)
  kw["stat_extent"] = float(model.stat.extent)
  host = mdl.HostModel(**kw)
  if host.ngeom != ngeom or host.ncam != ncam:
    raise ValueError("the model's geom or camera arrays disagree with ngeom / ncam")
  host.validate()
  return host


def _defaults(name, model):
  """Fields absent from some models: height fields (no height field) and sensor-size cameras."""
  n = {"hfield": int(getattr(model, "nhfield", 0)), "cam": int(model.ncam)}
  if name == "hfield_size":
    return np.zeros((n["hfield"], 4))
  if name.startswith("hfield_"):
    return np.zeros(n["hfield"])
  if name == "cam_sensorsize":
    return np.zeros((n["cam"], 2))
  if name == "cam_intrinsic":
    return np.zeros((n["cam"], 4))
  return None


def as_host_model(mjm, mjd=None) -> mdl.HostModel:
  """`mjm` itself when it is a HostModel, else its conversion (from_mujoco)."""
  return mjm if is_host_model(mjm) else from_mujoco(mjm, mjd)
