"""Host model and device model/data: the MuJoCo / MuJoCo Warp fields the renderer reads.

MuJoCo and MuJoCo Warp are not dependencies of this package. `HostModel` holds, in NumPy, the
fields of `mujoco.MjModel` that the reference's context creation reads (geoms, materials,
lights, cameras, meshes, textures, headlight, statistics; reference io.py:331-634,
material.py:93-182, texture.py:63-101). `put_model` / `put_data` give the device `Model` and
`Data` with MuJoCo Warp's field names and conventions: batched model arrays have a leading
dimension of 1 (replace them with leading size nworld to vary a value per world; kernels index
`worldid % shape[0]`), data arrays are per world.
"""

from __future__ import annotations

import dataclasses
from typing import Any

import numpy as np
import torch

# MuJoCo enums used by the renderer.
GEOM_PLANE, GEOM_HFIELD, GEOM_SPHERE, GEOM_CAPSULE, GEOM_ELLIPSOID, GEOM_CYLINDER, GEOM_BOX, GEOM_MESH, GEOM_SDF = range(9)
LIGHT_SPOT, LIGHT_DIRECTIONAL, LIGHT_POINT, LIGHT_IMAGE = range(4)
(
  ROLE_USER,
  ROLE_RGB,
  ROLE_OCCLUSION,
  ROLE_ROUGHNESS,
  ROLE_METALLIC,
  ROLE_NORMAL,
  ROLE_OPACITY,
  ROLE_EMISSIVE,
  ROLE_RGBA,
  ROLE_ORM,
) = range(10)
NTEXROLE = 10
TEXTURE_2D, TEXTURE_CUBE, TEXTURE_SKYBOX = range(3)
COLORSPACE_AUTO, COLORSPACE_LINEAR, COLORSPACE_SRGB = range(3)
OBJ_GEOM = 5


def _f32(x, shape=None):
  a = np.asarray(x, dtype=np.float32)
  return a.reshape(shape) if shape is not None else a


def _i32(x, shape=None):
  a = np.asarray(x, dtype=np.int32)
  return a.reshape(shape) if shape is not None else a


@dataclasses.dataclass
class HostModel:
  """The `mujoco.MjModel` fields the renderer reads (poses are those of `mj_forward` at qpos0)."""

  # geoms
  geom_type: np.ndarray
  geom_size: np.ndarray  # (ngeom, 3)
  geom_rgba: np.ndarray  # (ngeom, 4)
  geom_matid: np.ndarray
  geom_dataid: np.ndarray
  geom_group: np.ndarray
  geom_xpos: np.ndarray  # (ngeom, 3)
  geom_xmat: np.ndarray  # (ngeom, 9) row-major
  # materials
  mat_rgba: np.ndarray  # (nmat, 4)
  mat_emission: np.ndarray
  mat_reflectance: np.ndarray
  mat_metallic: np.ndarray  # -1: unset
  mat_roughness: np.ndarray  # -1: unset
  mat_shininess: np.ndarray
  mat_specular: np.ndarray
  mat_texid: np.ndarray  # (nmat, 10)
  mat_texuniform: np.ndarray
  mat_texrepeat: np.ndarray  # (nmat, 2)
  # lights
  light_type: np.ndarray
  light_castshadow: np.ndarray
  light_active: np.ndarray
  light_xpos: np.ndarray
  light_xdir: np.ndarray
  light_diffuse: np.ndarray
  light_specular: np.ndarray
  light_ambient: np.ndarray
  light_attenuation: np.ndarray
  light_cutoff: np.ndarray
  light_exponent: np.ndarray
  light_intensity: np.ndarray
  light_range: np.ndarray
  light_bulbradius: np.ndarray
  # cameras
  cam_xpos: np.ndarray
  cam_xmat: np.ndarray  # (ncam, 9)
  cam_fovy: np.ndarray
  cam_projection: np.ndarray
  cam_sensorsize: np.ndarray  # (ncam, 2)
  cam_intrinsic: np.ndarray  # (ncam, 4)
  cam_resolution: np.ndarray  # (ncam, 2)
  # meshes (mesh frame = geom frame)
  mesh_vertadr: np.ndarray
  mesh_vertnum: np.ndarray
  mesh_faceadr: np.ndarray
  mesh_facenum: np.ndarray
  mesh_normaladr: np.ndarray
  mesh_texcoordadr: np.ndarray  # -1: none
  mesh_vert: np.ndarray
  mesh_normal: np.ndarray
  mesh_face: np.ndarray
  mesh_facenormal: np.ndarray
  mesh_facetexcoord: np.ndarray
  mesh_texcoord: np.ndarray
  # textures
  tex_type: np.ndarray
  tex_width: np.ndarray
  tex_height: np.ndarray
  tex_nchannel: np.ndarray
  tex_adr: np.ndarray
  tex_data: np.ndarray  # uint8
  tex_colorspace: np.ndarray
  # visual / statistics
  headlight_active: bool = False
  headlight_ambient: np.ndarray = dataclasses.field(default_factory=lambda: np.full(3, 0.1, np.float32))
  headlight_diffuse: np.ndarray = dataclasses.field(default_factory=lambda: np.full(3, 0.4, np.float32))
  headlight_specular: np.ndarray = dataclasses.field(default_factory=lambda: np.full(3, 0.5, np.float32))
  stat_center: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3, np.float32))
  stat_extent: float = 1.0
  # height fields: (radius x, radius y, elevation z, base z), grids of normalised elevations
  hfield_size: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 4), np.float32))
  hfield_nrow: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
  hfield_ncol: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
  hfield_adr: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
  hfield_data: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.float32))
  # mesh_quat: rotation (w, x, y, z) from a mesh's MuJoCo frame (its principal inertia frame,
  # where mesh_vert lives) back to its authored frame; None: identity for every mesh
  mesh_quat: np.ndarray | None = None

  @property
  def nhfield(self) -> int:
    return int(self.hfield_nrow.shape[0])

  @property
  def ngeom(self) -> int:
    return int(self.geom_type.shape[0])

  @property
  def nmat(self) -> int:
    return int(self.mat_rgba.shape[0])

  @property
  def nlight(self) -> int:
    return int(self.light_type.shape[0])

  @property
  def ncam(self) -> int:
    return int(self.cam_xpos.shape[0])

  @property
  def nmesh(self) -> int:
    return int(self.mesh_vertadr.shape[0])

  @property
  def ntex(self) -> int:
    return int(self.tex_type.shape[0])

  def validate(self) -> None:
    """Raises ValueError when array shapes disagree."""
    n = self.ngeom
    for name, shape in (("geom_size", (n, 3)), ("geom_rgba", (n, 4)), ("geom_xpos", (n, 3)), ("geom_xmat", (n, 9))):
      if getattr(self, name).shape != shape:
        raise ValueError(f"{name} must have shape {shape}, got {getattr(self, name).shape}")
    if self.mat_texid.shape != (self.nmat, NTEXROLE):
      raise ValueError(f"mat_texid must have shape ({self.nmat}, {NTEXROLE})")
    if np.any(self.geom_matid >= self.nmat):
      raise ValueError("geom_matid out of range")
    nh = self.nhfield
    if self.hfield_size.shape != (nh, 4) or self.hfield_ncol.shape != (nh,) or self.hfield_adr.shape != (nh,):
      raise ValueError("hfield_size, hfield_nrow, hfield_ncol and hfield_adr must describe the same height fields")
    for h in range(nh):
      nr, nc, adr = int(self.hfield_nrow[h]), int(self.hfield_ncol[h]), int(self.hfield_adr[h])
      if nr < 2 or nc < 2 or adr < 0 or adr + nr * nc > self.hfield_data.shape[0]:
        raise ValueError(f"height field {h}: needs at least 2 x 2 samples inside hfield_data")
    for kind, count in ((GEOM_HFIELD, nh), (GEOM_MESH, self.nmesh), (GEOM_SDF, self.nmesh)):
      ids = self.geom_dataid[self.geom_type == kind]
      if np.any(ids >= count):
        raise ValueError("geom_dataid out of range")


def hfield_mesh(mjm: HostModel, h: int) -> tuple[np.ndarray, np.ndarray]:
  """Closed triangle surface of height field `h` in its geom frame: (verts (n, 3), faces (m, 3)).

  Sample (r, c) sits at x = sx (2c / (ncol - 1) - 1), y = sy (2r / (nrow - 1) - 1), z = elevation
  * data[r, c] (MuJoCo's hfield layout); the solid extends down to z = -base. Each grid cell is
  two triangles, each boundary segment a wall quad, and the base one quad; every face is wound
  so that cross(v1 - v0, v2 - v0) points out of the solid.
  """
  nr, nc, adr = int(mjm.hfield_nrow[h]), int(mjm.hfield_ncol[h]), int(mjm.hfield_adr[h])
  sx, sy, sz, base = (float(v) for v in mjm.hfield_size[h])
  z = np.asarray(mjm.hfield_data[adr : adr + nr * nc], np.float64).reshape(nr, nc) * sz
  ys, xs = np.meshgrid(np.linspace(-sy, sy, nr), np.linspace(-sx, sx, nc), indexing="ij")
  top = np.stack([xs, ys, z], axis=-1).reshape(-1, 3)
  idx = np.arange(nr * nc).reshape(nr, nc)
  a, b, c, d = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, 1:].ravel(), idx[1:, :-1].ravel()
  faces = [np.stack([a, b, c], 1), np.stack([a, c, d], 1)]
  # boundary loop (counter-clockwise seen from above) and the base vertices below it
  ring = np.concatenate([idx[0, :], idx[1:, -1], idx[-1, -2::-1], idx[-2:0:-1, 0]])
  bottom = top[ring].copy()
  bottom[:, 2] = -base
  nb0 = len(top)
  nxt = np.roll(np.arange(len(ring)), -1)
  # wall quad per boundary segment: top[i], top[i+1], bottom[i+1], bottom[i]
  t0, t1, b0, b1 = ring, ring[nxt], nb0 + np.arange(len(ring)), nb0 + nxt
  faces += [np.stack([t0, b1, t1], 1), np.stack([t0, b0, b1], 1)]
  corners = nb0 + np.array([0, nc - 1, nc + nr - 2, 2 * nc + nr - 3])
  faces += [np.array([[corners[0], corners[2], corners[1]], [corners[0], corners[3], corners[2]]])]
  verts = np.concatenate([top, bottom]).astype(np.float32)
  return verts, np.concatenate(faces).astype(np.int32)


def trace_mesh_arrays(mjm: HostModel) -> dict[str, np.ndarray]:
  """The model's mesh arrays followed by one triangle mesh per height field: height field h is
  traced as mesh nmesh + h (flat-shaded, no vertex normals or texture coordinates)."""
  out = {k: np.asarray(getattr(mjm, k)) for k in PLAIN if k.startswith("mesh_")}
  out["mesh_vertnum"] = np.asarray(mjm.mesh_vertnum)
  if mjm.nhfield == 0:
    return out
  lists = {k: [v] for k, v in out.items()}
  nv, nf = len(out["mesh_vert"]), len(out["mesh_face"])
  nn = len(out["mesh_normal"])
  for h in range(mjm.nhfield):
    verts, faces = hfield_mesh(mjm, h)
    lists["mesh_vertadr"].append(np.array([nv], np.int32))
    lists["mesh_vertnum"].append(np.array([len(verts)], np.int32))
    lists["mesh_faceadr"].append(np.array([nf], np.int32))
    lists["mesh_facenum"].append(np.array([len(faces)], np.int32))
    lists["mesh_normaladr"].append(np.array([nn], np.int32))
    lists["mesh_texcoordadr"].append(np.array([-1], np.int32))
    lists["mesh_vert"].append(verts)
    lists["mesh_face"].append(faces)
    lists["mesh_facenormal"].append(np.zeros_like(faces))
    lists["mesh_facetexcoord"].append(np.zeros_like(faces))
    nv += len(verts)
    nf += len(faces)
  lists["mesh_normal"].append(np.zeros((1, 3), np.float32))
  return {k: np.concatenate(v).astype(out[k].dtype) for k, v in lists.items()}


def empty_host_model() -> dict[str, Any]:
  """Keyword arguments of a HostModel with no geoms, materials, lights, cameras, meshes, textures."""
  z3 = np.zeros((0, 3), np.float32)
  return dict(
    geom_type=_i32([]),
    geom_size=z3.copy(),
    geom_rgba=np.zeros((0, 4), np.float32),
    geom_matid=_i32([]),
    geom_dataid=_i32([]),
    geom_group=_i32([]),
    geom_xpos=z3.copy(),
    geom_xmat=np.zeros((0, 9), np.float32),
    mat_rgba=np.zeros((0, 4), np.float32),
    mat_emission=_f32([]),
    mat_reflectance=_f32([]),
    mat_metallic=_f32([]),
    mat_roughness=_f32([]),
    mat_shininess=_f32([]),
    mat_specular=_f32([]),
    mat_texid=np.zeros((0, NTEXROLE), np.int32),
    mat_texuniform=np.zeros(0, np.int32),
    mat_texrepeat=np.zeros((0, 2), np.float32),
    light_type=_i32([]),
    light_castshadow=_i32([]),
    light_active=_i32([]),
    light_xpos=z3.copy(),
    light_xdir=z3.copy(),
    light_diffuse=z3.copy(),
    light_specular=z3.copy(),
    light_ambient=z3.copy(),
    light_attenuation=z3.copy(),
    light_cutoff=_f32([]),
    light_exponent=_f32([]),
    light_intensity=_f32([]),
    light_range=_f32([]),
    light_bulbradius=_f32([]),
    cam_xpos=z3.copy(),
    cam_xmat=np.zeros((0, 9), np.float32),
    cam_fovy=_f32([]),
    cam_projection=_i32([]),
    cam_sensorsize=np.zeros((0, 2), np.float32),
    cam_intrinsic=np.zeros((0, 4), np.float32),
    cam_resolution=np.zeros((0, 2), np.int32),
    mesh_vertadr=_i32([]),
    mesh_vertnum=_i32([]),
    mesh_faceadr=_i32([]),
    mesh_facenum=_i32([]),
    mesh_normaladr=_i32([]),
    mesh_texcoordadr=_i32([]),
    mesh_vert=z3.copy(),
    mesh_normal=z3.copy(),
    mesh_face=np.zeros((0, 3), np.int32),
    mesh_facenormal=np.zeros((0, 3), np.int32),
    mesh_facetexcoord=np.zeros((0, 3), np.int32),
    mesh_texcoord=np.zeros((0, 2), np.float32),
    tex_type=_i32([]),
    tex_width=_i32([]),
    tex_height=_i32([]),
    tex_nchannel=_i32([]),
    tex_adr=np.zeros(0, np.int64),
    tex_data=np.zeros(0, np.uint8),
    tex_colorspace=_i32([]),
    hfield_size=np.zeros((0, 4), np.float32),
    hfield_nrow=_i32([]),
    hfield_ncol=_i32([]),
    hfield_adr=_i32([]),
    hfield_data=_f32([]),
  )


# Batched model fields: leading dimension 1 on device, indexed worldid % shape[0].
BATCHED = (
  "geom_size",
  "geom_rgba",
  "geom_matid",
  "geom_dataid",
  "mat_rgba",
  "mat_emission",
  "mat_texid",
  "mat_texuniform",
  "mat_texrepeat",
  "light_type",
  "light_castshadow",
  "light_active",
  "light_attenuation",
  "light_cutoff",
  "light_exponent",
  "light_diffuse",
  "light_specular",
  "cam_fovy",
  "cam_intrinsic",
)
# Unbatched model fields copied to device as they are.
PLAIN = (
  "geom_type",
  "cam_projection",
  "cam_sensorsize",
  "mesh_vertadr",
  "mesh_faceadr",
  "mesh_facenum",
  "mesh_normaladr",
  "mesh_texcoordadr",
  "mesh_vert",
  "mesh_normal",
  "mesh_face",
  "mesh_facenormal",
  "mesh_facetexcoord",
  "mesh_texcoord",
)


class Model:
  """Device model (MuJoCo Warp `Model` field names). Attributes are torch tensors."""

  def __init__(self, mjm: HostModel, device):
    self.ngeom, self.nmat, self.nlight, self.ncam, self.nmesh = mjm.ngeom, mjm.nmat, mjm.nlight, mjm.ncam, mjm.nmesh
    self.nhfield = mjm.nhfield
    for name in BATCHED:
      a = np.ascontiguousarray(getattr(mjm, name))
      setattr(self, name, torch.as_tensor(a[None].copy()).to(device))
    meshes = trace_mesh_arrays(mjm)
    for name in PLAIN:
      a = np.ascontiguousarray(meshes[name] if name in meshes else getattr(mjm, name))
      setattr(self, name, torch.as_tensor(a.copy()).to(device))


class Data:
  """Device data (MuJoCo Warp `Data` field names): per-world poses."""

  def __init__(self, mjm: HostModel, nworld: int, device):
    self.nworld = int(nworld)

    def tile(a, tail):
      t = torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).reshape(1, -1, tail)
      return t.expand(self.nworld, -1, -1).contiguous().to(device)

    self.geom_xpos = tile(mjm.geom_xpos, 3)
    self.geom_xmat = tile(mjm.geom_xmat, 9)
    self.cam_xpos = tile(mjm.cam_xpos, 3)
    self.cam_xmat = tile(mjm.cam_xmat, 9)
    self.light_xpos = tile(mjm.light_xpos, 3)
    self.light_xdir = tile(mjm.light_xdir, 3)


def put_model(mjm, device="cuda") -> Model:
  """Device model of a HostModel or of a `mujoco.MjModel` (MuJoCo Warp's put_model)."""
  from mujoco_warp_pbr_amd.mjmodel import as_host_model

  mjm = as_host_model(mjm)
  mjm.validate()
  return Model(mjm, device)


def put_data(mjm, *args, device="cuda", nworld=None) -> Data:
  """Per-world device data at the model's poses (MuJoCo Warp's put_data).

  put_data(mjm, nworld=1, device) with a HostModel (its poses), or MuJoCo Warp's form
  put_data(mjm, mjd, nworld=...) with a `mujoco.MjModel` and its MjData (the MjData's poses).
  """
  from mujoco_warp_pbr_amd.mjmodel import as_host_model

  rest = list(args)
  mjd = rest.pop(0) if rest and not isinstance(rest[0], (int, np.integer)) else None
  if rest:
    if nworld is not None:
      raise TypeError("put_data: nworld given twice")
    nworld = rest.pop(0)
  if rest:
    device = rest.pop(0)
  if rest:
    raise TypeError("put_data takes (mjm, [mjd], [nworld], [device])")
  nworld = 1 if nworld is None else int(nworld)
  if nworld < 1:
    raise ValueError(f"nworld must be at least 1, got {nworld}")
  return Data(as_host_model(mjm, mjd), nworld, device)
