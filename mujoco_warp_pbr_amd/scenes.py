"""Seeded synthetic scenes of the benchmark configurations, emitted directly as HostModels.

- `primitives_scene` (small real-time case): ground plane + 2 spheres, 2 boxes, 2 capsules,
  2 cylinders, one directional and one point light, no textures, metallic 0, roughness
  U[0.2, 1], one 64x64 pinhole camera with fovy 45 degrees; lit by a constant grey sky
  (`constant_sky`).
- `full_scene` (full-feature real-time case): mirror plane (reflectance 0.3), 12 primitives,
  4 random convex-hull meshes of 500-2000 triangles, procedural 256x256 base / ORM / normal
  textures, a directional sun and a soft point light (bulb radius 0.05), one 128x128 camera;
  lit by `procedural_sky` (512x256 log-normal radiance plus a sun lobe).
- `mesh_scene` (denoised path tracing at scale): 64 instances of 16 procedural meshes
  (~200k triangles), textures and the procedural sky.
- `randomize_worlds`: per-world roughness U[0.05, 1], metallic Bernoulli(0.3), rgba U[0, 1].

Every generator is a pure function of its seed.
"""

from __future__ import annotations

import math

import numpy as np

from mujoco_warp_pbr_amd import model as mdl


def random_rotation(rng: np.random.Generator) -> np.ndarray:
  """Uniform random rotation matrix (3, 3)."""
  q = rng.normal(size=4)
  q /= np.linalg.norm(q)
  w, x, y, z = q
  return np.array(
    [
      [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ]
  )


def look_at(pos, target, up=(0.0, 0.0, 1.0)) -> np.ndarray:
  """Camera frame (3, 3) at `pos` looking at `target`: columns x right, y up, z backwards."""
  pos, target, up = (np.asarray(v, np.float64) for v in (pos, target, up))
  z = pos - target
  z /= np.linalg.norm(z)
  x = np.cross(up, z)
  x /= np.linalg.norm(x)
  y = np.cross(z, x)
  return np.stack([x, y, z], axis=1)


def constant_sky(value: float = 0.5, height: int = 16, width: int = 32) -> np.ndarray:
  return np.full((height, width, 3), value, np.float32)


def procedural_sky(seed: int = 0, height: int = 256, width: int = 512) -> np.ndarray:
  """Equirect HDR sky (height, width, 3) f32: log-normal radiance field plus a sun lobe."""
  rng = np.random.default_rng(seed)
  theta = (np.arange(height) + 0.5) / height * np.pi
  phi = ((np.arange(width) + 0.5) / width - 0.5) * 2.0 * np.pi
  st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
  d = np.stack(np.broadcast_arrays(st * np.cos(phi)[None], st * np.sin(phi)[None], ct), axis=-1)
  # smooth log-normal field from a few random low-frequency waves
  field = np.zeros((height, width))
  for _ in range(8):
    k = rng.normal(size=3) * rng.uniform(1.0, 4.0)
    field += np.sin(d @ k + rng.uniform(0, 2 * np.pi)) * rng.uniform(0.1, 0.35)
  sky_tint = np.array([0.55, 0.7, 1.0]) * (0.35 + 0.65 * np.clip(d[..., 2:3], 0.0, 1.0))
  ground_tint = np.array([0.35, 0.3, 0.25])
  base = np.where(d[..., 2:3] >= 0.0, sky_tint, ground_tint)
  radiance = base * np.exp(field)[..., None]
  sun_dir = np.array([0.4, -0.3, 0.86])
  sun_dir /= np.linalg.norm(sun_dir)
  cosang = np.clip(d @ sun_dir, -1.0, 1.0)
  radiance += np.array([30.0, 28.0, 24.0]) * np.exp((cosang - 1.0) / 0.0005)[..., None]
  radiance += np.array([1.0, 0.9, 0.7]) * np.exp((cosang - 1.0) / 0.05)[..., None]
  return radiance.astype(np.float32)


def _value_noise(rng: np.random.Generator, size: int, cells: int) -> np.ndarray:
  grid = rng.uniform(0.0, 1.0, (cells + 1, cells + 1))
  grid[-1, :] = grid[0, :]
  grid[:, -1] = grid[:, 0]
  t = (np.arange(size) + 0.5) / size * cells
  i = np.floor(t).astype(int)
  f = t - i
  f = f * f * (3 - 2 * f)
  a = grid[i][:, i] * (1 - f)[None, :] + grid[i][:, i + 1] * f[None, :]
  b = grid[i + 1][:, i] * (1 - f)[None, :] + grid[i + 1][:, i + 1] * f[None, :]
  return a * (1 - f)[:, None] + b * f[:, None]


def procedural_textures(seed: int = 0, size: int = 256) -> list[tuple[np.ndarray, int]]:
  """[(base RGBA8 sRGB), (ORM RGB8 linear), (normal RGB8 linear)] as (image, colorspace)."""
  rng = np.random.default_rng(seed)
  y, x = np.mgrid[0:size, 0:size]
  checker = ((x // (size // 8) + y // (size // 8)) % 2).astype(np.float64)
  noise = 0.6 * _value_noise(rng, size, 8) + 0.4 * _value_noise(rng, size, 32)
  c1, c2 = np.array([0.85, 0.8, 0.7]), np.array([0.25, 0.35, 0.55])
  base = (checker[..., None] * c1 + (1 - checker[..., None]) * c2) * (0.75 + 0.5 * noise[..., None])
  base_rgba = np.concatenate([np.clip(base, 0, 1), np.ones((size, size, 1))], axis=-1)
  base_u8 = np.round(base_rgba * 255).astype(np.uint8)
  occ = 0.7 + 0.3 * _value_noise(rng, size, 16)
  rough = 0.3 + 0.7 * _value_noise(rng, size, 16)
  metal = (_value_noise(rng, size, 4) > 0.7).astype(np.float64)
  orm_u8 = np.round(np.stack([occ, rough, metal], -1) * 255).astype(np.uint8)
  height = _value_noise(rng, size, 16) + 0.5 * _value_noise(rng, size, 48)
  gx = (np.roll(height, -1, 1) - np.roll(height, 1, 1)) * size / 64.0
  gy = (np.roll(height, -1, 0) - np.roll(height, 1, 0)) * size / 64.0
  n = np.stack([-gx, gy, np.ones_like(gx)], -1)
  n /= np.linalg.norm(n, axis=-1, keepdims=True)
  normal_u8 = np.round((n * 0.5 + 0.5) * 255).astype(np.uint8)
  return [(base_u8, mdl.COLORSPACE_AUTO), (orm_u8, mdl.COLORSPACE_AUTO), (normal_u8, mdl.COLORSPACE_AUTO)]


def convex_hull_mesh(rng: np.random.Generator, ntri: int, radius: float):
  """Convex hull of random points on an ellipsoid: every point is a hull vertex, so the hull has
  2 * points - 4 = `ntri` triangles (rounded down to even). (vertices, faces CCW outward)."""
  from scipy.spatial import ConvexHull

  npts = max(ntri // 2 + 2, 8)
  axes = rng.uniform(0.6, 1.0, 3) * radius
  p = rng.normal(size=(npts, 3))
  p /= np.linalg.norm(p, axis=1, keepdims=True)
  p *= axes
  hull = ConvexHull(p)
  verts = p[hull.vertices]
  remap = -np.ones(len(p), dtype=np.int64)
  remap[hull.vertices] = np.arange(len(hull.vertices))
  faces = remap[hull.simplices]
  centre = verts.mean(axis=0)
  verts = verts - centre
  tri = verts[faces]
  n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
  flip = np.einsum("ij,ij->i", n, tri.mean(axis=1)) < 0
  faces[flip] = faces[flip][:, [0, 2, 1]]
  return verts.astype(np.float32), faces.astype(np.int32)


def _mesh_attributes(verts: np.ndarray, faces: np.ndarray):
  """Smooth vertex normals (area weighted) and per-corner spherical texture coordinates."""
  tri = verts[faces].astype(np.float64)
  fn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
  vn = np.zeros_like(verts, dtype=np.float64)
  for k in range(3):
    np.add.at(vn, faces[:, k], fn)
  vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-12)
  d = tri / np.maximum(np.linalg.norm(tri, axis=-1, keepdims=True), 1e-12)
  u = np.arctan2(d[..., 1], d[..., 0]) / (2 * np.pi) + 0.5
  v = np.arccos(np.clip(d[..., 2], -1, 1)) / np.pi
  span = u.max(axis=1) - u.min(axis=1)
  seam = span > 0.5
  u[seam] = np.where(u[seam] < 0.5, u[seam] + 1.0, u[seam])
  tc = np.stack([u, v], -1).reshape(-1, 2)
  return vn.astype(np.float32), tc.astype(np.float32)


class _Builder:
  """Accumulates geoms, materials, lights, cameras, meshes and textures into a HostModel."""

  def __init__(self):
    self.kw = mdl.empty_host_model()
    self.lists = {k: [] for k in self.kw}

  def add(self, **fields):
    for k, v in fields.items():
      self.lists[k].append(np.asarray(v))

  def geom(self, type_, size, pos, mat=None, rgba=(0.5, 0.5, 0.5, 1.0), matid=-1, dataid=-1, group=0):
    mat = np.eye(3) if mat is None else np.asarray(mat)
    self.add(
      geom_type=type_,
      geom_size=np.asarray(size, np.float32),
      geom_rgba=np.asarray(rgba, np.float32),
      geom_matid=matid,
      geom_dataid=dataid,
      geom_group=group,
      geom_xpos=np.asarray(pos, np.float32),
      geom_xmat=mat.astype(np.float32).reshape(9),
    )
    return len(self.lists["geom_type"]) - 1

  def material(self, rgba=(1, 1, 1, 1), metallic=-1.0, roughness=-1.0, reflectance=0.0, emission=0.0, texid=None, repeat=(1, 1), uniform=0):
    tid = np.full(mdl.NTEXROLE, -1, np.int32)
    for role, t in (texid or {}).items():
      tid[role] = t
    self.add(
      mat_rgba=np.asarray(rgba, np.float32),
      mat_emission=emission,
      mat_reflectance=reflectance,
      mat_metallic=metallic,
      mat_roughness=roughness,
      mat_shininess=0.5,
      mat_specular=0.5,
      mat_texid=tid,
      mat_texuniform=uniform,
      mat_texrepeat=np.asarray(repeat, np.float32),
    )
    return len(self.lists["mat_rgba"]) - 1

  def light(self, type_, pos, direction, diffuse=(1, 1, 1), castshadow=1, intensity=0.0, range_=10.0, bulbradius=0.02):
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    self.add(
      light_type=type_,
      light_castshadow=castshadow,
      light_active=1,
      light_xpos=np.asarray(pos, np.float32),
      light_xdir=d.astype(np.float32),
      light_diffuse=np.asarray(diffuse, np.float32),
      light_specular=np.full(3, 0.3, np.float32),
      light_ambient=np.zeros(3, np.float32),
      light_attenuation=np.array([1.0, 0.0, 0.0], np.float32),
      light_cutoff=45.0,
      light_exponent=10.0,
      light_intensity=intensity,
      light_range=range_,
      light_bulbradius=bulbradius,
    )

  def camera(self, pos, target, fovy, res):
    self.add(
      cam_xpos=np.asarray(pos, np.float32),
      cam_xmat=look_at(pos, target).astype(np.float32).reshape(9),
      cam_fovy=fovy,
      cam_projection=0,
      cam_sensorsize=np.zeros(2, np.float32),
      cam_intrinsic=np.zeros(4, np.float32),
      cam_resolution=np.asarray(res, np.int32),
    )

  def mesh(self, verts, faces):
    nv0 = sum(len(v) for v in self.lists["mesh_vert"])
    nf0 = sum(len(f) for f in self.lists["mesh_face"])
    nt0 = sum(len(t) for t in self.lists["mesh_texcoord"])
    normals, tc = _mesh_attributes(verts, faces)
    self.add(
      mesh_vertadr=nv0,
      mesh_vertnum=len(verts),
      mesh_faceadr=nf0,
      mesh_facenum=len(faces),
      mesh_normaladr=nv0,
      mesh_texcoordadr=nt0,
      mesh_vert=verts,
      mesh_normal=normals,
      mesh_face=faces,
      mesh_facenormal=faces,
      mesh_facetexcoord=np.arange(3 * len(faces), dtype=np.int32).reshape(-1, 3),
      mesh_texcoord=tc,
    )
    return len(self.lists["mesh_vertadr"]) - 1

  def texture(self, image: np.ndarray, colorspace: int, tex_type: int = mdl.TEXTURE_2D):
    h, w, nc = image.shape
    adr = sum(len(t) for t in self.lists["tex_data"])
    self.add(
      tex_type=tex_type,
      tex_width=w,
      tex_height=h,
      tex_nchannel=nc,
      tex_adr=adr,
      tex_data=image.astype(np.uint8).ravel(),
      tex_colorspace=colorspace,
    )
    return len(self.lists["tex_type"]) - 1

  def hfield(self, elevation: np.ndarray, size):
    """Height field from a (nrow, ncol) grid of normalised elevations and (x, y, z, base) size."""
    nrow, ncol = elevation.shape
    adr = sum(len(t) for t in self.lists["hfield_data"])
    self.add(
      hfield_size=np.asarray(size, np.float32),
      hfield_nrow=nrow,
      hfield_ncol=ncol,
      hfield_adr=adr,
      hfield_data=np.asarray(elevation, np.float32).ravel(),
    )
    return len(self.lists["hfield_nrow"]) - 1

  def build(self, **extra) -> mdl.HostModel:
    kw = dict(self.kw)
    concat = {
      "mesh_vert",
      "mesh_normal",
      "mesh_face",
      "mesh_facenormal",
      "mesh_facetexcoord",
      "mesh_texcoord",
      "tex_data",
      "hfield_data",
    }
    for k, items in self.lists.items():
      if not items:
        continue
      dtype = kw[k].dtype
      if k in concat:
        kw[k] = np.concatenate(items).astype(dtype)
      else:
        kw[k] = np.stack(items).astype(dtype)
    kw.update(extra)
    return mdl.HostModel(**kw)


def _place(rng, n):
  return np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(0, 1, n)], axis=1)


def _random_rgb(rng):
  return np.concatenate([rng.uniform(0.1, 0.9, 3), [1.0]])


def primitives_scene(seed: int = 0, res: int = 64) -> mdl.HostModel:
  """Ground plane + 8 primitives, two lights, no textures, one pinhole camera (fovy 45)."""
  rng = np.random.default_rng(seed)
  b = _Builder()
  plane_mat = b.material(rgba=(0.6, 0.6, 0.6, 1), metallic=0.0, roughness=float(rng.uniform(0.2, 1.0)))
  b.geom(mdl.GEOM_PLANE, (3.0, 3.0, 0.1), (0, 0, 0), matid=plane_mat)
  kinds = [mdl.GEOM_SPHERE] * 2 + [mdl.GEOM_BOX] * 2 + [mdl.GEOM_CAPSULE] * 2 + [mdl.GEOM_CYLINDER] * 2
  pos = _place(rng, len(kinds))
  for k, p in zip(kinds, pos):
    if k == mdl.GEOM_SPHERE:
      size = (rng.uniform(0.05, 0.2), 0, 0)
    elif k == mdl.GEOM_BOX:
      size = rng.uniform(0.05, 0.2, 3)
    else:
      size = (rng.uniform(0.05, 0.2), rng.uniform(0.05, 0.2), 0)
    mat = b.material(rgba=_random_rgb(rng), metallic=0.0, roughness=float(rng.uniform(0.2, 1.0)))
    b.geom(k, size, p, mat=random_rotation(rng), matid=mat)
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (0.3, 0.2, -1.0))
  b.light(mdl.LIGHT_POINT, (0.8, -0.6, 2.0), (0, 0, -1))
  b.camera((0.0, -3.0, 2.0), (0.0, 0.0, 0.3), 45.0, (res, res))
  return b.build(stat_center=np.array([0, 0, 0.5], np.float32), stat_extent=2.0)


def full_scene(seed: int = 0, res: int = 128, tex_size: int = 256) -> mdl.HostModel:
  """Mirror plane, 12 primitives, 4 convex-hull meshes, textures, sun + soft point light."""
  rng = np.random.default_rng(seed)
  b = _Builder()
  tex = [b.texture(img, cs) for img, cs in procedural_textures(seed, tex_size)]
  roles = {mdl.ROLE_RGBA: tex[0], mdl.ROLE_ORM: tex[1], mdl.ROLE_NORMAL: tex[2]}
  plane_mat = b.material(rgba=(1, 1, 1, 1), metallic=0.0, roughness=0.6, reflectance=0.3, texid=roles, repeat=(4, 4))
  b.geom(mdl.GEOM_PLANE, (3.0, 3.0, 0.1), (0, 0, 0), matid=plane_mat)
  kinds = [mdl.GEOM_SPHERE] * 3 + [mdl.GEOM_BOX] * 3 + [mdl.GEOM_CAPSULE] * 2 + [mdl.GEOM_CYLINDER] * 2 + [mdl.GEOM_ELLIPSOID] * 2
  pos = _place(rng, len(kinds) + 4)
  for i, k in enumerate(kinds):
    if k == mdl.GEOM_SPHERE:
      size = (rng.uniform(0.05, 0.2), 0, 0)
    elif k in (mdl.GEOM_BOX, mdl.GEOM_ELLIPSOID):
      size = rng.uniform(0.05, 0.2, 3)
    else:
      size = (rng.uniform(0.05, 0.2), rng.uniform(0.05, 0.2), 0)
    textured = i % 2 == 0
    mat = b.material(
      rgba=_random_rgb(rng),
      metallic=0.0,
      roughness=float(rng.uniform(0.2, 1.0)),
      texid=roles if textured else None,
      repeat=(2, 2),
    )
    b.geom(k, size, pos[i], mat=random_rotation(rng), matid=mat)
  for j, ntri in enumerate(rng.integers(500, 2001, 4)):
    verts, faces = convex_hull_mesh(rng, int(ntri), float(rng.uniform(0.12, 0.25)))
    mid = b.mesh(verts, faces)
    mat = b.material(rgba=_random_rgb(rng), metallic=0.0, roughness=float(rng.uniform(0.2, 1.0)), texid=roles if j % 2 == 0 else None)
    b.geom(mdl.GEOM_MESH, (0, 0, 0), pos[len(kinds) + j], mat=random_rotation(rng), matid=mat, dataid=mid)
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (-0.4, 0.3, -0.86), bulbradius=0.05)
  b.light(mdl.LIGHT_POINT, (0.8, -0.6, 2.0), (0, 0, -1), bulbradius=0.05)
  b.camera((0.0, -3.0, 2.0), (0.0, 0.0, 0.3), 45.0, (res, res))
  return b.build(stat_center=np.array([0, 0, 0.5], np.float32), stat_extent=2.0)


def mesh_scene(seed: int = 0, res: int = 256, nmesh: int = 16, instances: int = 4, ntri: int = 3125) -> mdl.HostModel:
  """nmesh procedural meshes of ~ntri triangles, each instanced `instances` times, textured."""
  rng = np.random.default_rng(seed)
  b = _Builder()
  tex = [b.texture(img, cs) for img, cs in procedural_textures(seed, 256)]
  roles = {mdl.ROLE_RGBA: tex[0], mdl.ROLE_ORM: tex[1], mdl.ROLE_NORMAL: tex[2]}
  plane_mat = b.material(rgba=(1, 1, 1, 1), metallic=0.0, roughness=0.7, texid=roles, repeat=(4, 4))
  b.geom(mdl.GEOM_PLANE, (3.0, 3.0, 0.1), (0, 0, 0), matid=plane_mat)
  meshes = [b.mesh(*convex_hull_mesh(rng, ntri, float(rng.uniform(0.08, 0.16)))) for _ in range(nmesh)]
  n = nmesh * instances
  side = int(math.ceil(math.sqrt(n)))
  for i in range(n):
    gx, gy = i % side, i // side
    p = ((gx + 0.5) / side * 2.2 - 1.1, (gy + 0.5) / side * 2.2 - 1.1, rng.uniform(0.1, 0.4))
    mat = b.material(rgba=_random_rgb(rng), metallic=0.0, roughness=float(rng.uniform(0.2, 1.0)), texid=roles if i % 3 == 0 else None)
    b.geom(mdl.GEOM_MESH, (0, 0, 0), p, mat=random_rotation(rng), matid=mat, dataid=meshes[i % nmesh])
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (-0.4, 0.3, -0.86), bulbradius=0.05)
  b.light(mdl.LIGHT_POINT, (0.8, -0.6, 2.0), (0, 0, -1), bulbradius=0.05)
  b.camera((0.0, -2.8, 2.2), (0.0, 0.0, 0.2), 45.0, (res, res))
  return b.build(stat_center=np.array([0, 0, 0.3], np.float32), stat_extent=2.0)


def crowd_scene(seed: int = 0, res: int = 64, ngeom: int = 600, nmat: int = 12) -> mdl.HostModel:
  """A floor and `ngeom` small primitives of every type on a jittered grid, four of them mesh
  instances, sharing `nmat` materials; a sun and a point light. Worlds with hundreds of geoms:
  records beyond the LDS and scene BVHs of more than 256 leaves."""
  rng = np.random.default_rng(seed)
  b = _Builder()
  plane_mat = b.material(rgba=(0.7, 0.7, 0.7, 1), metallic=0.0, roughness=0.8)
  b.geom(mdl.GEOM_PLANE, (3.0, 3.0, 0.1), (0, 0, 0), matid=plane_mat)
  mats = [b.material(rgba=_random_rgb(rng), metallic=float(rng.choice([0.0, 1.0])), roughness=float(rng.uniform(0.2, 1.0)))
          for _ in range(nmat)]
  mesh = b.mesh(*convex_hull_mesh(rng, 200, 0.06))
  kinds = [mdl.GEOM_SPHERE, mdl.GEOM_BOX, mdl.GEOM_CAPSULE, mdl.GEOM_CYLINDER, mdl.GEOM_ELLIPSOID]
  side = int(math.ceil(math.sqrt(ngeom)))
  pitch = 2.4 / side
  for i in range(ngeom):
    gx, gy = i % side, i // side
    p = ((gx + 0.5) * pitch - 1.2 + rng.uniform(-0.2, 0.2) * pitch, (gy + 0.5) * pitch - 1.2 + rng.uniform(-0.2, 0.2) * pitch,
         rng.uniform(0.05, 0.5))
    r = 0.3 * pitch
    mat = mats[i % nmat]
    if i % (ngeom // 4) == 0:
      b.geom(mdl.GEOM_MESH, (0, 0, 0), p, mat=random_rotation(rng), matid=mat, dataid=mesh)
      continue
    k = kinds[i % len(kinds)]
    size = (r, 0, 0) if k == mdl.GEOM_SPHERE else ((r, r * 0.8, 0) if k in (mdl.GEOM_CAPSULE, mdl.GEOM_CYLINDER) else (r, 0.7 * r, 0.5 * r))
    b.geom(k, size, p, mat=random_rotation(rng), matid=mat)
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (-0.4, 0.3, -0.86))
  b.light(mdl.LIGHT_POINT, (0.6, -0.5, 1.5), (0, 0, -1))
  b.camera((0.0, -2.6, 2.0), (0.0, 0.0, 0.2), 50.0, (res, res))
  return b.build(stat_center=np.array([0, 0, 0.3], np.float32), stat_extent=2.0)


def xyaxes_frame(xyaxes) -> np.ndarray:
  """Camera frame (3, 3) from MJCF `xyaxes` (x axis, then y axis, in world coordinates): y is
  made orthogonal to x, z = x × y; columns are the axes."""
  v = np.asarray(xyaxes, np.float64)
  x = v[:3] / np.linalg.norm(v[:3])
  y = v[3:] - x * np.dot(x, v[3:])
  y /= np.linalg.norm(y)
  return np.stack([x, y, np.cross(x, y)], axis=1)


def checker_texture(size: int, rgb1, rgb2) -> np.ndarray:
  """MuJoCo's builtin 2 x 2 checker (rgb1 on the diagonal quadrants), RGB8."""
  half = size // 2
  y, x = np.mgrid[0:size, 0:size]
  first = (x < half) == (y < half)
  img = np.where(first[..., None], np.asarray(rgb1, np.float64), np.asarray(rgb2, np.float64))
  return np.round(np.clip(img, 0.0, 1.0) * 255).astype(np.uint8)


def suite_material_grid(res: int = 128, n: int = 7) -> mdl.HostModel:
  """The reference benchmark suite's `material_grid` (benchmarks/scenes.py:46-70): n x n spheres
  of radius 0.09 on a floor, metallic rising bottom to top and roughness left to right, a
  shadow-casting sun, the headlight at ambient 0.2 / diffuse 0.3 / specular 0.3."""
  b = _Builder()
  floor = b.material(rgba=(0.5, 0.5, 0.5, 1), roughness=0.9)
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 0), (-0.4, 0.6, -0.7), diffuse=(0.7, 0.7, 0.7))
  b.geom(mdl.GEOM_PLANE, (3.0, 3.0, 0.1), (0, 0, 0), matid=floor)
  spacing = 0.22
  for i in range(n):
    for j in range(n):
      mat = b.material(rgba=(0.95, 0.64, 0.54, 1), metallic=i / (n - 1), roughness=j / (n - 1))
      b.geom(mdl.GEOM_SPHERE, (0.09, 0, 0), ((j - (n - 1) / 2) * spacing, 0.0, 0.12 + i * spacing), matid=mat)
  height = 0.12 + (n - 1) * spacing
  pos = (0.0, -0.55 - 1.9 * height, 0.5 * height + 0.06)
  b.add(
    cam_xpos=np.asarray(pos, np.float32),
    cam_xmat=xyaxes_frame((1, 0, 0, 0, 0, 1)).astype(np.float32).reshape(9),
    cam_fovy=45.0,
    cam_projection=0,
    cam_sensorsize=np.zeros(2, np.float32),
    cam_intrinsic=np.zeros(4, np.float32),
    cam_resolution=np.asarray((res, res), np.int32),
  )
  return b.build(
    stat_center=np.array([0, 0, 0.5 * height], np.float32),
    stat_extent=2.0,
    headlight_active=True,
    headlight_ambient=np.full(3, 0.2, np.float32),
    headlight_diffuse=np.full(3, 0.3, np.float32),
    headlight_specular=np.full(3, 0.3, np.float32),
  )


def suite_primitives(res: int = 128) -> mdl.HostModel:
  """The reference benchmark suite's `primitives` (benchmarks/scenes.py:73-100, MuJoCo Warp's
  render benchmark layout): an infinite checker-textured plane with reflectance 0.2, a table
  box, 25 spheres and 25 capsules, a directional light straight down and MuJoCo's default
  headlight."""
  b = _Builder()
  tex = b.texture(checker_texture(512, (0.1, 0.2, 0.3), (0.2, 0.3, 0.4)), mdl.COLORSPACE_AUTO)
  grid = b.material(rgba=(1, 1, 1, 1), reflectance=0.2, texid={mdl.ROLE_RGBA: tex}, repeat=(1, 1), uniform=1)
  b.geom(mdl.GEOM_PLANE, (0.0, 0.0, 0.05), (0, 0, 0), matid=grid)
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (0, 0, -1), diffuse=(0.7, 0.7, 0.7))
  b.geom(mdl.GEOM_BOX, (0.5, 1.0, 0.01), (0.8, 0.0, 0.5))
  for kind, size, z, rgba in ((mdl.GEOM_SPHERE, (0.05, 0, 0), 0.8, (1, 0, 0, 1)), (mdl.GEOM_CAPSULE, (0.04, 0.05, 0), 0.95, (0, 1, 0, 1))):
    for i in range(5):
      for j in range(5):
        b.geom(kind, size, (0.6 + 0.1 * i, -0.4 + 0.2 * j, z), rgba=rgba)
  b.add(
    cam_xpos=np.asarray((3.0, 0.0, 2.0), np.float32),
    cam_xmat=xyaxes_frame((0, 1, 0, -0.6, 0, 1)).astype(np.float32).reshape(9),
    cam_fovy=45.0,
    cam_projection=0,
    cam_sensorsize=np.zeros(2, np.float32),
    cam_intrinsic=np.zeros(4, np.float32),
    cam_resolution=np.asarray((res, res), np.int32),
  )
  return b.build(stat_center=np.array([0.8, 0, 0.5], np.float32), stat_extent=2.0, headlight_active=True)


def randomize_worlds(m, rc, nworld: int, seed: int = 1) -> None:
  """Per-world appearance: roughness U[0.05, 1], metallic Bernoulli(0.3), rgba U[0, 1] (alpha 1)."""
  import torch

  rng = np.random.default_rng(seed)
  ngeom = m.ngeom
  device = m.geom_rgba.device
  rgba = np.concatenate([rng.uniform(0.0, 1.0, (nworld, ngeom, 3)), np.ones((nworld, ngeom, 1))], axis=-1)
  m.geom_rgba = torch.as_tensor(rgba.astype(np.float32), device=device)
  rc.geom_roughness = torch.as_tensor(rng.uniform(0.05, 1.0, (nworld, ngeom)).astype(np.float32), device=device)
  rc.geom_metallic = torch.as_tensor((rng.uniform(0, 1, (nworld, ngeom)) < 0.3).astype(np.float32), device=device)


def jitter_cameras(d, seed: int = 1, offset: float = 0.03, angle_deg: float = 2.0) -> None:
  """Per-world random camera offsets and small rotations so worlds render different images."""
  import torch

  rng = np.random.default_rng(seed)
  pos = d.cam_xpos.cpu().numpy().astype(np.float64)
  mat = d.cam_xmat.cpu().numpy().astype(np.float64).reshape(d.nworld, -1, 3, 3)
  pos += rng.normal(0.0, offset, pos.shape)
  for w in range(d.nworld):
    for c in range(mat.shape[1]):
      axis = rng.normal(size=3)
      axis /= np.linalg.norm(axis)
      a = math.radians(angle_deg) * rng.normal()
      k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
      rot = np.eye(3) + math.sin(a) * k + (1 - math.cos(a)) * k @ k
      mat[w, c] = rot @ mat[w, c]
  d.cam_xpos = torch.as_tensor(pos.astype(np.float32), device=d.cam_xpos.device)
  d.cam_xmat = torch.as_tensor(mat.reshape(d.nworld, -1, 9).astype(np.float32), device=d.cam_xmat.device)
