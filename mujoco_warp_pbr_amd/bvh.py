"""Acceleration structures: mesh BVHs and the per-world scene BVH, both built on the GPU.

Both are implicit perfect binary trees in heap order whose inner nodes store the boxes of their
two children (12 floats); see include/mjpbr.h. Mesh BVH leaves hold MJPBR_BLAS_LEAF = 4 triangle
slots `{v0, face id}, {e1 = v1 - v0}, {e2 = v2 - v0}` (12 floats, traced frame); padding slots
have face id -1. Triangles are assigned to leaves by recursive median split of their centroids
along the widest axis (csrc/blas.hip), which keeps the tree balanced (depth = ceil(log2(leaves))).
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd import model as mdl

BLAS_LEAF = 4
# A mesh is traced in its authored frame when that shrinks the summed area of its triangles'
# bounding boxes by at least this factor (reference io.py:62, :138-160).
MESH_FRAME_MIN_GAIN = 1.25


def _depth_for(nleaf: int) -> int:
  d = 1
  while (1 << d) < nleaf:
    d += 1
  return d


def mesh_depth(nface: int) -> int:
  """Depth of a mesh BVH: 2^depth leaves of BLAS_LEAF triangles, at least two leaves."""
  return _depth_for(max(2, -(-nface // BLAS_LEAF)))


def quat_to_mat(q) -> np.ndarray:
  """Rotation matrix (3, 3) of a unit quaternion (w, x, y, z)."""
  w, x, y, z = (float(v) for v in np.asarray(q, np.float64) / np.linalg.norm(q))
  return np.array(
    [
      [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ]
  )


def _box_area_sum(points: np.ndarray, faces: np.ndarray) -> float:
  """Summed surface area of the triangles' axis-aligned boxes (the cost a BVH pays for them)."""
  tri = points[faces]
  e = tri.max(axis=1) - tri.min(axis=1)
  return float(np.sum(e[:, 0] * e[:, 1] + e[:, 1] * e[:, 2] + e[:, 2] * e[:, 0]))


def select_mesh_frames(mjm: mdl.HostModel) -> dict[int, np.ndarray]:
  """{mesh id: R} of the meshes whose BVH is better built in their authored frame
  (mesh_bvh_frame="auto", reference io.py:138-160). MuJoCo stores mesh vertices in the mesh's
  principal inertia frame; R = rotation of mesh_quat takes them back to the authored frame (for
  CAD-style assets axis-aligned), where long triangles get tight boxes. Kept when the summed
  triangle-box area shrinks by MESH_FRAME_MIN_GAIN."""
  quats = getattr(mjm, "mesh_quat", None)
  if quats is None:
    return {}
  out = {}
  for m in range(mjm.nmesh):
    q = np.asarray(quats[m], np.float64)
    if np.allclose(q, (1.0, 0.0, 0.0, 0.0)):
      continue
    v0, nv = int(mjm.mesh_vertadr[m]), int(mjm.mesh_vertnum[m])
    f0, nf = int(mjm.mesh_faceadr[m]), int(mjm.mesh_facenum[m])
    points = np.asarray(mjm.mesh_vert[v0 : v0 + nv], np.float64)
    faces = np.asarray(mjm.mesh_face[f0 : f0 + nf])
    R = quat_to_mat(q)
    if _box_area_sum(points, faces) >= MESH_FRAME_MIN_GAIN * _box_area_sum(points @ R.T, faces):
      out[m] = R
  return out


@dataclasses.dataclass
class MeshBvhs:
  depth: np.ndarray  # (nmesh,)
  node_adr: np.ndarray  # (nmesh,)
  tri_adr: np.ndarray  # (nmesh,)
  box: torch.Tensor  # (ninner total, 12) f32
  tri: torch.Tensor  # (slots, 12) f32
  aabb: np.ndarray  # (nmesh, 6)
  frame: torch.Tensor | None  # (nmesh, 9) rotations of re-framed meshes (identity otherwise), or None
  vert: torch.Tensor | None  # traced vertices (re-framed meshes rotated), or None: the model's
  normal: torch.Tensor | None  # traced vertex normals, likewise
  host_vert: np.ndarray  # traced vertices on the host
  faces: list  # (vertex offset, vertex count, faces) of every traced mesh


def build_mesh_bvhs(mjm: mdl.HostModel, device, precompute: bool = True, frames: dict[int, np.ndarray] | None = None) -> MeshBvhs:
  """BVHs of the model's meshes followed by those of its height fields (mesh nmesh + h), built on
  the GPU (precompute) or left zero for the caller to fill. Meshes in `frames` are traced in the
  frame R (vertices R v, normals R n): their geoms' rotation becomes geom_xmat R^T, so distances,
  faces, barycentrics and world normals are unchanged (reference trace.py:20-30)."""
  frames = frames or {}
  tm = mdl.trace_mesh_arrays(mjm)
  verts = np.asarray(tm["mesh_vert"], np.float32)
  normals = np.asarray(tm["mesh_normal"], np.float32)
  nmesh = len(tm["mesh_vertadr"])
  if frames:
    verts = verts.copy()
    normals = normals.copy()
    for m, R in frames.items():
      v0, nv = int(tm["mesh_vertadr"][m]), int(tm["mesh_vertnum"][m])
      verts[v0 : v0 + nv] = (verts[v0 : v0 + nv].astype(np.float64) @ R.T).astype(np.float32)
      n0 = int(tm["mesh_normaladr"][m])
      normals[n0 : n0 + nv] = (normals[n0 : n0 + nv].astype(np.float64) @ R.T).astype(np.float32)
  dev = torch.device(device)
  if nmesh == 0:
    return MeshBvhs(np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1, np.int32), torch.zeros((1, 12), device=dev),
                    torch.zeros((4, 12), device=dev), np.zeros((1, 6), np.float32), None, None, None, verts, [])
  depth, node_adr, tri_adr, aabb, faces = [], [], [], [], []
  nb = nt = 0
  for m in range(nmesh):
    v0, nv = int(tm["mesh_vertadr"][m]), int(tm["mesh_vertnum"][m])
    f0, nf = int(tm["mesh_faceadr"][m]), int(tm["mesh_facenum"][m])
    d = mesh_depth(nf)
    depth.append(d)
    node_adr.append(nb)
    tri_adr.append(nt)
    nb += (1 << d) - 1
    nt += BLAS_LEAF << d
    pts = verts[v0 : v0 + nv]
    aabb.append(np.concatenate([pts.min(axis=0), pts.max(axis=0)]))
    faces.append((v0, nv, np.ascontiguousarray(tm["mesh_face"][f0 : f0 + nf], np.int32)))
  box = torch.zeros((nb, 12), dtype=torch.float32, device=dev)
  tri = torch.zeros((nt, 12), dtype=torch.float32, device=dev)
  frame = None
  if frames:
    f = np.tile(np.eye(3, dtype=np.float32).reshape(1, 9), (nmesh, 1))
    for m, R in frames.items():
      f[m] = R.astype(np.float32).reshape(9)
    frame = torch.as_tensor(f).to(dev)
  bvhs = MeshBvhs(np.array(depth, np.int32), np.array(node_adr, np.int32), np.array(tri_adr, np.int32), box, tri,
                  np.array(aabb, np.float32), frame, torch.as_tensor(verts).to(dev) if frames else None,
                  torch.as_tensor(normals).to(dev) if frames else None, verts, faces)
  if precompute:
    _build_on_device(bvhs)
  return bvhs


def _build_on_device(b: MeshBvhs) -> None:
  dev = b.box.device
  stream = int(torch.cuda.current_stream(dev).cuda_stream)
  lib = _abi.library()
  need = max(lib.mjpbr_mesh_bvh_workspace(len(f), int(d)) for (_, _, f), d in zip(b.faces, b.depth))
  if need < 0:
    raise _abi.MjpbrError(lib.mjpbr_last_error().decode())
  work = torch.empty(int(need), dtype=torch.uint8, device=dev)
  verts = torch.as_tensor(b.host_vert).to(dev)
  for m, (v0, nv, faces) in enumerate(b.faces):
    f = torch.as_tensor(faces).to(dev)
    _abi.call(
      "mjpbr_mesh_bvh_build",
      verts.data_ptr() + 12 * v0,
      _abi.ptr(f),
      len(faces),
      int(b.depth[m]),
      b.box.data_ptr() + 48 * int(b.node_adr[m]),
      b.tri.data_ptr() + 48 * int(b.tri_adr[m]),
      _abi.ptr(work),
      work.numel(),
      stream,
    )
  # the per-mesh face tensors and the shared workspace are reused by the stream in order; keep
  # the host-side uploads alive until the builds have run
  torch.cuda.current_stream(dev).synchronize()


def scene_depth(nbvh: int) -> int:
  return _depth_for(max(nbvh, 2))


def alloc_scene_bvh(nworld: int, depth: int, device):
  nleaf = 1 << depth
  box = torch.zeros((nworld, nleaf - 1, 12), dtype=torch.float32, device=device)
  leaf = torch.full((nworld, nleaf), -1, dtype=torch.int32, device=device)
  return box, leaf


def geometry_struct(m, d, rc):
  """The MjpbrGeometry of a model, data and context (device pointers)."""
  from mujoco_warp_pbr_amd.render import geometry_struct as _g

  return _g(m, d, rc)


def refit_bvh(m, d, rc, rebuild: bool = False) -> None:
  """Updates the per-world scene BVH from the current geom poses (MuJoCo Warp's refit_bvh).

  rebuild=True re-assigns geoms to leaves (use after large motions); the default keeps the
  assignment and recomputes the boxes.
  """
  from mujoco_warp_pbr_amd.render import geometry_struct as _g
  from mujoco_warp_pbr_amd.render import _stream_of

  geo = _g(m, d, rc)
  # The first call builds (the context holds no leaf assignment yet).
  built = rc._host_cache.get("bvh_built", False)
  _abi.call("mjpbr_scene_bvh", _abi.byref(geo), 0 if (rebuild or not built) else 1, _stream_of(d.geom_xpos))
  rc._host_cache["bvh_built"] = True
