"""Acceleration structures: mesh BVHs (host build) and the per-world scene BVH (GPU build/refit).

Both are implicit perfect binary trees in heap order whose inner nodes store the boxes of their
two children (12 floats); see include/mjpbr.h. Mesh BVH leaves hold MJPBR_BLAS_LEAF = 4
triangle slots `{v0, face id}, {e1 = v1 - v0}, {e2 = v2 - v0}` (12 floats, mesh frame); padding
slots have face id -1. Triangles are assigned to leaves by recursive median split of their
centroids along the widest axis, which keeps the tree balanced (depth = ceil(log2(leaves))).
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd import model as mdl

BLAS_LEAF = 4
_EMPTY = 1.0e30


def _depth_for(nleaf: int) -> int:
  d = 1
  while (1 << d) < nleaf:
    d += 1
  return d


def _inflate(lo: np.ndarray, hi: np.ndarray):
  ext = np.max(hi - lo, axis=-1, keepdims=True)
  pad = 1.0e-5 * ext + 1.0e-6 * (np.abs(lo) + np.abs(hi)) + 1.0e-6
  return lo - pad, hi + pad


def _assign(centroids: np.ndarray, depth: int, capacity: int) -> np.ndarray:
  """Leaf slot (0 .. 2^depth * capacity) of every item by recursive median split."""
  n = len(centroids)
  slots = np.empty(n, np.int64)

  def rec(idx: np.ndarray, first_leaf: int, levels: int):
    if levels == 0:
      slots[idx] = first_leaf * capacity + np.arange(len(idx))
      return
    if len(idx) == 0:
      return
    c = centroids[idx]
    axis = int(np.argmax(c.max(axis=0) - c.min(axis=0)))
    order = idx[np.argsort(c[:, axis], kind="stable")]
    half_cap = (1 << (levels - 1)) * capacity
    nl = min((len(idx) + 1) // 2, half_cap)
    nl = max(nl, len(idx) - half_cap)
    rec(order[:nl], first_leaf, levels - 1)
    rec(order[nl:], first_leaf + (1 << (levels - 1)), levels - 1)

  rec(np.arange(n), 0, depth)
  return slots


def _inner_boxes(leaf_lo: np.ndarray, leaf_hi: np.ndarray, depth: int) -> np.ndarray:
  """(2^depth - 1, 12) child boxes of every inner node from the (2^depth, 3) leaf boxes."""
  ninner = (1 << depth) - 1
  boxes = np.zeros((ninner, 12), np.float32)
  lo, hi = leaf_lo, leaf_hi
  for level in range(depth - 1, -1, -1):
    first = (1 << level) - 1
    count = 1 << level
    l0, h0 = lo[0::2], hi[0::2]
    l1, h1 = lo[1::2], hi[1::2]
    boxes[first : first + count] = np.concatenate([l0, h0, l1, h1], axis=1)
    lo = np.minimum(l0, l1)
    hi = np.maximum(h0, h1)
  return boxes


@dataclasses.dataclass
class MeshBvhs:
  depth: np.ndarray  # (nmesh,)
  node_adr: np.ndarray  # (nmesh,)
  tri_adr: np.ndarray  # (nmesh,)
  box: np.ndarray  # (ninner total, 12) f32
  tri: np.ndarray  # (slots, 12) f32
  aabb: np.ndarray  # (nmesh, 6)


def build_mesh_bvh(verts: np.ndarray, faces: np.ndarray):
  """(depth, inner boxes (n, 12), triangle slots (m, 12)) of one mesh."""
  verts = np.asarray(verts, np.float32)
  faces = np.asarray(faces, np.int64)
  nf = len(faces)
  nleaf = max(2, -(-nf // BLAS_LEAF))
  depth = _depth_for(nleaf)
  nleaf = 1 << depth
  v0 = verts[faces[:, 0]]
  v1 = verts[faces[:, 1]]
  v2 = verts[faces[:, 2]]
  cen = (v0.astype(np.float64) + v1 + v2) / 3.0
  slots = _assign(cen, depth, BLAS_LEAF)
  tri = np.zeros((nleaf * BLAS_LEAF, 12), np.float32)
  tri[:, 3] = np.array([-1], np.int32).view(np.float32)[0]
  tri[slots, 0:3] = v0
  tri[slots, 3] = np.arange(nf, dtype=np.int32).view(np.float32)
  tri[slots, 4:7] = v1 - v0
  tri[slots, 8:11] = v2 - v0
  tlo = np.minimum(np.minimum(v0, v1), v2).astype(np.float64)
  thi = np.maximum(np.maximum(v0, v1), v2).astype(np.float64)
  leaf_lo = np.full((nleaf, BLAS_LEAF, 3), np.inf)
  leaf_hi = np.full((nleaf, BLAS_LEAF, 3), -np.inf)
  leaf_lo.reshape(-1, 3)[slots] = tlo
  leaf_hi.reshape(-1, 3)[slots] = thi
  llo = leaf_lo.min(axis=1)
  lhi = leaf_hi.max(axis=1)
  empty = ~np.isfinite(llo[:, 0])
  llo, lhi = _inflate(np.where(empty[:, None], 0.0, llo), np.where(empty[:, None], 0.0, lhi))
  llo[empty] = _EMPTY
  lhi[empty] = _EMPTY
  # propagate with empty leaves not widening their parents
  boxes = _inner_boxes_masked(llo, lhi, empty, depth)
  return depth, boxes, tri


def _inner_boxes_masked(llo, lhi, empty, depth):
  ninner = (1 << depth) - 1
  boxes = np.zeros((ninner, 12), np.float32)
  lo, hi, emp = llo.copy(), lhi.copy(), empty.copy()
  for level in range(depth - 1, -1, -1):
    first = (1 << level) - 1
    count = 1 << level
    l0, h0, e0 = lo[0::2], hi[0::2], emp[0::2]
    l1, h1, e1 = lo[1::2], hi[1::2], emp[1::2]
    boxes[first : first + count] = np.concatenate([l0, h0, l1, h1], axis=1)
    lo = np.where(e0[:, None], l1, np.where(e1[:, None], l0, np.minimum(l0, l1)))
    hi = np.where(e0[:, None], h1, np.where(e1[:, None], h0, np.maximum(h0, h1)))
    emp = e0 & e1
  return boxes


def build_mesh_bvhs(mjm: mdl.HostModel) -> MeshBvhs:
  """BVHs of the model's meshes followed by those of its height fields (mesh nmesh + h)."""
  tm = mdl.trace_mesh_arrays(mjm)
  depth, node_adr, tri_adr, boxes, tris, aabb = [], [], [], [], [], []
  nb = nt = 0
  for m in range(len(tm["mesh_vertadr"])):
    v0, nv = int(tm["mesh_vertadr"][m]), int(tm["mesh_vertnum"][m])
    f0, nf = int(tm["mesh_faceadr"][m]), int(tm["mesh_facenum"][m])
    verts = tm["mesh_vert"][v0 : v0 + nv]
    d, b, t = build_mesh_bvh(verts, tm["mesh_face"][f0 : f0 + nf])
    depth.append(d)
    node_adr.append(nb)
    tri_adr.append(nt)
    boxes.append(b)
    tris.append(t)
    nb += len(b)
    nt += len(t)
    aabb.append(np.concatenate([verts.min(axis=0), verts.max(axis=0)]))
  if not depth:
    return MeshBvhs(
      np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros((1, 12), np.float32), np.zeros((4, 12), np.float32), np.zeros((1, 6), np.float32)
    )
  return MeshBvhs(
    np.array(depth, np.int32),
    np.array(node_adr, np.int32),
    np.array(tri_adr, np.int32),
    np.concatenate(boxes).astype(np.float32),
    np.concatenate(tris).astype(np.float32),
    np.array(aabb, np.float32),
  )


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
