"""Host restatement of the mesh BVH build (TEST INFRASTRUCTURE ONLY, like the rest of oracle/):
the tree the GPU builder makes (mujoco_warp_pbr_amd/csrc/blas.hip), built with NumPy by recursion.
It checks the GPU build bit for bit and fills the mesh BVHs of contexts created on the host
(precompute=False)."""

from __future__ import annotations

import numpy as np

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




def fill_mesh_bvhs(base, verts: np.ndarray, faces_by_mesh: list) -> None:
  """Writes every mesh's BVH into a host context's blas_box / blas_tri at its addresses.
  faces_by_mesh[m] = (vertex offset, vertex count, face array)."""
  box = base.blas_box.numpy()
  tri = base.blas_tri.numpy()
  for m, (v0, nv, faces) in enumerate(faces_by_mesh):
    d, b, t = build_mesh_bvh(verts[v0 : v0 + nv], faces)
    nb, nt = int(base.blas_node_adr[m]), int(base.blas_tri_adr[m])
    box[nb : nb + len(b)] = b
    tri[nt : nt + len(t)] = t
