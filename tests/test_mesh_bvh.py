"""Mesh BVHs: the host restatement of the median-split build (oracle/blas.py) pinned by the tree's
invariants, the GPU builder (csrc/blas.hip) compared with it bit for bit, and re-framed mesh BVHs
(mesh_bvh_frame="auto", reference io.py:138-211) rendering what the MuJoCo-frame ones render."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
import oracle
from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd import bvh
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd import scenes
from mujoco_warp_pbr_amd.scenes import _Builder
from oracle import blas
from tests.helpers import compare_rt
from tests.helpers import gpu_outputs
from tests.helpers import make


def _check_tree(depth, boxes, tri, nface):
  ids = tri[:, 3].view(np.int32)
  assert sorted(ids[ids >= 0].tolist()) == list(range(nface))
  assert (1 << depth) * blas.BLAS_LEAF == len(tri) and boxes.shape == ((1 << depth) - 1, 12)
  nleaf = 1 << depth
  for slot in np.nonzero(ids >= 0)[0]:  # every triangle inside the box its parent stores for its leaf
    node = nleaf - 1 + slot // blas.BLAS_LEAF
    parent, side = (node - 1) // 2, (node - 1) % 2
    lo, hi = boxes[parent, 6 * side : 6 * side + 3], boxes[parent, 6 * side + 3 : 6 * side + 6]
    v0 = tri[slot, 0:3]
    for p in (v0, v0 + tri[slot, 4:7], v0 + tri[slot, 8:11]):
      assert np.all(p >= lo) and np.all(p <= hi)
  for i in range(1, (1 << depth) - 1):  # parents contain their children
    parent, side = (i - 1) // 2, (i - 1) % 2
    plo, phi = boxes[parent, 6 * side : 6 * side + 3], boxes[parent, 6 * side + 3 : 6 * side + 6]
    for c in (0, 1):
      clo, chi = boxes[i, 6 * c : 6 * c + 3], boxes[i, 6 * c + 3 : 6 * c + 6]
      if clo[0] < 1e29:
        assert np.all(clo >= plo) and np.all(chi <= phi)


def _meshes():
  rng = np.random.default_rng(3)
  out = [scenes.convex_hull_mesh(rng, 300, 0.2), scenes.convex_hull_mesh(rng, 6000, 0.3)]
  # ties everywhere: a flat grid of identical-centroid-coordinate triangles
  g = np.stack(np.meshgrid(np.arange(9.0), np.arange(9.0), indexing="ij"), -1).reshape(-1, 2)
  verts = np.concatenate([g, np.zeros((81, 1))], axis=1).astype(np.float32)
  idx = np.arange(81).reshape(9, 9)
  a, b, c, d = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, 1:].ravel(), idx[1:, :-1].ravel()
  out.append((verts, np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)]).astype(np.int32)))
  out.append((np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.int32)))  # one triangle
  return out


@pytest.mark.parametrize("k", range(4))
def test_host_build_is_a_valid_median_split_tree(k):
  verts, faces = _meshes()[k]
  depth, boxes, tri = blas.build_mesh_bvh(verts, faces)
  assert depth == bvh.mesh_depth(len(faces))
  _check_tree(depth, boxes, tri, len(faces))
  # balanced: no leaf holds more than BLAS_LEAF, sibling leaf counts differ by at most the capacity rule
  counts = (tri[:, 3].view(np.int32) >= 0).reshape(-1, blas.BLAS_LEAF).sum(axis=1)
  assert counts.max() <= blas.BLAS_LEAF


def test_quat_to_mat_and_frame_selection():
  q = np.array([np.cos(0.4), np.sin(0.4) * 0.6, 0.0, np.sin(0.4) * 0.8])
  R = bvh.quat_to_mat(q)
  np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
  assert np.linalg.det(R) == pytest.approx(1.0)
  mjm = _rotated_bar_model(q)
  frames = bvh.select_mesh_frames(mjm)
  assert set(frames) == {0}
  np.testing.assert_allclose(frames[0], R)
  mjm.mesh_quat = np.array([[1.0, 0, 0, 0]])
  assert bvh.select_mesh_frames(mjm) == {}


def _bar(n=40, length=0.8, width=0.04):
  """An axis-aligned bar along x, its long faces cut into n slivers (thin, long triangles)."""
  xs = np.linspace(-length / 2, length / 2, n + 1)
  verts, faces = [], []
  for y, z in ((-width, -width), (width, -width), (width, width), (-width, width)):
    verts += [[x, y, z] for x in xs]
  verts = np.array(verts, np.float32)
  m = n + 1
  for s in range(4):
    a0, b0 = s * m, ((s + 1) % 4) * m
    for i in range(n):
      faces += [[a0 + i, b0 + i + 1, a0 + i + 1], [a0 + i, b0 + i, b0 + i + 1]]  # outward
  c = len(verts)
  verts = np.concatenate([verts, [[-length / 2, 0, 0], [length / 2, 0, 0]]]).astype(np.float32)
  for s in range(4):  # end caps as fans
    a, b = s * m, ((s + 1) % 4) * m
    faces += [[c, b, a], [c + 1, a + n, b + n]]
  return verts, np.array(faces, np.int32)


def _rotated_bar_model(q):
  """Ground plane and a bar mesh stored in a rotated frame: mesh_quat maps it back to the
  axis-aligned bar (what MuJoCo does to CAD meshes: it stores them in their inertia frame)."""
  R = bvh.quat_to_mat(q)
  verts, faces = _bar()
  stored = (verts.astype(np.float64) @ R).astype(np.float32)  # R^T v: the stored (rotated) vertices
  b = _Builder()
  b.geom(mdl.GEOM_PLANE, (2, 2, 0.1), (0, 0, 0), matid=b.material(rgba=(0.6, 0.6, 0.6, 1), roughness=0.8))
  mid = b.mesh(stored, faces)
  rot = scenes.random_rotation(np.random.default_rng(5))
  b.geom(mdl.GEOM_MESH, (0, 0, 0), (0.0, 0.0, 0.3), mat=rot, matid=b.material(rgba=(0.8, 0.4, 0.2, 1), roughness=0.5), dataid=mid)
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (0.3, 0.2, -1.0))
  b.camera((0.0, -1.4, 1.0), (0, 0, 0.3), 45.0, (32, 32))
  mjm = b.build()
  mjm.mesh_quat = np.array([q])
  return mjm


Q = np.array([np.cos(0.6), 0.3 * np.sin(0.6), 0.5 * np.sin(0.6), np.sqrt(1 - 0.34) * np.sin(0.6)])


def test_reframed_mesh_renders_like_the_mujoco_frame_on_the_oracle():
  mjm = _rotated_bar_model(Q)
  outs = {}
  for frame in ("auto", "mujoco"):
    m, d, rc = make(mjm, nworld=2, cam_res=(32, 32), mesh_bvh_frame=frame, render_guides=True)
    assert (rc.base.mesh_frame is not None) == (frame == "auto")
    outs[frame] = oracle.rt_render(m, d, rc)
  a, b = outs["auto"], outs["mujoco"]
  np.testing.assert_array_equal(a["seg"], b["seg"])
  assert (a["seg"][..., 0] == 1).sum() > 20
  np.testing.assert_allclose(a["depth"], b["depth"], rtol=2e-5, atol=1e-6)
  np.testing.assert_allclose(a["normal"], b["normal"], atol=2e-4)
  np.testing.assert_allclose(a["hdr"], b["hdr"], atol=2e-3)


# ----------------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("k", range(4))
def test_gpu_mesh_bvh_matches_host_build_bit_for_bit(cuda, k):
  verts, faces = _meshes()[k]
  depth = bvh.mesh_depth(len(faces))
  want_depth, want_box, want_tri = blas.build_mesh_bvh(verts, faces)
  assert depth == want_depth
  box = torch.zeros(((1 << depth) - 1, 12), device=cuda)
  tri = torch.zeros(((1 << depth) * blas.BLAS_LEAF, 12), device=cuda)
  need = _abi.library().mjpbr_mesh_bvh_workspace(len(faces), depth)
  work = torch.empty(int(need), dtype=torch.uint8, device=cuda)
  v = torch.as_tensor(verts).to(cuda)
  f = torch.as_tensor(faces).to(cuda)
  _abi.call("mjpbr_mesh_bvh_build", _abi.ptr(v), _abi.ptr(f), len(faces), depth, _abi.ptr(box), _abi.ptr(tri), _abi.ptr(work),
            work.numel(), 0)
  torch.cuda.synchronize()
  np.testing.assert_array_equal(tri.cpu().numpy().view(np.uint32), want_tri.view(np.uint32))
  np.testing.assert_array_equal(box.cpu().numpy().view(np.uint32), want_box.view(np.uint32))


@pytest.mark.gpu
def test_gpu_context_mesh_bvhs_match_host_build(cuda):
  """Every mesh and height field of a context, built on the GPU at context creation."""
  from tests.test_hfield import _terrain_scene

  for mjm in (scenes.full_scene(0, res=16), _terrain_scene()):
    _, _, rc = make(mjm, nworld=1, device=cuda, cam_res=(16, 16))
    torch.cuda.synchronize()
    mb = rc.base.mesh_bvhs
    for m, (v0, nv, faces) in enumerate(mb.faces):
      d, b, t = blas.build_mesh_bvh(mb.host_vert[v0 : v0 + nv], faces)
      nb, nt = int(mb.node_adr[m]), int(mb.tri_adr[m])
      np.testing.assert_array_equal(rc.base.blas_box[nb : nb + len(b)].cpu().numpy(), b)
      np.testing.assert_array_equal(rc.base.blas_tri[nt : nt + len(t)].cpu().numpy().view(np.uint32), t.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("tier", [mj.Tier.REALTIME, mj.Tier.PATH])
def test_gpu_reframed_mesh_matches_oracle(cuda, tier):
  mjm = _rotated_bar_model(Q)
  kw = {"tier": tier, "path_samples": 2, "max_depth": 4} if tier == mj.Tier.PATH else {"samples_per_pixel": 2}
  m, d, rc = make(mjm, nworld=2, device=cuda, cam_res=(32, 32), mesh_bvh_frame="auto", **kw)
  assert rc.base.mesh_frame is not None
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc)
  gpu = gpu_outputs(rc)
  ref = oracle.pt_render(m, d, rc) if tier == mj.Tier.PATH else oracle.rt_render(m, d, rc)
  if tier == mj.Tier.PATH:
    np.testing.assert_array_equal(gpu["seg"], ref["seg"])
    diff = np.abs(gpu["hdr"] - ref["hdr"]).max(axis=-1)
    assert np.mean(diff <= 1e-3) >= 0.999
  else:
    compare_rt(gpu, ref)
  assert (gpu["seg"][..., 0] == 1).sum() > 20
