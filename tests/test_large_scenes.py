"""Worlds with hundreds of geoms: scene BVHs past 256 leaves (one workgroup a world builds and
refits them) and per-world records beyond the 64 KB the render kernels stage in LDS (read in place
from the context's world_stage instead).

The reference traverses any BVH MuJoCo Warp builds (reference traverse.py:194-323); segmentation
against the oracle's brute-force closest hit pins the traversal, the tree invariants pin the
builder's output.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
import oracle
from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd import io
from mujoco_warp_pbr_amd import scenes
from tests.helpers import compare_rt
from tests.helpers import gpu_outputs
from tests.helpers import make

EMPTY = np.float32(1.0e30)
PATH = {"tier": mj.Tier.PATH, "render_hdr": True}


def test_crowd_scene_records_exceed_the_lds():
  mjm = scenes.crowd_scene(0, 16, 600)
  m, d, rc = make(mjm, nworld=1, cam_res=(16, 16))
  assert rc.base.tlas_depth == 10
  ndirect = int(rc.base.direct_geom_ids.numel())
  assert ndirect == 1 and int(rc.base.bvh_geom_ids.numel()) == 600  # the floor is tested before the BVH
  floats = io.world_stage_floats(rc.base.tlas_depth, mjm.ngeom, mjm.nlight, ndirect)
  assert floats * 4 > 64 * 1024
  assert tuple(rc.base.world_stage.shape) == (1, floats)


def test_oversized_world_without_world_stage_is_refused_without_a_gpu():
  """The launcher checks on the host: records beyond the LDS need geo.world_stage."""
  lib = _abi.library()
  S = _abi.structs()
  sc = S["MjpbrScene"]()
  s = S["MjpbrSettings"]()
  out = S["MjpbrOutputs"]()
  sc.cams.nrender = 1
  sc.cams.cam_res[0][0] = sc.cams.cam_res[0][1] = 16
  sc.cams.total_rays = 256
  sc.geo.nworld = 1
  sc.geo.tlas_depth = 10
  sc.geo.ngeom = 601
  sc.lit.nlight = 2
  s.samples_per_axis = 1
  s.path_samples = 1
  for fn in ("mjpbr_rt_render", "mjpbr_pt_render"):
    rc = getattr(lib, fn)(ctypes.addressof(sc), ctypes.addressof(s), ctypes.addressof(out), 0, None)
    assert rc == -1
    assert b"pass geo.world_stage" in lib.mjpbr_last_error(), fn
  # the wavefront batch refuses such a group in prepare, before it touches the device
  plan = S["MjpbrWavefrontPlan"]()
  scenes_arr = (S["MjpbrScene"] * 1)(sc)
  outs_arr = (S["MjpbrOutputs"] * 1)(out)
  need = lib.mjpbr_pt_wavefront_workspace(ctypes.addressof(scenes_arr), 1, ctypes.addressof(s), 16)
  assert need > 0
  rc = lib.mjpbr_pt_wavefront_prepare(ctypes.addressof(scenes_arr), ctypes.addressof(outs_arr), 1, ctypes.addressof(s), ctypes.c_void_p(256),
                                      need, ctypes.addressof(plan), None)
  assert rc == -1
  assert b"pass geo.world_stage" in lib.mjpbr_last_error()


def test_scene_bvh_limit_is_4096_geoms():
  lib = _abi.library()
  S = _abi.structs()
  geo = S["MjpbrGeometry"]()
  geo.nbvh = 4097
  geo.tlas_depth = 12
  assert lib.mjpbr_scene_bvh(ctypes.addressof(geo), 0, None) == -1
  assert b"1..4096" in lib.mjpbr_last_error()
  geo.nbvh = 600
  geo.tlas_depth = 9
  assert lib.mjpbr_scene_bvh(ctypes.addressof(geo), 0, None) == -1
  assert b"too small" in lib.mjpbr_last_error()


# ------------------------------------------------------------------------------------------- GPU
def _check_tree(box: np.ndarray, leaf: np.ndarray, bvh_ids: np.ndarray) -> None:
  """One world's scene BVH: every BVH geom in exactly one leaf; each inner node's child box is
  the union of that child's non-empty children (leaf boxes as stored); empty subtrees are the far
  point and widen nothing."""
  depth = int(np.log2(leaf.size))
  ninner = leaf.size - 1
  placed = leaf[leaf >= 0]
  assert sorted(placed.tolist()) == sorted(bvh_ids.tolist())
  # leaves fill left-first: a node's left subtree never holds fewer geoms than its right one
  counts = (leaf >= 0).astype(int)
  for level in range(depth - 1, -1, -1):
    pairs = counts.reshape(-1, 2)
    assert np.all(pairs[:, 0] >= pairs[:, 1])
    counts = pairs.sum(axis=1)
  for i in range(ninner):
    for side in range(2):
      c = 2 * i + 1 + side
      stored = box[i, 6 * side : 6 * side + 6]
      if c >= ninner:
        gi = leaf[c - ninner]
        if gi < 0:
          assert np.all(stored == EMPTY)
        else:
          assert np.all(stored[:3] <= stored[3:]) and np.all(np.abs(stored) < 1e9)
        continue
      kids = box[c].reshape(2, 6)
      full = [k for k in kids if k[0] != EMPTY]
      if not full:
        assert np.all(stored == EMPTY)
        continue
      want = np.concatenate([np.min([k[:3] for k in full], axis=0), np.max([k[3:] for k in full], axis=0)])
      np.testing.assert_array_equal(stored, want, err_msg=f"node {i} side {side}")


@pytest.mark.gpu
@pytest.mark.parametrize("ngeom", [40, 150, 600, 3000])
def test_gpu_scene_bvh_build_and_refit_invariants(cuda, ngeom):
  """Build (median split, block-sorted segments above 64 geoms, thread-sorted below) and refit
  after the geoms move: the tree invariants hold and refit keeps the leaf assignment."""
  mjm = scenes.crowd_scene(1, 16, ngeom)
  m, d, rc = make(mjm, nworld=3, device="cuda", cam_res=(16, 16))
  mj.refit_bvh(m, d, rc)
  torch.cuda.synchronize()
  nleaf = 1 << rc.base.tlas_depth
  bvh_ids = rc.base.bvh_geom_ids.cpu().numpy()
  assert bvh_ids.size == ngeom
  box = rc.base.tlas_box.cpu().numpy().reshape(3, nleaf - 1, 12)
  leaf = rc.base.tlas_leaf.cpu().numpy().reshape(3, nleaf)
  for w in range(3):
    _check_tree(box[w], leaf[w], bvh_ids)
  # move every geom, refit: same leaves, boxes follow
  d.geom_xpos += torch.tensor([0.3, -0.2, 0.1], device=cuda)
  mj.refit_bvh(m, d, rc)
  torch.cuda.synchronize()
  box2 = rc.base.tlas_box.cpu().numpy().reshape(3, nleaf - 1, 12)
  np.testing.assert_array_equal(rc.base.tlas_leaf.cpu().numpy().reshape(3, nleaf), leaf)
  for w in range(3):
    _check_tree(box2[w], leaf[w], bvh_ids)
  shift = np.array([0.3, -0.2, 0.1], np.float32)
  np.testing.assert_allclose(box2[0, 0, :3], box[0, 0, :3] + shift, atol=1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("ngeom", [150, 600])
def test_gpu_crowd_realtime_matches_oracle(cuda, ngeom):
  """150 geoms: records staged in LDS, 256-leaf BVH; 600: records read from world_stage (the
  global-records kernel instance), 1024-leaf BVH. Supersampling and soft shadows on."""
  mjm = scenes.crowd_scene(2, 48, ngeom)
  m, d, rc = make(mjm, nworld=2, device="cuda", cam_res=(48, 48), envmap=scenes.procedural_sky(1, 64, 128),
                  samples_per_pixel=2, shadow_samples=2)
  scenes.randomize_worlds(m, rc, 2, seed=3)
  rc.light_bulbradius = torch.full_like(rc.light_bulbradius, 0.05)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=1)
  gpu = gpu_outputs(rc)
  ref = oracle.rt_render(m, d, rc, seed=1, threads=8)
  compare_rt(gpu, ref)
  assert len(np.unique(gpu["seg"][..., 0])) > ngeom // 8


@pytest.mark.gpu
def test_gpu_crowd_path_traced_matches_oracle(cuda):
  """The path-traced kernel's global-records instance: one path a pixel follows the oracle."""
  mjm = scenes.crowd_scene(3, 32, 600)
  m, d, rc = make(mjm, nworld=2, device="cuda", cam_res=(32, 32), envmap=scenes.procedural_sky(2, 64, 128), path_samples=1,
                  max_depth=6, **PATH)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=4)
  gpu = gpu_outputs(rc)
  ref = oracle.pt_render(m, d, rc, seed=4, threads=8)
  np.testing.assert_array_equal(gpu["seg"], ref["seg"])
  np.testing.assert_allclose(gpu["depth"], ref["depth"], rtol=1e-5, atol=1e-7)
  diff = np.abs(gpu["hdr"] - ref["hdr"]).max(axis=-1)
  assert np.mean(diff <= 1e-3) >= 0.999, f"{np.mean(diff > 1e-3):.4%} of samples differ, max {diff.max()}"


@pytest.mark.gpu
def test_gpu_wavefront_batch_with_oversized_worlds_matches_megakernel(cuda):
  """A wavefront batch whose first group's records exceed the LDS (600 geoms: staged into its
  world_stage by the launch and read from there) beside a group staged in LDS: both groups'
  images equal the per-pixel megakernel's bit for bit, also on a second launch after the geoms
  moved (the staging runs every launch)."""
  from mujoco_warp_pbr_amd.wavefront import WavefrontBatch

  big = scenes.crowd_scene(4, 32, 600)
  small = scenes.crowd_scene(5, 16, 40)
  kw = dict(envmap=scenes.procedural_sky(2, 64, 128), path_samples=2, max_depth=4, **PATH)
  groups = [make(big, nworld=2, device="cuda", cam_res=(32, 32), **kw), make(small, nworld=3, device="cuda", cam_res=(16, 16), **kw)]
  assert io.world_stage_floats(groups[0][2].base.tlas_depth, big.ngeom, big.nlight, int(groups[0][2].base.direct_geom_ids.numel())) * 4 > 64 * 1024
  batch = WavefrontBatch(groups)
  for step in range(2):
    if step:
      for m, d, rc in groups:
        d.geom_xpos += torch.tensor([0.05, -0.03, 0.0], device=cuda)
    ref = []
    for m, d, rc in groups:
      mj.refit_bvh(m, d, rc)
      mj.render(m, d, rc, seed=9)
      ref.append(gpu_outputs(rc))
      for t in (rc.base.rgb_data, rc.hdr_data, rc.base.depth_data, rc.base.seg_data):
        t.fill_(7)
    batch.render(seed=9)
    for i, (m, d, rc) in enumerate(groups):
      got = gpu_outputs(rc)
      for k in ref[i]:
        np.testing.assert_array_equal(got[k], ref[i][k], err_msg=f"launch {step} group {i} {k}")
