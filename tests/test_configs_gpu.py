"""The benchmark workloads' configurations on the GPU against the CPU oracle, at reduced world
counts (the oracle renders every compared pixel on the host):

- the mixed-resolution wavefront batch (full-feature scene, 4 paths a pixel, depth 8, Russian
  roulette from depth 3, one world group each at 64^2, 128^2 and 256^2);
- the mesh-heavy scene (64 instances of 16 procedural meshes, ~200k triangles, textures, HDR sky)
  in the real-time and the path-traced tier;
- a spot light and the ACES tone mapper in the real-time tier.

Bars: segmentation bit-exact; depth 1e-5 relative; real-time HDR 1e-3 absolute on every pixel
and sRGB8 within 1; path-traced HDR within 1e-3 absolute on at least 99.9 % of pixels (a path
whose float sequence diverges in one random decision takes a different route).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
import oracle
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd import scenes
from mujoco_warp_pbr_amd.wavefront import WavefrontBatch
from tests.helpers import compare_rt
from tests.helpers import gpu_outputs
from tests.helpers import make

pytestmark = pytest.mark.gpu

PT_BATCH = {"tier": mj.Tier.PATH, "path_samples": 4, "max_depth": 8, "rr_depth": 3, "render_hdr": True}


def _randomized(m, d, rc, nworld, seed):
  scenes.randomize_worlds(m, rc, nworld, seed=seed)
  scenes.jitter_cameras(d, seed=seed + 1)
  rc.light_bulbradius = torch.full_like(rc.light_bulbradius, 0.05)
  mj.refit_bvh(m, d, rc)


def _pt_parity(gpu: dict, ref: dict, what: str) -> None:
  np.testing.assert_array_equal(gpu["seg"], ref["seg"], err_msg=what)
  np.testing.assert_allclose(gpu["depth"], ref["depth"], rtol=1e-5, atol=1e-7, err_msg=what)
  diff = np.abs(gpu["hdr"] - ref["hdr"]).max(axis=-1)
  frac = float(np.mean(diff <= 1e-3))
  assert frac >= 0.999, f"{what}: {1 - frac:.4%} of pixels differ by more than 1e-3 (max {diff.max()})"


def test_gpu_wavefront_mixed_resolution_batch_matches_oracle(cuda):
  """Mixed-resolution path-traced batch: the persistent wavefront launch renders three world
  groups (64^2, 128^2, 256^2; 2 worlds each) of the full-feature scene in one launch; each group
  matches the oracle's path tracer, and the per-group megakernel bit for bit."""
  sky = scenes.procedural_sky(0, 64, 128)
  groups = []
  for i, res in enumerate((64, 128, 256)):
    m, d, rc = make(scenes.full_scene(0, res=res), 2, device=cuda, cam_res=(res, res), envmap=sky, **PT_BATCH)
    _randomized(m, d, rc, 2, seed=10 * i + 1)
    groups.append((m, d, rc))
  batch = WavefrontBatch(groups)
  batch.render(seed=3)
  wave = [gpu_outputs(rc) for _, _, rc in groups]
  for i, (m, d, rc) in enumerate(groups):
    print(f"wavefront group {i}: oracle", flush=True)
    ref = oracle.pt_render(m, d, rc, seed=3, threads=16)
    _pt_parity(wave[i], ref, f"group {i}")
  for i, (m, d, rc) in enumerate(groups):
    mj.render(m, d, rc, seed=3)
    mega = gpu_outputs(rc)
    for k in ("rgb", "hdr", "depth", "seg"):
      assert np.array_equal(mega[k], wave[i][k]), f"group {i}: {k} differs from the megakernel"


@pytest.fixture(scope="module")
def mesh_heavy():
  return scenes.mesh_scene(0, res=64)


def test_gpu_mesh_heavy_realtime_matches_oracle(cuda, mesh_heavy):
  """Mesh-heavy scene (~200k triangles in 16 meshes, 64 instances) in the real-time tier."""
  m, d, rc = make(mesh_heavy, 2, device=cuda, cam_res=(64, 64), envmap=scenes.procedural_sky(1, 64, 128), render_guides=True)
  _randomized(m, d, rc, 2, seed=5)
  mj.render(m, d, rc, seed=0)
  gpu = gpu_outputs(rc)
  ref = oracle.rt_render(m, d, rc, seed=0, threads=16)
  compare_rt(gpu, ref)
  np.testing.assert_allclose(gpu["normal"], ref["normal"], atol=1e-5)
  segs = set(np.unique(gpu["seg"][..., 0]).tolist())
  assert len(segs & set(np.nonzero(mesh_heavy.geom_type == mdl.GEOM_MESH)[0].tolist())) > 10


def test_gpu_mesh_heavy_path_traced_matches_oracle(cuda, mesh_heavy):
  """Mesh-heavy scene in the path-traced tier: 1 path a pixel, depth 8."""
  m, d, rc = make(mesh_heavy, 2, device=cuda, cam_res=(64, 64), envmap=scenes.procedural_sky(1, 64, 128), tier=mj.Tier.PATH,
                  path_samples=1, max_depth=8, rr_depth=3)
  _randomized(m, d, rc, 2, seed=6)
  mj.render(m, d, rc, seed=1)
  gpu = gpu_outputs(rc)
  ref = oracle.pt_render(m, d, rc, seed=1, threads=16)
  _pt_parity(gpu, ref, "mesh-heavy PT")


def test_gpu_mesh_heavy_wavefront_matches_megakernel(cuda, mesh_heavy):
  """render() of the mesh-heavy scene at 2 paths a pixel takes the wavefront kernel at 8 waves a
  SIMD (more than 65,536 mesh triangles): the images equal the per-pixel kernel's bit for bit."""
  from mujoco_warp_pbr_amd import wavefront

  m, d, rc = make(mesh_heavy, 2, device=cuda, cam_res=(64, 64), envmap=scenes.procedural_sky(1, 64, 128), tier=mj.Tier.PATH,
                  path_samples=2, max_depth=8, rr_depth=3, render_hdr=True)
  _randomized(m, d, rc, 2, seed=7)
  assert wavefront.render_config(rc) == wavefront.MESH_HEAVY_CONFIG
  mj.render(m, d, rc, seed=2)
  assert "wavefront" in rc._host_cache
  wave = gpu_outputs(rc)
  rc.path_kernel = "megakernel"
  mj.render(m, d, rc, seed=2)
  mega = gpu_outputs(rc)
  for k in ("rgb", "hdr", "depth", "seg"):
    assert np.array_equal(mega[k], wave[k]), f"{k} differs from the megakernel"


def _spot_scene():
  mjm = scenes.primitives_scene(4, 48)
  mjm.light_type[1] = mdl.LIGHT_SPOT
  mjm.light_xpos[1] = (0.2, -0.1, 1.6)
  d = np.array([-0.1, 0.05, -1.0])
  mjm.light_xdir[1] = (d / np.linalg.norm(d)).astype(np.float32)
  mjm.light_cutoff[1] = 30.0
  mjm.light_exponent[1] = 4.0
  mjm.light_intensity[1] = 4.0
  return mjm


@pytest.mark.parametrize(
  "name,kw",
  [
    ("spot_physical", {"light_model": mj.LightModel.PHYSICAL}),
    ("spot_mujoco", {"light_model": mj.LightModel.MUJOCO}),
    ("aces", {"tone_mapper": mj.ToneMapper.ACES}),
  ],
)
def test_gpu_spot_and_aces_match_oracle(cuda, name, kw):
  """Spot lights (smooth physical cone, MuJoCo's cutoff + exponent) and the ACES tone mapper."""
  mjm = _spot_scene() if name.startswith("spot") else scenes.primitives_scene(4, 48)
  m, d, rc = make(mjm, 3, device=cuda, cam_res=(48, 48), envmap=scenes.procedural_sky(2, 64, 128), **kw)
  assert bool(rc.base.has_spot_lights) == name.startswith("spot")
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=0)
  gpu = gpu_outputs(rc)
  ref = oracle.rt_render(m, d, rc, seed=0, threads=16)
  compare_rt(gpu, ref)
  if name == "spot_physical":
    # the cone edge is in the frame: floor inside it is lit by the spot, floor outside is not
    lum = gpu["hdr"][0].reshape(48, 48, 3).sum(-1)
    floor = gpu["seg"][0].reshape(48, 48, 2)[..., 0] == 0
    assert lum[floor].max() - np.percentile(lum[floor], 10) > 0.5
