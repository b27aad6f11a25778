"""Path-traced tier: the oracle pinned against the reference's invariants
(tests/test_pathtrace.py:39-115), and the GPU kernel against the oracle.

GPU bar: segmentation bit-exact, depth 1e-5 relative, per-sample radiance within 1e-3 (sample 0
over several seeds: each seed draws a different random path), 4-sample means within 1e-3, and
high-sample means within 0.5%.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
import oracle
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd import scenes
from mujoco_warp_pbr_amd.scenes import _Builder
from tests.helpers import gpu_outputs
from tests.helpers import make
from tests.helpers import plane_and_light

PATH = {"tier": mj.Tier.PATH, "render_hdr": True}


def _sphere(metallic, roughness, res=8):
  b = _Builder()
  mat = b.material(rgba=(1, 1, 1, 1), metallic=metallic, roughness=roughness)
  b.geom(mdl.GEOM_SPHERE, (0.5, 0, 0), (0, 0, 0), matid=mat)
  b.camera((0, -3, 0), (0, 0, 0), 8.0, (res, res))
  return b.build()


@pytest.mark.parametrize("metallic,roughness", [(0.0, 0.3), (1.0, 0.3), (0.0, 1.0), (1.0, 1.0)])
def test_white_furnace_oracle(metallic, roughness):
  """tests/test_pathtrace.py:39-45: a white sphere in a white environment looks white."""
  m, d, rc = make(_sphere(metallic, roughness), cam_res=(8, 8), envmap=np.ones((32, 64, 3), np.float32), envmap_intensity=1.0,
                  path_samples=256, max_depth=8, **PATH)
  out = oracle.pt_render(m, d, rc, threads=8)
  patch = out["hdr"][0].reshape(8, 8, 3)[2:6, 2:6].mean(axis=(0, 1))
  np.testing.assert_allclose(patch, 1.0, rtol=0.03)


def test_direct_lambert_matches_realtime_oracle():
  """tests/test_pathtrace.py:48-63: with one interaction both tiers integrate the same light."""
  mjm = plane_and_light(mdl.LIGHT_POINT, pos=(0, 0, 1.5), rgba=(0.8, 0.8, 0.8, 1), intensity=3.0, range_=1000.0, cam_height=2.0)
  mjm.light_bulbradius[:] = 0.0
  m, d, rc = make(mjm, cam_res=(8, 8))
  rt = oracle.rt_render(m, d, rc)["hdr"]
  m, d, rc = make(mjm, cam_res=(8, 8), path_samples=16, max_depth=1, **PATH)
  pt = oracle.pt_render(m, d, rc)["hdr"]
  np.testing.assert_allclose(pt, rt, rtol=0.01)


def test_seeded_and_deterministic_oracle():
  env = np.random.default_rng(0).uniform(0, 2, (32, 64, 3)).astype(np.float32)
  m, d, rc = make(_sphere(0.0, 0.5), cam_res=(8, 8), envmap=env, envmap_intensity=1.0, path_samples=4, **PATH)
  a = oracle.pt_render(m, d, rc, seed=0)["hdr"]
  b = oracle.pt_render(m, d, rc, seed=0)["hdr"]
  c = oracle.pt_render(m, d, rc, seed=1)["hdr"]
  np.testing.assert_array_equal(a, b)
  assert not np.array_equal(a, c)


def _corner():
  b = _Builder()
  b.geom(mdl.GEOM_PLANE, (2, 2, 0.1), (0, 0, 0), rgba=(0.9, 0.9, 0.9, 1))
  b.geom(mdl.GEOM_BOX, (2, 0.01, 1), (0, 1, 1), rgba=(0.9, 0.9, 0.9, 1))
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (0, 0.5, -1), castshadow=1)
  b.camera((0, -1.5, 0.5), (0, 0.3, 0.9), 40.0, (16, 16))
  return b.build()


def test_interreflection_brightens_concave_corner_oracle():
  """tests/test_pathtrace.py:100-115."""
  m, d, rc = make(_corner(), cam_res=(16, 16), path_samples=64, max_depth=1, **PATH)
  one = oracle.pt_render(m, d, rc, threads=8)["hdr"].mean()
  m, d, rc = make(_corner(), cam_res=(16, 16), path_samples=64, max_depth=4, **PATH)
  four = oracle.pt_render(m, d, rc, threads=8)["hdr"].mean()
  assert four > one * 1.1


def test_depth_and_segmentation_match_realtime_oracle():
  """tests/test_pathtrace.py:81-97: sample 0 goes through the pixel centre."""
  mjm = scenes.primitives_scene(4, 24)
  m, d, rc = make(mjm, cam_res=(24, 20), envmap=scenes.constant_sky())
  rt = oracle.rt_render(m, d, rc)
  m, d, rc = make(mjm, cam_res=(24, 20), envmap=scenes.constant_sky(), path_samples=2, **PATH)
  pt = oracle.pt_render(m, d, rc)
  np.testing.assert_array_equal(pt["seg"], rt["seg"])
  np.testing.assert_allclose(pt["depth"], rt["depth"], rtol=1e-5)


def test_denoise_features_oracle():
  """The denoiser's features: normalised normals, albedo in [0, 1], non-negative variance."""
  mjm = scenes.primitives_scene(5, 16)
  m, d, rc = make(mjm, cam_res=(16, 16), envmap=scenes.procedural_sky(0, 32, 64), path_samples=4, denoise=True, **PATH)
  out = oracle.pt_render(m, d, rc)
  n = out["aov_normal"][0]
  lens = np.linalg.norm(n, axis=-1)
  hit = out["seg"][0, :, 0] >= 0
  np.testing.assert_allclose(lens[hit], 1.0, atol=2e-3)
  # a pixel whose samples all miss has a zero normal; any hit gives a normalised one
  assert np.all((lens == 0) | (np.abs(lens - 1.0) < 2e-3))
  assert np.any(lens[~hit] == 0)
  assert np.all(out["aov_variance"] >= 0)
  assert np.all(out["aov_depth"][0][hit] > 0)


# ------------------------------------------------------------------------------------------- GPU


def _pt_both(mjm, nworld, res, seed=0, setup=None, **kw):
  m, d, rc = make(mjm, nworld=nworld, device="cuda", cam_res=res, **kw)
  if setup:
    setup(m, d, rc)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=seed)
  gpu = gpu_outputs(rc)
  if rc.denoise:
    torch.cuda.synchronize()
    gpu["aov_albedo"] = rc.aov_albedo.cpu().numpy().view(np.float16).astype(np.float32)
    gpu["aov_normal"] = rc.aov_normal.cpu().numpy().view(np.float16).astype(np.float32)
    gpu["aov_depth"] = rc.aov_depth.cpu().numpy()
    gpu["aov_emission"] = rc.aov_emission.cpu().numpy()
    gpu["aov_variance"] = rc.aov_variance.cpu().numpy()
  ref = oracle.pt_render(m, d, rc, seed=seed, threads=8)
  return gpu, ref, (m, d, rc)


def _full_setup(nworld):
  def setup(m, d, rc):
    scenes.randomize_worlds(m, rc, nworld, seed=1)
    scenes.jitter_cameras(d, seed=2)
    rc.light_bulbradius = torch.full_like(rc.light_bulbradius, 0.05)

  return setup


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["wavefront", "megakernel"])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gpu_per_sample_radiance_matches_oracle(cuda, seed, kernel):
  """One path per pixel (sample 0): the GPU follows the oracle's random decisions, through the
  wavefront kernel (work items of four tiles) and the per-pixel kernel (its 6-wave instance of
  scenes with light meshes)."""
  mjm = scenes.full_scene(0, res=32)
  gpu, ref, _ = _pt_both(mjm, 4, (32, 32), seed=seed, setup=_full_setup(4), envmap=scenes.procedural_sky(0, 64, 128),
                         path_samples=1, max_depth=8, path_kernel=kernel, **PATH)
  np.testing.assert_array_equal(gpu["seg"], ref["seg"])
  np.testing.assert_allclose(gpu["depth"], ref["depth"], rtol=1e-5, atol=1e-7)
  diff = np.abs(gpu["hdr"] - ref["hdr"]).max(axis=-1)
  # per-sample radiance within 1e-3 (absolute, exposure 1) on every pixel
  assert np.mean(diff <= 1e-3) >= 0.999, f"{np.mean(diff > 1e-3):.4%} of samples differ, max {diff.max()}"


@pytest.mark.gpu
def test_gpu_per_sample_radiance_without_meshes_matches_oracle(cuda):
  """One path per pixel on a scene without meshes: the per-pixel kernel's instance for such
  scenes (4 waves a SIMD; the full-feature scene above runs the 6-wave one, the mesh-heavy
  scene of test_configs_gpu.py the 8-wave one) follows the oracle's random decisions too."""
  mjm = scenes.primitives_scene(3, 32)
  gpu, ref, (_, _, rc) = _pt_both(mjm, 3, (32, 32), seed=4, envmap=scenes.procedural_sky(2, 64, 128), path_samples=1, max_depth=8,
                                  path_kernel="megakernel", **PATH)
  assert not rc.base.has_meshes
  np.testing.assert_array_equal(gpu["seg"], ref["seg"])
  np.testing.assert_allclose(gpu["depth"], ref["depth"], rtol=1e-5, atol=1e-7)
  diff = np.abs(gpu["hdr"] - ref["hdr"]).max(axis=-1)
  assert np.mean(diff <= 1e-3) >= 0.999, f"{np.mean(diff > 1e-3):.4%} of samples differ, max {diff.max()}"


@pytest.mark.gpu
def test_gpu_sample_means_match_oracle(cuda):
  mjm = scenes.full_scene(1, res=24)
  gpu, ref, _ = _pt_both(mjm, 2, (24, 24), setup=_full_setup(2), envmap=scenes.procedural_sky(1, 64, 128), path_samples=4,
                         max_depth=8, rr_depth=3, **PATH)
  diff = np.abs(gpu["hdr"] - ref["hdr"]).max(axis=-1)
  assert np.mean(diff <= 1e-3) >= 0.999, f"max {diff.max()}"
  a = gpu["rgb"].view(np.uint32)
  b = ref["rgb"].view(np.uint32)
  for sh in (0, 8, 16):
    assert np.mean(np.abs(((a >> sh) & 255).astype(int) - ((b >> sh) & 255).astype(int)) <= 1) >= 0.999


@pytest.mark.gpu
@pytest.mark.slow
def test_gpu_4096_sample_means_match_oracle_per_pixel(cuda):
  """4096 paths a pixel, depth 8 (Russian roulette from 3): every pixel's mean radiance agrees
  with the oracle's within 0.5 % (plus 1e-4 absolute for near-black pixels), per channel."""
  mjm = scenes.primitives_scene(6, 16)
  gpu, ref, _ = _pt_both(mjm, 2, (16, 16), envmap=scenes.procedural_sky(2, 32, 64), path_samples=4096, max_depth=8, rr_depth=3,
                         **PATH)
  g, r = gpu["hdr"].astype(np.float64), ref["hdr"].astype(np.float64)
  err = np.abs(g - r) - (5e-3 * np.abs(r) + 1e-4)
  worst = np.unravel_index(np.argmax(err), err.shape)
  assert err.max() <= 0.0, f"pixel {worst[:2]}: gpu {g[worst[:2]]} oracle {r[worst[:2]]}"
  assert np.mean(r) > 0.05


@pytest.mark.gpu
def test_gpu_denoise_features_match_oracle(cuda):
  mjm = scenes.primitives_scene(5, 32)
  gpu, ref, _ = _pt_both(mjm, 2, (32, 32), envmap=scenes.procedural_sky(0, 32, 64), path_samples=4, denoise=True, **PATH)
  for k in ("aov_albedo", "aov_normal", "aov_depth", "aov_emission"):
    np.testing.assert_allclose(gpu[k], ref[k], atol=2e-3, err_msg=k)
  np.testing.assert_allclose(gpu["aov_variance"], ref["aov_variance"], rtol=1e-2, atol=1e-5)
