"""Real-time tier on the GPU against the CPU oracle (same seeded scenes).

Bar: segmentation bit-exact, depth within 1e-5 relative, linear HDR within 1e-3 absolute per
pixel (exposure 1, so HDR is display-relative), packed sRGB8 within 1 LSB.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
import oracle
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd import scenes
from tests.helpers import EV100_UNIT
from tests.helpers import compare_rt
from tests.helpers import gpu_outputs
from tests.helpers import make
from tests.helpers import many_cameras_scene
from tests.helpers import plane_and_light

pytestmark = pytest.mark.gpu


def _render_both(mjm, nworld, cam_res, seed=0, setup=None, **kw):
  m, d, rc = make(mjm, nworld=nworld, device="cuda", cam_res=cam_res, **kw)
  if setup is not None:
    setup(m, d, rc)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=seed)
  gpu = gpu_outputs(rc)
  ref = oracle.rt_render(m, d, rc, seed=seed, threads=8)
  return gpu, ref, (m, d, rc)


def test_primitives_scene_matches_oracle(cuda):
  """Small real-time case: plane + 8 primitives, directional + point light, grey sky."""
  mjm = scenes.primitives_scene(0, 64)
  gpu, ref, _ = _render_both(mjm, 16, (64, 64), envmap=scenes.constant_sky())
  compare_rt(gpu, ref)
  assert np.mean(gpu["seg"][..., 0] >= 0) > 0.3


@pytest.mark.parametrize("scene", ["material_grid", "primitives"])
def test_reference_suite_scenes_match_oracle(cuda, scene):
  """The reference suite's material_grid (no textures, no mirror: the leanest kernel instance)
  and primitives (checker texture, reflective plane) scenes with an HDR sky and the headlight."""
  fn = scenes.suite_material_grid if scene == "material_grid" else scenes.suite_primitives
  mjm = fn(48)
  gpu, ref, (m, d, rc) = _render_both(mjm, 3, (48, 48), envmap=scenes.procedural_sky(3, 64, 128))
  from mujoco_warp_pbr_amd import _abi
  from mujoco_warp_pbr_amd.render import settings_struct

  expected = 0 if scene == "material_grid" else 40
  assert _abi.library().mjpbr_rt_kernel_features(_abi.byref(settings_struct(rc))) == expected
  compare_rt(gpu, ref)
  assert np.mean(gpu["seg"][..., 0] >= 0) > 0.3


def test_full_scene_matches_oracle(cuda):
  """Full-feature real-time case at reduced size: meshes, textures, HDR sky, mirror plane,
  2x2 supersampling, soft shadows, per-world materials and cameras."""
  mjm = scenes.full_scene(0, res=32)

  def setup(m, d, rc):
    scenes.randomize_worlds(m, rc, 4, seed=1)
    scenes.jitter_cameras(d, seed=2)
    rc.light_bulbradius = torch.full_like(rc.light_bulbradius, 0.05)

  gpu, ref, (m, d, rc) = _render_both(
    mjm, 4, (32, 32), setup=setup, envmap=scenes.procedural_sky(0, 64, 128), samples_per_pixel=2, shadow_samples=4,
    render_guides=True,
  )
  compare_rt(gpu, ref)
  np.testing.assert_allclose(gpu["albedo"], ref["albedo"], atol=1e-5)
  np.testing.assert_allclose(gpu["normal"], ref["normal"], atol=1e-5)
  segs = set(np.unique(gpu["seg"][..., 0]).tolist())
  mesh_geoms = set(np.nonzero(mjm.geom_type == mdl.GEOM_MESH)[0].tolist())
  assert segs & mesh_geoms, "no mesh visible"


def test_render_orders_only_change_seeds(cuda):
  """tests/test_render.py:232-250: without random effects every render order gives identical images."""
  mjm = scenes.primitives_scene(2, 40)
  outs = []
  for order in (mj.RenderOrder.ROW, mj.RenderOrder.TILED, mj.RenderOrder.WORLD):
    m, d, rc = make(mjm, nworld=3, device="cuda", cam_res=(40, 24), render_order=order, envmap=scenes.constant_sky())
    mj.refit_bvh(m, d, rc)
    mj.render(m, d, rc)
    outs.append(gpu_outputs(rc))
  for o in outs[1:]:
    np.testing.assert_array_equal(o["rgb"], outs[0]["rgb"])
    np.testing.assert_array_equal(o["seg"], outs[0]["seg"])


def test_lambertian_plane_point_light_gpu(cuda):
  intensity, height, albedo = 3.0, 1.5, 0.8
  mjm = plane_and_light(mdl.LIGHT_POINT, pos=(0, 0, height), rgba=(albedo, albedo, albedo, 1), intensity=intensity, range_=1000.0)
  m, d, rc = make(mjm, device="cuda", cam_res=(8, 8), tone_mapper=mj.ToneMapper.CLAMP)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc)
  out = torch.zeros((1, 8, 8, 3), device="cuda")
  mj.get_hdr(rc, 0, out)
  np.testing.assert_allclose(out[0, 4, 4].cpu().numpy(), albedo / np.pi * intensity / height**2, rtol=2e-3)


def test_gpu_tables_match_oracle(cuda):
  """DFG LUT and prefiltered stack built by the HIP kernels vs the oracle's."""
  lut = torch.zeros((32, 32, 2), device="cuda")
  from mujoco_warp_pbr_amd import _abi

  _abi.call("mjpbr_dfg_lut", _abi.ptr(lut), 32, 256, 0)
  torch.cuda.synchronize()
  np.testing.assert_allclose(lut.cpu().numpy(), oracle.dfg_lut(32, 256), atol=2e-6)
  mjm = scenes.primitives_scene(0, 8)
  _, _, rc = make(mjm, device="cuda", cam_res=(8, 8), envmap=scenes.procedural_sky(1, 32, 64))
  torch.cuda.synchronize()
  stack = rc.env.prefiltered.cpu().numpy().view(np.float16).astype(np.float32)
  rng = np.random.default_rng(0)
  L, H, W = stack.shape[1:4]
  for _ in range(40):
    lv, y, x = int(rng.integers(L)), int(rng.integers(H)), int(rng.integers(W))
    ref = oracle.prefilter_texel(rc.env, 0, lv, y, x, 512)
    np.testing.assert_allclose(stack[0, lv, y, x, :3], ref, rtol=2e-3, atol=1e-4)


def test_refit_follows_moving_geoms(cuda):
  """Moving a geom and refitting changes what the camera sees exactly as the oracle says."""
  mjm = scenes.primitives_scene(3, 32)
  m, d, rc = make(mjm, nworld=2, device="cuda", cam_res=(32, 32), envmap=scenes.constant_sky())
  mj.refit_bvh(m, d, rc)
  d.geom_xpos[1, 1:] += torch.tensor([0.3, -0.2, 0.1], device="cuda")
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc)
  gpu = gpu_outputs(rc)
  ref = oracle.rt_render(m, d, rc)
  compare_rt(gpu, ref)
  assert np.any(gpu["seg"][0] != gpu["seg"][1])


@pytest.mark.gpu
@pytest.mark.parametrize("scene", ["per_world_metallic", "display"])
def test_render_behaviour_scenes_match_oracle(cuda, scene):
  """tests/test_render.py:186-219 scenes on the GPU: per-world metallic (two worlds) and the
  HDR / display consistency scene, every output against the oracle."""
  from tests.helpers import Shot
  from tests.test_oracle_rt import display_scene
  from tests.test_oracle_rt import per_world_metallic_scene

  if scene == "per_world_metallic":
    s = Shot(per_world_metallic_scene(), cuda, nworld=2)
    s.rc.geom_metallic = torch.tensor([[0.0], [1.0]], device=cuda)
  else:
    s = Shot(display_scene(), cuda, nworld=2, tone_mapper=mj.ToneMapper.CLAMP)
  s.render()
  compare_rt(s.out, oracle.rt_render(s.m, s.d, s.rc))


def _instance_cases():
  def translucent_full():
    mjm = scenes.full_scene(0, res=32)
    mjm.geom_rgba[3, 3] = 0.5
    return mjm

  return [
    (0, lambda: scenes.primitives_scene(1, 32), {}),
    (32, lambda: scenes.suite_primitives(32), {"enable_reflections": False}),
    (40, lambda: scenes.suite_primitives(32), {}),
    (1, lambda: scenes.mesh_scene(1, res=32, nmesh=4, instances=2, ntri=600), {"use_textures": False}),
    (33, lambda: scenes.mesh_scene(1, res=32, nmesh=4, instances=2, ntri=600), {}),
    (47, lambda: scenes.full_scene(1, res=32), {"samples_per_pixel": 3, "shadow_samples": 4}),
    (63, translucent_full, {"samples_per_pixel": 3, "shadow_samples": 4}),
    (47 | 128, lambda: scenes.full_scene(1, res=32), {"samples_per_pixel": 2, "shadow_samples": 4}),
    (63 | 128, translucent_full, {"samples_per_pixel": 2, "shadow_samples": 4}),
  ]


@pytest.mark.parametrize("mask,scene,kw", _instance_cases(), ids=[str(c[0]) for c in _instance_cases()])
def test_every_kernel_instance_matches_oracle(cuda, mask, scene, kw):
  """Each compiled real-time kernel instance (feature mask: 1 meshes, 2 supersampling, 4 soft
  shadows, 8 mirror, 16 layers, 32 textures, 128 one lane a sample at 2x2 supersampling) renders
  its own scene like the oracle: instances are separately compiled code, so each needs its own
  parity check."""
  from mujoco_warp_pbr_amd import _abi
  from mujoco_warp_pbr_amd.render import settings_struct

  mjm = scene()

  def setup(m, d, rc):
    scenes.randomize_worlds(m, rc, 2, seed=mask + 1)
    if (mask & 63) == 63:
      m.geom_rgba[..., 3, 3] = 0.5  # randomize_worlds made every geom opaque
    rc.light_bulbradius = torch.full_like(rc.light_bulbradius, 0.05)

  gpu, ref, (m, d, rc) = _render_both(mjm, 2, (32, 32), setup=setup, envmap=scenes.procedural_sky(mask % 5, 64, 128), **kw)
  assert _abi.library().mjpbr_rt_kernel_features(_abi.byref(settings_struct(rc))) == mask
  compare_rt(gpu, ref)


@pytest.mark.parametrize("ncam", [20, 40])
def test_many_cameras_match_oracle(cuda, ncam):
  """Contexts with 20 and 40 rendering cameras of mixed resolutions (40: two launches of 32 and 8
  cameras, the second's ray ids continuing the first's), soft shadows and supersampling seeded by
  the ray ids: every camera's images follow the oracle, and get_rgb reads the last camera."""
  mjm = many_cameras_scene(ncam)
  res = [(8 + 8 * (i % 2), 8 + 4 * (i % 3 == 0)) for i in range(ncam)]

  def soft(m, d, rc):
    rc.light_bulbradius = torch.full_like(rc.light_bulbradius, 0.05)

  gpu, ref, (_, _, rc) = _render_both(mjm, 2, res, setup=soft, envmap=scenes.constant_sky(), samples_per_pixel=2, shadow_samples=4)
  assert rc.base.nrender == ncam
  compare_rt(gpu, ref)
  seg = gpu["seg"][..., 0]
  assert len(np.unique(seg)) >= 5  # plane, three primitives, background
  w, h = res[-1]
  img = torch.zeros((2, h, w, 3), dtype=torch.float32, device="cuda")
  mj.get_rgb(rc, ncam - 1, img)
  adr = int(rc.base.rgb_adr[ncam - 1])
  packed = gpu["rgb"][:, adr : adr + w * h].astype(np.int64).reshape(2, h, w)
  want = np.stack([(packed >> 16) & 255, (packed >> 8) & 255, packed & 255], axis=-1)
  np.testing.assert_array_equal(np.rint(img.cpu().numpy() * 255.0).astype(np.int64), want)


def test_many_cameras_path_traced_with_atrous_match_oracle(cuda):
  """40 cameras on the path-traced tier with a-trous denoising: both launches' samples and the
  per-launch filter follow the oracle."""
  mjm = many_cameras_scene(40)
  res = [(8 + 8 * (i % 2), 8 + 4 * (i % 3 == 0)) for i in range(40)]
  m, d, rc = make(mjm, nworld=2, device="cuda", cam_res=res, envmap=scenes.constant_sky(), tier=mj.Tier.PATH, path_samples=2, max_depth=3,
                  denoise=True, render_hdr=True)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=3)
  gpu = gpu_outputs(rc)
  ref = oracle.pt_render(m, d, rc, seed=3, threads=8)
  np.testing.assert_array_equal(gpu["seg"], ref["seg"])
  np.testing.assert_allclose(gpu["depth"], ref["depth"], rtol=1e-5, atol=1e-7)
  diff = np.abs(gpu["hdr"] - ref["hdr"]).max(axis=-1)
  assert np.mean(diff <= 1e-3) >= 0.999, f"{np.mean(diff > 1e-3):.4%} of pixels differ, max {diff.max()}"


def test_one_lane_a_sample_with_a_depth_only_camera_matches_oracle(cuda):
  """2x2 supersampling (one lane a sample: a pixel's four lanes) with a ragged 24x12 camera that
  renders depth and segmentation only (one lane of the four traces it) beside a 16x16 RGB camera,
  soft shadows on: every buffer follows the oracle."""
  from mujoco_warp_pbr_amd import _abi
  from mujoco_warp_pbr_amd.render import settings_struct
  from tests.test_readback import _two_camera_scene

  mjm, _ = _two_camera_scene()

  def soft(m, d, rc):
    rc.light_bulbradius = torch.full_like(rc.light_bulbradius, 0.05)

  gpu, ref, (_, _, rc) = _render_both(mjm, 2, None, setup=soft, render_rgb=[True, False], render_depth=[True, True],
                                      render_seg=[True, True], render_guides=True, samples_per_pixel=2, shadow_samples=4)
  assert _abi.library().mjpbr_rt_kernel_features(_abi.byref(settings_struct(rc))) & 128
  compare_rt(gpu, ref)
  np.testing.assert_allclose(gpu["normal"], ref["normal"], atol=1e-5)
