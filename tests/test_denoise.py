"""À-trous denoiser: the oracle against the reference's invariants (tests/test_denoise.py:49-73)
and the GPU filter against the oracle on identical inputs."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
import oracle
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd.scenes import _Builder
from tests.helpers import EV100_UNIT
from tests.helpers import gpu_outputs
from tests.helpers import make

RES = 64
PATH = {"tier": mj.Tier.PATH, "render_hdr": True, "ev100": EV100_UNIT}


def _scene(second_camera=False):
  """tests/test_denoise.py:24-40 restated: checkered rough floor, emissive sphere, random sky.
  second_camera: a second, side-on camera (for per-camera image layouts)."""
  from mujoco_warp_pbr_amd.scenes import procedural_textures

  b = _Builder()
  y, x = np.mgrid[0:64, 0:64]
  checker = ((x // 32 + y // 32) % 2).astype(np.float64)  # MuJoCo's builtin checker: 2x2
  img = (checker[..., None] * 0.9 + (1 - checker[..., None]) * 0.1) * np.ones(3)
  tex = b.texture(np.round(np.concatenate([img, np.ones((64, 64, 1))], -1) * 255).astype(np.uint8), mdl.COLORSPACE_AUTO)
  floor = b.material(rgba=(1, 1, 1, 1), roughness=0.8, texid={mdl.ROLE_RGBA: tex}, repeat=(4, 4))
  glow = b.material(rgba=(0.2, 0.2, 0.2, 1), roughness=0.5, emission=1.0)
  b.geom(mdl.GEOM_PLANE, (2, 2, 0.1), (0, 0, 0), matid=floor)
  b.geom(mdl.GEOM_SPHERE, (0.3, 0, 0), (0, 0, 0.3), matid=glow)
  b.camera((0, -2.5, 1.5), (0, 0, 0.2), 45.0, (RES, RES))
  if second_camera:
    b.camera((2.2, -1.0, 1.0), (0, 0, 0.3), 50.0, (RES, RES))
  return b.build()


def _env():
  return np.random.default_rng(0).uniform(0, 2, (32, 64, 3)).astype(np.float32)


def _render(samples, **kw):
  m, d, rc = make(_scene(), cam_res=(RES, RES), envmap=_env(), envmap_intensity=1.0, path_samples=samples, **PATH, **kw)
  return oracle.pt_render(m, d, rc, threads=8)


def test_zero_iterations_is_identity():
  """tests/test_denoise.py:54-59."""
  raw = _render(4)
  passthrough = _render(4, denoise=True, denoise_iterations=0)
  np.testing.assert_allclose(passthrough["hdr"], raw["hdr"], rtol=1e-4, atol=1e-6)
  a, b = passthrough["rgb"].view(np.uint32), raw["rgb"].view(np.uint32)
  for sh in (0, 8, 16):
    assert np.abs(((a >> sh) & 255).astype(int) - ((b >> sh) & 255).astype(int)).max() <= 1


def test_reduces_error_and_keeps_energy():
  """tests/test_denoise.py:62-73: error < 0.6x raw against a high-sample render (1024 samples
  here, 2048 in the reference), energy within 3%."""
  reference = _render(1024)["hdr"]
  raw = _render(4)["hdr"]
  denoised = _render(4, denoise=True)["hdr"]

  def error(x):
    return np.mean(np.abs(x - reference) / (reference + 0.05))

  assert error(denoised) < 0.6 * error(raw)
  np.testing.assert_allclose(denoised.mean(), reference.mean(), rtol=0.03)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["one_camera", "two_ragged_cameras"])
def test_gpu_atrous_matches_oracle_on_same_inputs(cuda, layout):
  """Same noisy inputs through the GPU filter and the oracle. two_ragged_cameras: two images per
  world whose sizes are not multiples of the 16-pixel tiles of the small-step passes (37 x 21 and
  20 x 45), so partial tiles, halos past the image edge and the second camera's offset are covered."""
  two = layout == "two_ragged_cameras"
  m, d, rc = make(_scene(two), nworld=2, device="cuda", cam_res=[(37, 21), (20, 45)] if two else (RES, RES), envmap=_env(),
                  envmap_intensity=1.0, path_samples=4, denoise=True, **PATH)
  rc.render_hdr = False  # keep the noisy HDR in hdr_data as the filter's input
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=0)
  torch.cuda.synchronize()
  noisy = rc.hdr_data.cpu().numpy().copy()
  aov = {
    "aov_albedo": rc.aov_albedo.cpu().numpy().view(np.float16),
    "aov_normal": rc.aov_normal.cpu().numpy().view(np.float16),
    "aov_depth": rc.aov_depth.cpu().numpy(),
    "aov_emission": rc.aov_emission.cpu().numpy(),
    "aov_variance": rc.aov_variance.cpu().numpy(),
  }
  gpu_rgb = rc.base.rgb_data.cpu().numpy().view(np.uint32).copy()
  # filtered HDR too: run the GPU filter again writing it
  from mujoco_warp_pbr_amd import denoise as dn
  from mujoco_warp_pbr_amd.render import _stream_of

  rc.render_hdr = True
  dn.denoise(rc, 2, _stream_of(rc.hdr_data))
  torch.cuda.synchronize()
  gpu_hdr = rc.hdr_data.cpu().numpy()
  ref = oracle.atrous(rc, noisy, aov, render_hdr=True)
  np.testing.assert_allclose(gpu_hdr, ref["hdr"], rtol=1e-4, atol=1e-5)
  for sh in (0, 8, 16):
    assert np.abs(((gpu_rgb >> sh) & 255).astype(int) - ((ref["rgb"] >> sh) & 255).astype(int)).max() <= 1


@pytest.mark.gpu
def test_gpu_denoised_path_tracing_end_to_end(cuda):
  m, d, rc = make(_scene(), device="cuda", cam_res=(RES, RES), envmap=_env(), envmap_intensity=1.0, path_samples=4, denoise=True,
                  **PATH)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=0)
  gpu = gpu_outputs(rc)
  ref = oracle.pt_render(m, d, rc, seed=0)
  diff = np.abs(gpu["hdr"] - ref["hdr"]).max(axis=-1)
  assert np.mean(diff <= 1e-3) >= 0.99
