"""mjpbr_rt_render_noalloc: the real-time render with its per-world staging in the library's
static device array instead of a caller-provided world_stage buffer. Called through the C ABI
and compared with mjpbr_rt_render (which stages into the context's world_stage tensor) bit for
bit, on two calls in a row, the second with fewer worlds than the first; the static array keeps
the first call's contents, which the second call must not read."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd import scenes
from mujoco_warp_pbr_amd.render import outputs_struct
from mujoco_warp_pbr_amd.render import scene_struct
from mujoco_warp_pbr_amd.render import settings_struct
from tests.helpers import gpu_outputs
from tests.helpers import make

pytestmark = pytest.mark.gpu


def _clear(rc):
  for t in (rc.base.rgb_data, rc.base.depth_data, rc.base.seg_data, rc.hdr_data, rc.guide_albedo, rc.guide_normal):
    t.fill_(3)


def _noalloc(m, d, rc, seed):
  sc, s, out = scene_struct(m, d, rc), settings_struct(rc), outputs_struct(rc)
  sc.geo.world_stage = None  # ignored by the entry point; it must not need it
  stream = int(torch.cuda.current_stream().cuda_stream)
  _abi.call("mjpbr_rt_render_noalloc", _abi.byref(sc), _abi.byref(s), _abi.byref(out), ctypes.c_int32(seed), stream)


def _case(cuda, mjm, nworld, res, seed, **kw):
  m, d, rc = make(mjm, nworld=nworld, device=cuda, cam_res=(res, res), render_guides=True, **kw)
  scenes.randomize_worlds(m, rc, nworld, seed=seed)
  scenes.jitter_cameras(d, seed=seed + 1)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=seed)
  want = gpu_outputs(rc)
  _clear(rc)
  print(f"no-alloc render: {nworld} worlds at {res}^2, seed {seed}", flush=True)
  _noalloc(m, d, rc, seed)
  got = gpu_outputs(rc)
  for k in want:
    assert np.array_equal(got[k], want[k]), f"{k} differs in {np.sum(got[k] != want[k])} entries"
  return m, d, rc


def test_gpu_noalloc_render_matches_staged_render_on_consecutive_calls(cuda):
  sky = scenes.procedural_sky(0, 64, 128)
  # first call: 8 worlds of the full-feature scene (meshes, textures, mirror, soft shadows, 2x2)
  _case(cuda, scenes.full_scene(0, res=32), 8, 32, seed=3, envmap=sky, samples_per_pixel=2, shadow_samples=4)
  # second call: fewer worlds of another scene, whose staged records differ in size and content
  _case(cuda, scenes.primitives_scene(1, 24), 3, 24, seed=7, envmap=sky)


def test_gpu_noalloc_rejects_worlds_beyond_the_static_maximum(cuda):
  m, d, rc = make(scenes.primitives_scene(1, 8), nworld=2, device=cuda, cam_res=(8, 8))
  mj.refit_bvh(m, d, rc)
  sc, s, out = scene_struct(m, d, rc), settings_struct(rc), outputs_struct(rc)
  floats = _abi.library().mjpbr_world_stage_floats(_abi.byref(sc))
  limit = int(_abi._defines(_abi._header_text())["MJPBR_NOALLOC_STAGE_BYTES"])
  sc.geo.nworld = limit // (4 * floats) + 1  # refused on the host, before any launch
  with pytest.raises(_abi.MjpbrError, match="allocation-free entry point holds"):
    _abi.call("mjpbr_rt_render_noalloc", _abi.byref(sc), _abi.byref(s), _abi.byref(out), ctypes.c_int32(0), 0)
