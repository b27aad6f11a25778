"""Persistent wavefront path tracer: batches of world groups with mixed camera resolutions (edge
tiles, two cameras per world) must reproduce the per-pixel megakernel bit for bit, for every
launch shape the autotuner can pick, eagerly and as a replayed graph; render() of a path-traced
context takes it, denoiser features included."""

from __future__ import annotations

import contextlib
import ctypes

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
from mujoco_warp_pbr_amd import autotune
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd import scenes
from mujoco_warp_pbr_amd.render import render_path_samples
from mujoco_warp_pbr_amd.scenes import _Builder
from mujoco_warp_pbr_amd.wavefront import WavefrontBatch
from mujoco_warp_pbr_amd.wavefront import WavefrontConfig
from tests.helpers import make

PT = {"tier": mj.Tier.PATH, "path_samples": 3, "max_depth": 6, "rr_depth": 2, "render_guides": True}


def _scene():
  """Textured mirror plane, primitives, one convex mesh, sun + soft point light, two cameras."""
  rng = np.random.default_rng(7)
  b = _Builder()
  tex = [b.texture(img, cs) for img, cs in scenes.procedural_textures(7, 64)]
  roles = {mdl.ROLE_RGBA: tex[0], mdl.ROLE_ORM: tex[1], mdl.ROLE_NORMAL: tex[2]}
  floor = b.material(rgba=(1, 1, 1, 1), roughness=0.6, reflectance=0.3, texid=roles, repeat=(4, 4))
  b.geom(mdl.GEOM_PLANE, (3, 3, 0.1), (0, 0, 0), matid=floor)
  for i, k in enumerate((mdl.GEOM_SPHERE, mdl.GEOM_BOX, mdl.GEOM_CAPSULE, mdl.GEOM_CYLINDER)):
    mat = b.material(rgba=(*rng.uniform(0.2, 1, 3), 1), roughness=float(rng.uniform(0.1, 1)), metallic=float(i == 1))
    b.geom(k, (0.15, 0.12, 0.1), (rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6), 0.2), mat=scenes.random_rotation(rng), matid=mat)
  verts, faces = scenes.convex_hull_mesh(rng, 200, 0.2)
  b.geom(mdl.GEOM_MESH, (0, 0, 0), (0.2, 0.3, 0.25), matid=b.material(rgba=(0.8, 0.5, 0.3, 1), roughness=0.4), dataid=b.mesh(verts, faces))
  b.light(mdl.LIGHT_DIRECTIONAL, (0, 0, 3), (-0.4, 0.3, -0.86), bulbradius=0.05)
  b.light(mdl.LIGHT_POINT, (0.8, -0.6, 2.0), (0, 0, -1), bulbradius=0.05)
  b.camera((0.0, -2.5, 1.6), (0, 0, 0.2), 45.0, (32, 32))
  b.camera((1.8, -1.2, 1.0), (0, 0, 0.2), 60.0, (32, 32))
  return b.build()


def _group(device, nworld, res, seed):
  m, d, rc = make(_scene(), nworld, device=device, cam_res=res, envmap=scenes.procedural_sky(1, 32, 64), envmap_intensity=1.0, **PT)
  if device != "cpu":
    scenes.randomize_worlds(m, rc, nworld, seed=seed)
    scenes.jitter_cameras(d, seed=seed + 1, offset=0.1, angle_deg=3.0)
    mj.refit_bvh(m, d, rc)
  return m, d, rc


GROUPS = [(3, [(16, 16), (24, 8)]), (2, [(40, 24), (16, 16)]), (2, [(48, 48), (32, 16)])]


def _outputs(rc):
  torch.cuda.synchronize()
  return {
    "rgb": rc.base.rgb_data.cpu().numpy().copy(),
    "hdr": rc.hdr_data.cpu().numpy().copy(),
    "depth": rc.base.depth_data.cpu().numpy().copy(),
    "seg": rc.base.seg_data.cpu().numpy().copy(),
    "albedo": rc.guide_albedo.cpu().numpy().copy(),
    "normal": rc.guide_normal.cpu().numpy().copy(),
  }


def _clear(rc):
  for t in (rc.base.rgb_data, rc.hdr_data, rc.base.depth_data, rc.base.seg_data, rc.guide_albedo, rc.guide_normal):
    t.fill_(7)


@contextlib.contextmanager
def _path_kernel(rc, name):
  old = rc.path_kernel
  rc.path_kernel = name
  try:
    yield
  finally:
    rc.path_kernel = old


def _mega_render(m, d, rc, seed):
  """render() through the per-pixel megakernel (the reference's scheme)."""
  with _path_kernel(rc, "megakernel"):
    mj.render(m, d, rc, seed=seed)


def _assert_same(a, b, what):
  for k in a:
    assert np.array_equal(a[k], b[k]), f"{what}: {k} differs in {np.sum(a[k] != b[k])} entries"


def test_batch_validation():
  groups = [_group("cpu", 2, [(16, 16), (16, 16)], 0)]
  m, d, rc = make(_scene(), 2, device="cpu", cam_res=[(16, 16), (16, 16)], tier=mj.Tier.REALTIME)
  with pytest.raises(ValueError, match="path-traced"):
    WavefrontBatch([(m, d, rc)])
  m2, d2, rc2 = make(_scene(), 2, device="cpu", cam_res=[(16, 16), (16, 16)], envmap=scenes.procedural_sky(1, 32, 64), **{**PT, "path_samples": 5})
  with pytest.raises(ValueError, match="settings differ"):
    WavefrontBatch(groups + [(m2, d2, rc2)])
  with pytest.raises(ValueError, match="runs on a GPU"):
    WavefrontBatch(groups)
  with pytest.raises(ValueError, match="1 to 64"):
    WavefrontBatch(groups * 65)


@pytest.mark.gpu
def test_gpu_wavefront_matches_megakernel(cuda):
  groups = [_group(cuda, n, res, 10 * i) for i, (n, res) in enumerate(GROUPS)]
  ref = []
  for m, d, rc in groups:
    _mega_render(m, d, rc, 11)
    ref.append(_outputs(rc))
  batch = WavefrontBatch(groups)
  for cfg in (WavefrontConfig(), WavefrontConfig(4, 2, 1), WavefrontConfig(8, 8, 64), WavefrontConfig(8, 1, 32)):
    for _, _, rc in groups:
      _clear(rc)
    batch.render(seed=11, config=cfg)
    for i, (_, _, rc) in enumerate(groups):
      _assert_same(_outputs(rc), ref[i], f"group {i} {cfg}")


@pytest.mark.gpu
def test_gpu_wavefront_batch_of_many_groups_matches_megakernel(cuda):
  """Twelve world groups (more than the first builds allowed) of mixed resolutions, camera counts
  and world counts in one launch: every group's images equal the megakernel's bit for bit."""
  shapes = [(1, [(16, 16), (8, 8)]), (2, [(16, 16), (8, 24)]), (1, [(32, 16), (16, 16)]), (3, [(16, 16), (16, 8)]),
            (1, [(24, 24), (16, 16)]), (2, [(8, 8), (24, 16)])] * 2
  groups = [_group(cuda, n, res, 31 + 7 * i) for i, (n, res) in enumerate(shapes)]
  ref = []
  for m, d, rc in groups:
    _mega_render(m, d, rc, 5)
    ref.append(_outputs(rc))
  batch = WavefrontBatch(groups)
  for _, _, rc in groups:
    _clear(rc)
  batch.render(seed=5)
  for i, (_, _, rc) in enumerate(groups):
    _assert_same(_outputs(rc), ref[i], f"group {i}")


@pytest.mark.gpu
def test_gpu_wavefront_graph_replays_match_eager(cuda):
  groups = [_group(cuda, n, res, 10 * i) for i, (n, res) in enumerate(GROUPS)]
  batch = WavefrontBatch(groups)
  batch.render(seed=4)
  eager = [_outputs(rc) for _, _, rc in groups]
  graph = batch.capture(seed=4)
  for rep in range(3):
    for _, _, rc in groups:
      _clear(rc)
    torch.cuda.synchronize()
    graph.replay()
    for i, (_, _, rc) in enumerate(groups):
      _assert_same(_outputs(rc), eager[i], f"replay {rep} group {i}")
    # eager work between replays
    mj.render(*groups[0], seed=5)
    torch.cuda.synchronize()


@pytest.mark.gpu
def test_gpu_autotuner_searches_and_caches(cuda, tmp_path, monkeypatch):
  monkeypatch.setenv("MJPBR_CACHE_DIR", str(tmp_path))
  groups = [_group(cuda, n, res, 10 * i) for i, (n, res) in enumerate(GROUPS)]
  batch = WavefrontBatch(groups)
  best, table = autotune.tune(batch, seed=2)
  assert len(table) >= 8
  for knob, values in autotune.SPACE.items():
    assert getattr(best, knob) in values
  again, table2 = autotune.tune(WavefrontBatch(groups), seed=2)
  assert again == best and table2[0][0] == "cached"
  # the tuned shape renders the same image as the default one
  batch.render(seed=2, config=WavefrontConfig())
  a = [_outputs(rc) for _, _, rc in groups]
  batch.render(seed=2, config=best)
  for i, (_, _, rc) in enumerate(groups):
    _assert_same(_outputs(rc), a[i], f"tuned group {i}")


@pytest.mark.gpu
def test_gpu_ray_counts_agree_between_kernels_and_leave_images_unchanged(cuda):
  """The counting kernel instances render the same images and count the same rays in the
  per-pixel megakernel and the wavefront tracer (the same paths); every path traces at least its
  camera segment, and each light and the sky add shadow rays."""
  groups = [_group(cuda, n, res, 10 * i) for i, (n, res) in enumerate(GROUPS)]
  mega = {k: 0 for k in ("segments", "light_shadow", "env_shadow", "total")}
  single = dict(mega)
  ref = []
  for m, d, rc in groups:
    _mega_render(m, d, rc, 5)
    ref.append(_outputs(rc))
    _clear(rc)
    with _path_kernel(rc, "megakernel"):
      for k, v in mj.count_path_rays(m, d, rc, seed=5).items():
        mega[k] += v
  for i, (m, d, rc) in enumerate(groups):
    _assert_same(_outputs(rc), ref[i], f"counting megakernel, group {i}")
    _clear(rc)
    for k, v in mj.count_path_rays(m, d, rc, seed=5).items():  # render()'s wavefront launch
      single[k] += v
    _assert_same(_outputs(rc), ref[i], f"counting render() wavefront, group {i}")
    _clear(rc)
  batch = WavefrontBatch(groups)
  wf = batch.count_rays(seed=5)
  for i, (_, _, rc) in enumerate(groups):
    _assert_same(_outputs(rc), ref[i], f"counting wavefront, group {i}")
  assert wf == mega and single == mega
  paths = sum(rc.base.nworld * rc.base.total_rays for _, _, rc in groups) * groups[0][2].path_samples
  assert mega["segments"] >= paths
  assert mega["light_shadow"] > 0 and mega["env_shadow"] > 0
  assert mega["total"] == mega["segments"] + mega["light_shadow"] + mega["env_shadow"]


@pytest.mark.gpu
def test_gpu_ray_counts_of_camera_segments_only(cuda):
  """max_depth 0: a path ends at its first hit before any next-event estimation, so the count is
  exactly one segment a path and no shadow rays."""
  m, d, rc = make(_scene(), 3, device=cuda, cam_res=(32, 32), envmap=scenes.procedural_sky(1, 32, 64),
                  **{**PT, "max_depth": 0, "path_samples": 3})
  mj.refit_bvh(m, d, rc)
  counts = mj.count_path_rays(m, d, rc, seed=2)
  assert counts == {"segments": 3 * rc.base.total_rays * 3, "light_shadow": 0, "env_shadow": 0, "total": 3 * rc.base.total_rays * 3}
  with pytest.raises(ValueError, match="path-traced"):
    m2, d2, rc2 = make(_scene(), 1, device=cuda, cam_res=(32, 32))
    mj.count_path_rays(m2, d2, rc2)


def _denoise_outputs(rc):
  torch.cuda.synchronize()
  out = {"rgb": rc.base.rgb_data.cpu().numpy().copy(), "hdr": rc.hdr_data.cpu().numpy().copy()}
  for k in ("aov_albedo", "aov_normal", "aov_depth", "aov_emission", "aov_variance"):
    out[k] = getattr(rc, k).cpu().numpy().copy()
  out["depth"] = rc.base.depth_data.cpu().numpy().copy()
  out["seg"] = rc.base.seg_data.cpu().numpy().copy()
  return out


@pytest.mark.gpu
@pytest.mark.parametrize("spp", [1, 5])
def test_gpu_render_takes_wavefront_with_denoiser_features(cuda, spp):
  """render() of a path-traced context with the a-trous denoiser and path_kernel="wavefront"
  launches the wavefront kernel (its workspace appears in the context), whose noisy HDR, first-hit features and filtered image
  equal the megakernel's bit for bit; edge tiles (40x24) and two cameras per world."""
  m, d, rc = make(_scene(), 3, device=cuda, cam_res=[(40, 24), (16, 16)], envmap=scenes.procedural_sky(1, 32, 64), envmap_intensity=1.0,
                  **{**PT, "path_samples": spp, "denoise": True, "render_hdr": True, "path_kernel": "wavefront"})
  scenes.randomize_worlds(m, rc, 3, seed=3)
  mj.refit_bvh(m, d, rc)
  _mega_render(m, d, rc, 9)
  ref = _denoise_outputs(rc)
  assert "wavefront" not in rc._host_cache
  for t in (rc.base.rgb_data, rc.hdr_data, rc.aov_albedo, rc.aov_normal, rc.aov_depth, rc.aov_emission, rc.aov_variance):
    t.fill_(3)
  mj.render(m, d, rc, seed=9)
  assert "wavefront" in rc._host_cache
  _assert_same(_denoise_outputs(rc), ref, f"render() wavefront, {spp} spp")
  # only the noisy HDR and features (no filter): the same through both kernels
  render_path_samples(m, d, rc, seed=4)
  a = _denoise_outputs(rc)
  with _path_kernel(rc, "megakernel"):
    render_path_samples(m, d, rc, seed=4)
  _assert_same(_denoise_outputs(rc), a, "render_path_samples")


@pytest.mark.gpu
def test_gpu_render_wavefront_follows_replaced_buffers_and_graph_capture(cuda):
  """A replaced per-world array re-prepares the context's wavefront descriptors; render() under
  graph capture replays the eager images (prepared before the capture), and a capture of a
  context never rendered eagerly falls back to the megakernel (the same images)."""
  m, d, rc = make(_scene(), 2, device=cuda, cam_res=(24, 24), envmap=scenes.procedural_sky(1, 32, 64), envmap_intensity=1.0, **PT)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=1)
  key = rc._host_cache["wavefront"].key
  m.mat_rgba = m.mat_rgba.clone() * 0.5
  _mega_render(m, d, rc, 1)
  ref = _outputs(rc)
  _clear(rc)
  mj.render(m, d, rc, seed=1)
  assert rc._host_cache["wavefront"].key != key
  _assert_same(_outputs(rc), ref, "after a replaced array")
  torch.cuda.synchronize()
  graph = torch.cuda.CUDAGraph()
  with torch.cuda.graph(graph):
    mj.render(m, d, rc, seed=1)
  for rep in range(3):
    _clear(rc)
    torch.cuda.synchronize()
    graph.replay()
    _assert_same(_outputs(rc), ref, f"replay {rep}")
    mj.render(m, d, rc, seed=2)  # eager work between replays
    torch.cuda.synchronize()
  m2, d2, rc2 = make(_scene(), 2, device=cuda, cam_res=(24, 24), envmap=scenes.procedural_sky(1, 32, 64), envmap_intensity=1.0, **PT)
  mj.refit_bvh(m2, d2, rc2)
  torch.cuda.synchronize()
  graph2 = torch.cuda.CUDAGraph()
  with torch.cuda.graph(graph2):
    mj.render(m2, d2, rc2, seed=1)
  assert "wavefront" not in rc2._host_cache
  graph2.replay()
  b = _outputs(rc2)
  _mega_render(m2, d2, rc2, 1)
  _assert_same(_outputs(rc2), b, "captured megakernel fallback")


def _bare_scene(S, nworld, res):
  sc = S["MjpbrScene"]()
  sc.cams.nrender = 1
  sc.cams.cam_res[0][0], sc.cams.cam_res[0][1] = res
  sc.cams.total_rays = res[0] * res[1]
  sc.geo.nworld = nworld
  sc.geo.tlas_depth = 1
  return sc


def test_workspace_holds_one_sample_region_a_workgroup():
  """mjpbr_pt_wavefront_workspace: a fixed header plus one region a persistent workgroup (at most
  one a work item) of tiles x 256 pixels x samples x 3 floats (12 with the denoiser features); a
  work item is a run of one camera's tiles, 4 of them at 1 sample a pixel, else 1."""
  from mujoco_warp_pbr_amd import _abi

  lib, S = _abi.library(), _abi.structs()
  scenes_arr = (S["MjpbrScene"] * 1)(_bare_scene(S, 3, (40, 24)))  # 3 x 2 tiles a world
  s = S["MjpbrSettings"]()

  def size(spp, denoise=0, gpc=16):
    s.path_samples, s.denoise = spp, denoise
    return lib.mjpbr_pt_wavefront_workspace(ctypes.addressof(scenes_arr), 1, ctypes.addressof(s), gpc)

  def regions(spp, tiles_per_item, floats=3):
    items = 3 * -(-6 // tiles_per_item)  # worlds x runs of the camera's 6 tiles
    return items * tiles_per_item * 256 * spp * floats * 4

  fixed = size(4) - regions(4, 1)
  assert fixed > 0
  assert size(7) == fixed + regions(7, 1)
  assert size(2) == fixed + regions(2, 1) and size(3) == fixed + regions(3, 1)
  assert size(1) == fixed + regions(1, 4)
  assert size(5, denoise=1) == fixed + regions(5, 1, 12)
  assert size(1, gpc=1) == size(1)  # 6 items: fewer regions than any grid
  for bad in (0, 17):
    assert size(1, gpc=bad) == -1
  assert size(0) == -1


@pytest.mark.gpu
def test_gpu_render_beyond_the_workspace_budget_takes_the_megakernel(cuda, monkeypatch):
  from mujoco_warp_pbr_amd import wavefront

  m, d, rc = make(_scene(), 2, device=cuda, cam_res=(24, 24), envmap=scenes.procedural_sky(1, 32, 64), envmap_intensity=1.0, **PT)
  mj.refit_bvh(m, d, rc)
  monkeypatch.setattr(wavefront, "WORKSPACE_BUDGET", 1 << 10)
  mj.render(m, d, rc, seed=6)
  assert "wavefront" not in rc._host_cache
  a = _outputs(rc)
  monkeypatch.undo()
  _clear(rc)
  mj.render(m, d, rc, seed=6)
  assert "wavefront" in rc._host_cache
  _assert_same(_outputs(rc), a, "megakernel beyond the budget")


def test_auto_path_kernel_choice():
  """path_kernel "auto": the wavefront kernel, except the per-pixel kernel at 1 sample a pixel on
  scenes of more than 65,536 rendered mesh triangles (counted once a geom); "wavefront" and
  "megakernel" force one. The wavefront launch shape: 8 waves a SIMD on such mesh-heavy scenes,
  4 otherwise."""
  from mujoco_warp_pbr_amd import io
  from mujoco_warp_pbr_amd import wavefront
  from mujoco_warp_pbr_amd.render import AUTO_MEGAKERNEL_MESH_TRIANGLES
  from mujoco_warp_pbr_amd.render import _wavefront_pays

  mjm = _scene()
  tris = io._rendered_mesh_triangles(mjm, np.ones(mjm.ngeom, bool))
  mesh_geom = int(np.nonzero(mjm.geom_type == mdl.GEOM_MESH)[0][0])
  assert tris == int(mjm.mesh_facenum[mjm.geom_dataid[mesh_geom]]) > 0
  hidden = np.ones(mjm.ngeom, bool)
  hidden[mesh_geom] = False
  assert io._rendered_mesh_triangles(mjm, hidden) == 0
  m, d, rc = make(mjm, 1, device="cpu", cam_res=(16, 16), **PT)
  assert rc.path_kernel == "auto" and rc.base.mesh_triangles == tris
  assert _wavefront_pays(rc)
  assert wavefront.render_config(rc) == wavefront.RENDER_CONFIG
  rc.path_samples = 1
  assert _wavefront_pays(rc)
  rc.base.mesh_triangles = AUTO_MEGAKERNEL_MESH_TRIANGLES + 1
  assert not _wavefront_pays(rc)
  assert wavefront.render_config(rc) == wavefront.MESH_HEAVY_CONFIG
  rc.path_samples = 2
  assert _wavefront_pays(rc)
  rc.path_kernel = "megakernel"
  assert not _wavefront_pays(rc)
  rc.path_kernel, rc.path_samples = "wavefront", 1
  assert _wavefront_pays(rc)
  assert wavefront.MESH_HEAVY_CONFIG.waves_per_eu == 8 and wavefront.RENDER_CONFIG.waves_per_eu == 4
  with pytest.raises(ValueError, match="path_kernel"):
    make(mjm, 1, device="cpu", cam_res=(16, 16), **{**PT, "path_kernel": "fast"})


@pytest.mark.gpu
def test_gpu_captured_render_keeps_its_buffers_after_a_reprepare(cuda):
  """A graph of render() keeps rendering the buffers it was captured with after render() moved to
  new ones (a larger sample count needs a larger workspace): the captured workspace stays alive
  and unmodified."""
  m, d, rc = make(_scene(), 2, device=cuda, cam_res=(24, 24), envmap=scenes.procedural_sky(1, 32, 64), envmap_intensity=1.0, **PT)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc, seed=3)
  ref = _outputs(rc)
  graph = torch.cuda.CUDAGraph()
  with torch.cuda.graph(graph):
    mj.render(m, d, rc, seed=3)
  old_rgb = rc.base.rgb_data
  rc.base.rgb_data = torch.zeros_like(old_rgb)
  rc.path_samples = 6  # a new settings key and a larger workspace
  mj.render(m, d, rc, seed=3)
  torch.cuda.synchronize()
  assert len(rc._host_cache["wavefront"].retired) == 1
  old_rgb.fill_(7)
  graph.replay()
  torch.cuda.synchronize()
  assert torch.equal(old_rgb.cpu(), torch.as_tensor(ref["rgb"]))
