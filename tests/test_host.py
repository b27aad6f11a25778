"""Host-side logic: material resolution, RGBE decoding, argument validation, exposure, mip
chains, BVH layout (reference tests/test_material.py, tests/test_context_validation.py,
tests/test_render.py:226-229)."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
from tests.helpers import many_cameras_scene
from oracle import tables as oracle_tables
from mujoco_warp_pbr_amd import bvh
from mujoco_warp_pbr_amd import envmap
from mujoco_warp_pbr_amd import material
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd import scenes
from mujoco_warp_pbr_amd import texture
from mujoco_warp_pbr_amd.scenes import _Builder
from mujoco_warp_pbr_amd.types import UnsetMaterial


def _material_model():
  """tests/test_material.py:20-38 restated as a HostModel."""
  b = _Builder()
  authored = b.material(metallic=0.3, roughness=0.6)
  b.lists["mat_specular"][-1] = np.asarray(0.7)
  legacy = b.material(rgba=(0.5, 0.5, 0.5, 1))
  b.lists["mat_shininess"][-1] = np.asarray(0.8)
  white = b.material(rgba=(0.9, 0.9, 0.9, 1))
  b.geom(mdl.GEOM_SPHERE, (0.1, 0, 0), (0, 0, 0), matid=authored)
  b.geom(mdl.GEOM_SPHERE, (0.1, 0, 0), (0, 0, 0), matid=legacy)
  b.geom(mdl.GEOM_SPHERE, (0.1, 0, 0), (0, 0, 0), matid=white)
  b.geom(mdl.GEOM_SPHERE, (0.1, 0, 0), (0, 0, 0), rgba=(0.1, 0.1, 0.1, 1))
  b.geom(mdl.GEOM_PLANE, (1, 1, 0.1), (0, 0, 0), matid=legacy)
  b.camera((0, -2, 1), (0, 0, 0), 45, (8, 8))
  return b.build()


def _resolve(**kw):
  mjm = _material_model()
  m, r, f = material.resolve_geom_materials(mjm, **kw)
  names = ["authored", "legacy", "white", "nomat", "floor"]
  return {n: (m[i], r[i], f[i]) for i, n in enumerate(names)}


def test_gltf_policy():
  r = _resolve()
  np.testing.assert_allclose(r["authored"], (0.3, 0.6, 0.7), rtol=1e-6)
  np.testing.assert_allclose(r["legacy"], (0.0, 0.2, 0.5), rtol=1e-6)
  np.testing.assert_allclose(r["nomat"], (0.0, 0.5, 0.5), rtol=1e-6)


def test_default_material_only_fills_unset_fields():
  r = _resolve(default_material={"metallic": 1.0, "roughness": 0.1})
  np.testing.assert_allclose(r["authored"][:2], (0.3, 0.6), rtol=1e-6)
  np.testing.assert_allclose(r["legacy"][:2], (1.0, 0.1), rtol=1e-6)


def test_heuristic_policy():
  r = _resolve(unset_material=UnsetMaterial.HEURISTIC)
  np.testing.assert_allclose(r["authored"], (0.3, 0.6, 0.7), rtol=1e-6)
  np.testing.assert_allclose(r["legacy"], (0.6, 0.40 * (1.3 - 0.8), 0.7), rtol=1e-5)
  np.testing.assert_allclose(r["white"], (0.0, 0.38, 0.65), rtol=1e-6)
  np.testing.assert_allclose(r["nomat"], (0.4, 0.45, 0.35), rtol=1e-6)
  np.testing.assert_allclose(r["floor"], (0.0, 1.0, 0.0), rtol=1e-6)


def test_exposure_from_ev100():
  """tests/test_render.py:226-229: Filament's 1 / (1.2 * 2^EV100)."""
  assert material.exposure_from_ev100(0.0) == pytest.approx(1.0 / 1.2)
  assert material.exposure_from_ev100(math.log2(1 / 1.2)) == pytest.approx(1.0)
  assert material.exposure_from_ev100(15.0) == pytest.approx(1.0 / (1.2 * 32768.0))


def _rgbe(image):
  peak = image.max(axis=-1)
  exponent = np.where(peak > 1e-32, np.ceil(np.log2(np.maximum(peak, 1e-32))), -128).astype(np.int32)
  mantissa = np.clip(image / np.ldexp(1.0, exponent)[..., None] * 256.0, 0, 255).astype(np.uint8)
  return np.concatenate([mantissa, (exponent + 128).clip(0, 255).astype(np.uint8)[..., None]], axis=-1)


def _rle_channel(values):
  out = bytearray()
  i, n = 0, len(values)
  while i < n:
    run = 1
    while i + run < n and run < 127 and values[i + run] == values[i]:
      run += 1
    if run >= 3:
      out += bytes([128 + run, int(values[i])])
      i += run
      continue
    j = i
    while j < n and j - i < 128 and not (j + 2 < n and values[j] == values[j + 1] == values[j + 2]):
      j += 1
    out += bytes([j - i]) + bytes(values[i:j])
    i = j
  return bytes(out)


def test_load_hdr_run_length(tmp_path):
  """tests/test_context_validation.py:135-143: RLE scanlines decode bit-exactly."""
  rng = np.random.default_rng(0)
  image = rng.uniform(0.0, 4.0, (5, 40, 3)).astype(np.float32)
  image[:, 10:30] = 1.5
  h, w, _ = image.shape
  rgbe = _rgbe(image)
  body = bytearray()
  for y in range(h):
    body += bytes([2, 2, w >> 8, w & 0xFF])
    for c in range(4):
      body += _rle_channel(rgbe[y, :, c])
  (tmp_path / "env.hdr").write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode() + bytes(body))
  decoded = mj.load_hdr(tmp_path / "env.hdr")
  scale = np.ldexp(1.0, rgbe[..., 3].astype(np.int32) - 136)
  np.testing.assert_array_equal(decoded, (rgbe[..., :3] * scale[..., None]).astype(np.float32))


def test_hdr_round_trip(tmp_path):
  image = np.random.default_rng(1).uniform(0.01, 8.0, (6, 12, 3)).astype(np.float32)
  envmap.save_hdr(tmp_path / "a.hdr", image)
  back = mj.load_hdr(tmp_path / "a.hdr")
  # RGBE keeps 8 mantissa bits relative to the pixel's largest channel.
  peak = image.max(axis=-1, keepdims=True)
  assert np.all(np.abs(back - image) <= peak * 2.0 / 256)


@pytest.mark.parametrize(
  "data,match",
  [
    (b"P6\n", "not a Radiance .hdr file"),
    (b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n", "truncated .hdr header"),
    (b"#?RADIANCE\n\n+Y 2 +X 2\n" + bytes(16), "only '-Y H \\+X W'"),
    (b"#?RADIANCE\n\n-Y 2 +X 2\n" + bytes(10), "truncated pixel data"),
    (b"#?RADIANCE\n\n-Y 1 +X 8\n" + bytes([2, 2, 0, 8, 200, 1]), "bad run-length data"),
    (b"#?RADIANCE\n\n-Y 1 +X 8\n" + bytes([2, 2, 0, 9]), "width does not match"),
    (b"#?RADIANCE\n\n-Y 1 +X 8\n" + bytes([2, 2, 0, 8, 136]), "bad run-length data"),
  ],
)
def test_load_hdr_malformed(tmp_path, data, match):
  (tmp_path / "bad.hdr").write_bytes(data)
  with pytest.raises(ValueError, match=match):
    mj.load_hdr(tmp_path / "bad.hdr")


@pytest.mark.parametrize(
  "kwargs,match",
  [
    ({"samples_per_pixel": 0}, "samples_per_pixel must be at least 1"),
    ({"shadow_samples": 0}, "shadow_samples must be at least 1"),
    ({"max_depth": -1}, "max_depth must be at least 0"),
    ({"shadow_light_fraction": 1.5}, r"shadow_light_fraction must be in \[0, 1\]"),
    ({"tone_mapper": 7}, "tone_mapper must be one of"),
    ({"denoise": True}, "denoise applies to the path-traced tier"),
    ({"default_material": {"shiny": 1.0}}, "default_material keys"),
    ({"default_material": {"metallic": 2.0}}, r"default_material\['metallic'\] must be in \[0, 1\]"),
    ({"tier": mj.Tier.PATH, "max_depth": 40, "max_layers": 6}, "max_depth \\+ 4 \\* max_layers"),
    ({"traversal": "rtx"}, 'traversal must be "software"'),
    ({"mesh_bvh_frame": "x"}, "mesh_bvh_frame"),
    ({"envmap": np.zeros((4, 4))}, "envmap\\[0\\] must be a path"),
  ],
)
def test_create_render_context_validation(kwargs, match):
  """tests/test_context_validation.py:41-69: the same ValueError messages."""
  mjm = scenes.primitives_scene(0, 8)
  with pytest.raises(ValueError, match=match):
    mj.create_render_context(mjm, device="cpu", precompute=False, **kwargs)


def test_context_on_cpu_needs_precompute_flag():
  with pytest.raises(ValueError, match="HIP kernels"):
    mj.create_render_context(scenes.primitives_scene(0, 8), device="cpu")


def test_srgb_textures_decode_to_linear_f16():
  mjm = scenes.full_scene(0, res=8, tex_size=16)
  pool = texture.upload_textures(mjm, np.ones(mjm.ntex, bool), "cpu", precompute=False)
  oracle_tables.fill_textures(pool)
  desc = pool.desc.numpy()
  assert desc[0, 3] == texture.FORMAT_RGBA16F  # base colour role: sRGB -> linear f16
  assert desc[1, 3] == texture.FORMAT_RGBA8 and desc[2, 3] == texture.FORMAT_RGBA8
  lv0 = pool.data.numpy()[pool.level_adr.numpy()[0, 0] :].view(np.float16)[: 16 * 16 * 4].reshape(16, 16, 4)
  raw = mjm.tex_data[: 16 * 16 * 4].reshape(16, 16, 4)
  np.testing.assert_allclose(lv0[..., :3].astype(np.float64), material.srgb_to_linear(raw[..., :3] / 255.0), atol=1e-3)


def test_per_world_tables_are_batched():
  mjm = scenes.primitives_scene(0, 8)
  m = mj.put_model(mjm, "cpu")
  assert m.geom_rgba.shape == (1, mjm.ngeom, 4)
  d = mj.put_data(mjm, 3, "cpu")
  assert d.geom_xpos.shape == (3, mjm.ngeom, 3)
  rc = mj.create_render_context(mjm, nworld=3, device="cpu", precompute=False)
  assert rc.geom_roughness.shape == (1, mjm.ngeom)
  assert rc.base.rgb_data.shape == (3, 8 * 8)


@pytest.mark.parametrize(
  "kwargs,match",
  [
    ({"tier": 5}, "tier must be one of"),
    ({"light_model": 3}, "light_model must be one of"),
    ({"render_order": 9}, "render_order must be one of"),
    ({"unset_material": 2}, "unset_material must be one of"),
    ({"path_samples": 0}, "path_samples must be at least 1"),
    ({"max_layers": 0}, "max_layers must be at least 1"),
    ({"default_material": {"roughness": 2.0}}, r"default_material\['roughness'\] must be in"),
    ({"envmap": [np.zeros((8, 16, 3), np.float32), np.zeros((8, 16, 5), np.float32)]}, r"envmap\[1\]"),
    ({"mesh_bvh_constructor": "median"}, "mesh_bvh_constructor must be one of"),
    ({"block_dim": 0}, "block_dim must be at least 1"),
  ],
)
def test_create_render_context_validation_more(kwargs, match):
  """tests/test_context_validation.py:41-69, the remaining cases."""
  with pytest.raises(ValueError, match=match):
    mj.create_render_context(scenes.primitives_scene(0, 8), cam_res=(8, 8), device="cpu", precompute=False, **kwargs)


def test_enum_arguments_accept_ints():
  """tests/test_context_validation.py:72-74."""
  rc = mj.create_render_context(scenes.primitives_scene(0, 8), cam_res=(8, 8), device="cpu", precompute=False, tone_mapper=1, render_order=0, tier=0)
  assert rc.tone_mapper == mj.ToneMapper.ACES and rc.render_order == mj.RenderOrder.ROW


def test_get_hdr_checks():
  """tests/test_context_validation.py:77-82."""
  rc = mj.create_render_context(scenes.primitives_scene(0, 8), cam_res=(8, 6), device="cpu", precompute=False, render_hdr=True)
  with pytest.raises(ValueError, match="camera_index must be in"):
    mj.get_hdr(rc, 1, torch.zeros((1, 6, 8, 3)))
  with pytest.raises(ValueError, match=r"hdr_out must have shape \(<= 1, 6, 8\)"):
    mj.get_hdr(rc, 0, torch.zeros((1, 8, 6, 3)))


def test_env_id_range():
  """tests/test_context_validation.py:85-95: an env_id past the last map is refused at render."""
  from mujoco_warp_pbr_amd.render import check_env_id

  maps = [np.full((8, 16, 3), 0.5, np.float32)] * 2
  rc = mj.create_render_context(scenes.primitives_scene(0, 8), nworld=2, cam_res=(8, 8), device="cpu", precompute=False, envmap=maps)
  check_env_id(rc)
  rc.env_id = torch.tensor([0, 2], dtype=torch.int32)
  with pytest.raises(ValueError, match=r"env_id must be in \[0, 2\)"):
    check_env_id(rc)


def test_many_cameras_render_in_launches_of_at_most_32():
  """Any number of rendering cameras: the renderer launches at most MJPBR_MAX_CAMS (32)
  consecutive cameras at a time, each launch's ray ids continuing the context's (ray_base)."""
  from mujoco_warp_pbr_amd.render import camera_chunks
  from mujoco_warp_pbr_amd.render import outputs_struct
  from mujoco_warp_pbr_amd.render import scene_struct

  mjm = many_cameras_scene(40)
  res = [(8 + 8 * (i % 2), 8 + 4 * (i % 3 == 0)) for i in range(40)]
  rc = mj.create_render_context(mjm, nworld=1, cam_res=res, device="cpu", precompute=False)
  assert rc.base.nrender == 40
  assert camera_chunks(rc) == [(0, 32), (32, 40)]
  m, d = mj.put_model(mjm, "cpu"), mj.put_data(mjm, 1, "cpu")
  pixels = [w * h for w, h in res]
  for first, end in camera_chunks(rc):
    sc = scene_struct(m, d, rc, (first, end))
    out = outputs_struct(rc, (first, end))
    assert sc.cams.nrender == end - first and sc.cams.total_rays == sum(pixels)
    assert sc.cams.ray_base == sum(pixels[:first])
    assert [tuple(sc.cams.cam_res[k]) for k in range(end - first)] == res[first:end]
    assert [out.rgb_adr[k] for k in range(end - first)] == [int(a) for a in rc.base.rgb_adr[first:end]]
  with pytest.raises(ValueError, match="1 to 32"):
    scene_struct(m, d, rc)
