"""Readback (reference output.py:34-152) and the per-entry-point world-count check.

The get_rgb / get_depth / get_segmentation paths are tensor views and copies, so their
semantics are checked here on CPU tensors filled with the oracle's outputs; the HIP extraction
of get_hdr / get_albedo / get_normal is compared on a GPU in test_rt_gpu.py.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import mujoco_warp_pbr_amd as mj
import oracle
from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd import scenes
from mujoco_warp_pbr_amd.render import render_path_samples
from tests.helpers import make


def _two_camera_scene():
  """Two cameras of different resolutions (16 x 16 and 24 x 12); the second renders depth and
  segmentation only."""
  mjm = scenes.primitives_scene(0, 16)

  def cat(a, row):
    return np.concatenate([a, np.asarray(row, a.dtype)[None]])

  mjm.cam_xpos = cat(mjm.cam_xpos, mjm.cam_xpos[0] + np.array([0.1, 0.0, 0.0], np.float32))
  mjm.cam_xmat = cat(mjm.cam_xmat, mjm.cam_xmat[0])
  mjm.cam_fovy = cat(mjm.cam_fovy, 50.0)
  mjm.cam_projection = cat(mjm.cam_projection, 0)
  mjm.cam_sensorsize = cat(mjm.cam_sensorsize, (0.0, 0.0))
  mjm.cam_intrinsic = cat(mjm.cam_intrinsic, (0.0, 0.0, 0.0, 0.0))
  mjm.cam_resolution = cat(mjm.cam_resolution, (24, 12))
  return mjm, 1


def _shot(**kw):
  mjm, _ = _two_camera_scene()
  m, d, rc = make(
    mjm,
    nworld=2,
    cam_res=None,
    render_rgb=[True, False],
    render_depth=[True, True],
    render_seg=[True, True],
    render_guides=True,
    **kw,
  )
  out = oracle.rt_render(m, d, rc)
  rc.base.rgb_data.copy_(torch.from_numpy(out["rgb"].view(np.int32)))
  rc.base.depth_data.copy_(torch.from_numpy(out["depth"]))
  rc.base.seg_data.copy_(torch.from_numpy(out["seg"]))
  return m, d, rc, out


def test_get_rgb_unpacks_mujoco_warp_byte_order():
  m, d, rc, out = _shot()
  w, h = (int(x) for x in rc.base.cam_res[0])
  rgb = torch.zeros((2, h, w, 3))
  mj.get_rgb(rc, 0, rgb)
  words = out["rgb"][:, : w * h].reshape(2, h, w)
  for c, shift in enumerate((16, 8, 0)):
    np.testing.assert_array_equal(rgb[..., c].numpy(), ((words >> shift) & 0xFF).astype(np.float32) / 255.0)
  # fewer worlds than the context is allowed (reference: out.shape[0] <= nworld)
  one = torch.zeros((1, h, w, 3))
  mj.get_rgb(rc, 0, one)
  np.testing.assert_array_equal(one.numpy(), rgb[:1].numpy())


def test_get_depth_is_scaled_and_clamped():
  """output.py:41-45: planar depth / depth_scale clamped to [0, 1]; background (depth 0) is 0."""
  m, d, rc, out = _shot()
  for cam in (0, 1):
    w, h = (int(x) for x in rc.base.cam_res[cam])
    adr = int(rc.base.depth_adr[cam])
    raw = out["depth"][:, adr : adr + w * h].reshape(2, h, w)
    assert (raw == 0).any() and (raw > 0).any()
    for scale in (1.0, 2.5, 10.0):
      got = torch.zeros((2, h, w))
      mj.get_depth(rc, cam, scale, got)
      np.testing.assert_array_equal(got.numpy(), np.clip(raw / np.float32(scale), 0.0, 1.0))
    assert (raw > 1.0).any(), "some pixels lie beyond depth_scale 1 and must clamp"


def test_get_segmentation_copies_ids_and_types():
  m, d, rc, out = _shot()
  w, h = (int(x) for x in rc.base.cam_res[1])
  adr = int(rc.base.seg_adr[1])
  seg = torch.zeros((2, h, w, 2), dtype=torch.int32)
  mj.get_segmentation(rc, 1, seg)
  np.testing.assert_array_equal(seg.numpy(), out["seg"][:, adr : adr + w * h].reshape(2, h, w, 2))
  assert (seg[..., 0] == -1).any() and (seg[..., 0] >= 0).any()
  bg = seg[..., 0] == -1
  assert (seg[..., 1][bg] == -1).all()
  assert (seg[..., 1][~bg] == mdl.OBJ_GEOM).all()


@pytest.mark.parametrize(
  "fn,cam,shape,dtype,match",
  [
    ("get_rgb", 0, (2, 16, 17, 3), torch.float32, "rgb_out must have shape"),
    ("get_rgb", 0, (3, 16, 16, 3), torch.float32, "rgb_out must have shape"),
    ("get_rgb", 0, (2, 16, 16, 3), torch.float64, "rgb_out must have dtype"),
    ("get_rgb", 1, (2, 12, 24, 3), torch.float32, "does not render rgb"),
    ("get_rgb", 2, (2, 16, 16, 3), torch.float32, "camera_index must be in"),
    ("get_rgb", -1, (2, 16, 16, 3), torch.float32, "camera_index must be in"),
    ("get_depth", 1, (2, 24, 12), torch.float32, "depth_out must have shape"),
    ("get_depth", 1, (2, 12, 24), torch.int32, "depth_out must have dtype"),
    ("get_segmentation", 0, (2, 16, 16), torch.int32, "segmentation_out must have shape"),
    ("get_segmentation", 0, (2, 16, 16, 2), torch.int64, "segmentation_out must have dtype"),
    ("get_hdr", 0, (2, 16, 16, 4), torch.float32, "hdr_out must have shape"),
    ("get_albedo", 1, (2, 12, 24, 3), torch.float32, "does not render rgb"),
    ("get_normal", 0, (2, 16, 16, 3), torch.float16, "normal_out must have dtype"),
  ],
)
def test_readback_argument_errors(fn, cam, shape, dtype, match):
  m, d, rc, _ = _shot()
  out = torch.zeros(shape, dtype=dtype)
  with pytest.raises(ValueError, match=match):
    if fn == "get_depth":
      mj.get_depth(rc, cam, 1.0, out)
    else:
      getattr(mj, fn)(rc, cam, out)


def test_guides_and_hdr_need_their_flags():
  mjm, _ = _two_camera_scene()
  m, d, rc = make(mjm, nworld=1, cam_res=None, render_hdr=False)
  out = torch.zeros((1, 16, 16, 3))
  with pytest.raises(ValueError, match="render_guides=True"):
    mj.get_albedo(rc, 0, out)
  with pytest.raises(ValueError, match="render_guides=True"):
    mj.get_normal(rc, 0, out)
  with pytest.raises(ValueError, match="render_hdr=True"):
    mj.get_hdr(rc, 0, out)


def test_world_count_mismatch_is_refused_before_any_launch():
  """Every entry point that launches refuses data with another world count than the context's
  buffers were sized for (the kernels would write past them)."""
  mjm = scenes.primitives_scene(0, 8)
  m, d, rc = make(mjm, nworld=2, cam_res=(8, 8))
  d3 = mj.put_data(mjm, 3, "cpu")
  with pytest.raises(ValueError, match="data has 3 worlds, the context 2"):
    mj.refit_bvh(m, d3, rc)
  with pytest.raises(ValueError, match="data has 3 worlds, the context 2"):
    mj.render(m, d3, rc)
  m, d, prc = make(mjm, nworld=2, cam_res=(8, 8), tier=mj.Tier.PATH, path_samples=1)
  with pytest.raises(ValueError, match="data has 3 worlds, the context 2"):
    render_path_samples(m, d3, prc)


def test_views_alias_the_context_buffers_and_match_the_copies():
  """The *_view functions return the camera's pixels in place: same values as the get_* copies
  (colour as BGRA bytes of the 0xAARRGGBB words), same storage as the context's buffers, and a
  DLPack export shares it too."""
  m, d, rc, out = _shot()
  rc.hdr_data.copy_(torch.from_numpy(out["hdr"]))
  rc.guide_albedo.copy_(torch.from_numpy(out["albedo"]))
  rc.guide_normal.copy_(torch.from_numpy(out["normal"]))
  w, h = (int(x) for x in rc.base.cam_res[0])
  bgra = mj.bgra_view(rc, 0)
  assert bgra.shape == (2, h, w, 4) and bgra.dtype == torch.uint8
  assert bgra.data_ptr() == rc.base.rgb_data.data_ptr()
  rgb = torch.zeros((2, h, w, 3))
  mj.get_rgb(rc, 0, rgb)
  np.testing.assert_array_equal(bgra[..., [2, 1, 0]].numpy().astype(np.float32) / 255.0, rgb.numpy())
  assert (bgra[..., 3] == 255).all()
  np.testing.assert_array_equal(mj.packed_rgb_view(rc, 0).numpy(), out["rgb"][:, : w * h].reshape(2, h, w).view(np.int32))
  for name in ("hdr", "albedo", "normal"):  # (their get_* copies run a HIP kernel: compared on the GPU)
    view = getattr(mj, f"{name}_view")(rc, 0)
    assert view.shape == (2, h, w, 3)
    np.testing.assert_array_equal(view.numpy(), out[name][:, : w * h].reshape(2, h, w, 3))
  for cam in (0, 1):
    cw_, ch_ = (int(x) for x in rc.base.cam_res[cam])
    adr = int(rc.base.depth_adr[cam])
    depth = mj.depth_view(rc, cam)
    np.testing.assert_array_equal(depth.numpy(), out["depth"][:, adr : adr + cw_ * ch_].reshape(2, ch_, cw_))
    seg = torch.zeros((2, ch_, cw_, 2), dtype=torch.int32)
    mj.get_segmentation(rc, cam, seg)
    np.testing.assert_array_equal(mj.segmentation_view(rc, cam).numpy(), seg.numpy())
  # aliasing: a write into the context shows through the view and through a DLPack export
  depth = mj.depth_view(rc, 1)
  exported = torch.utils.dlpack.from_dlpack(torch.utils.dlpack.to_dlpack(depth))
  rc.base.depth_data[1, int(rc.base.depth_adr[1])] = 123.0
  assert depth[1, 0, 0] == 123.0 and exported[1, 0, 0] == 123.0
  with pytest.raises(ValueError, match="does not render rgb"):
    mj.bgra_view(rc, 1)
  with pytest.raises(ValueError, match="camera_index"):
    mj.depth_view(rc, 2)


@pytest.mark.gpu
def test_gpu_readback_matches_oracle_buffers(cuda):
  """Every get_* on GPU tensors after a GPU render of 2 worlds x 2 cameras (16 x 16 with RGB,
  24 x 12 depth and segmentation only), against the oracle's buffers read with the reference's
  semantics (output.py:34-152)."""
  mjm, _ = _two_camera_scene()
  m, d, rc = make(mjm, nworld=2, device=cuda, cam_res=None, render_rgb=[True, False], render_depth=[True, True],
                  render_seg=[True, True], render_guides=True)
  mj.refit_bvh(m, d, rc)
  mj.render(m, d, rc)
  ref = oracle.rt_render(m, d, rc)
  w0, h0 = 16, 16
  rgb = torch.zeros((2, h0, w0, 3), device=cuda)
  mj.get_rgb(rc, 0, rgb)
  words = ref["rgb"][:, : w0 * h0].reshape(2, h0, w0)
  want = np.stack([((words >> s) & 0xFF) for s in (16, 8, 0)], -1).astype(np.float32) / 255.0
  assert np.abs(rgb.cpu().numpy() - want).max() <= 1.0 / 255.0 + 1e-7
  for name, key in (("get_hdr", "hdr"), ("get_albedo", "albedo"), ("get_normal", "normal")):
    out = torch.zeros((2, h0, w0, 3), device=cuda)
    getattr(mj, name)(rc, 0, out)
    np.testing.assert_allclose(out.cpu().numpy(), ref[key][:, : w0 * h0].reshape(2, h0, w0, 3), atol=1e-3 if key == "hdr" else 1e-5)
  for cam in (0, 1):
    w, h = (int(x) for x in rc.base.cam_res[cam])
    adr = int(rc.base.depth_adr[cam])
    dep = torch.zeros((2, h, w), device=cuda)
    mj.get_depth(rc, cam, 2.0, dep)
    raw = ref["depth"][:, adr : adr + w * h].reshape(2, h, w)
    np.testing.assert_allclose(dep.cpu().numpy(), np.clip(raw / 2.0, 0, 1), rtol=1e-5, atol=1e-7)
    seg = torch.zeros((2, h, w, 2), dtype=torch.int32, device=cuda)
    mj.get_segmentation(rc, cam, seg)
    sadr = int(rc.base.seg_adr[cam])
    np.testing.assert_array_equal(seg.cpu().numpy(), ref["seg"][:, sadr : sadr + w * h].reshape(2, h, w, 2))
    # the zero-copy views of the same render
    np.testing.assert_array_equal(mj.segmentation_view(rc, cam).cpu().numpy(), seg.cpu().numpy())
    np.testing.assert_allclose(mj.depth_view(rc, cam).cpu().numpy(), raw, rtol=1e-5, atol=1e-7)
  # the view's bytes are get_rgb's values x 255 (on the GPU, torch divides by 255 as a multiply by
  # its reciprocal: compare bytes)
  np.testing.assert_array_equal(mj.bgra_view(rc, 0)[..., [2, 1, 0]].cpu().numpy(), np.rint(rgb.cpu().numpy() * 255.0).astype(np.uint8))
  for name in ("hdr", "albedo", "normal"):
    copy = torch.zeros((2, h0, w0, 3), device=cuda)
    getattr(mj, f"get_{name}")(rc, 0, copy)
    torch.testing.assert_close(getattr(mj, f"{name}_view")(rc, 0), copy, rtol=0, atol=0)


@pytest.mark.parametrize("flag", ["render_hdr", "render_guides", "denoise"])
def test_flags_switched_on_after_creation_are_refused(flag):
  """Output buffers are sized when the context is created; switching a flag on afterwards would
  make the kernels write past them, so every entry point refuses it."""
  mjm = scenes.primitives_scene(0, 8)
  kw = {"tier": mj.Tier.PATH, "path_samples": 1} if flag == "denoise" else {}
  m, d, rc = make(mjm, nworld=2, cam_res=(8, 8), render_hdr=False, **kw)
  setattr(rc, flag, True)
  with pytest.raises(ValueError, match=f"rc.{flag} is set but"):
    oracle.rt_render(m, d, rc) if flag != "denoise" else oracle.pt_render(m, d, rc)
  with pytest.raises(ValueError, match=f"rc.{flag} is set but"):
    mj.render(m, d, rc)
