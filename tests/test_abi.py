"""The C ABI library loads (no GPU needed), exports every entry point include/mjpbr.h declares,
and its struct sizes match the generated ctypes mirrors."""

from __future__ import annotations

import ctypes
import re

from mujoco_warp_pbr_amd import _abi


def _declared() -> list[str]:
  text = _abi._header_text()
  return sorted(set(re.findall(r"\b(mjpbr_\w+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
  lib = _abi.library()
  names = _declared()
  assert "mjpbr_rt_render" in names and "mjpbr_scene_bvh" in names
  missing = [n for n in names if not hasattr(lib, n)]
  assert not missing, missing


def test_abi_version_and_struct_sizes():
  lib = _abi.library()
  assert lib.mjpbr_abi_version() == 1
  sizes = _abi.check_struct_sizes()
  assert set(sizes) == set(_abi.STRUCT_ORDER)


def test_errors_are_reported_without_a_gpu():
  """Argument checks run on the host before any launch and fill mjpbr_last_error."""
  lib = _abi.library()
  S = _abi.structs()
  geo = S["MjpbrGeometry"]()
  geo.nbvh = 0
  rc = lib.mjpbr_scene_bvh(ctypes.addressof(geo), 0, None)
  assert rc == -1
  assert b"scene BVH holds" in lib.mjpbr_last_error()


def test_realtime_kernel_instance_covers_the_settings():
  """mjpbr_rt_kernel_features names the compiled real-time instance a launch uses: it covers
  every feature the settings need (mesh 1, supersampling 2, soft shadows 4, mirror 8, layers 16,
  textures 32), runs one lane a sample (128) exactly at 2x2 supersampling, and is the smallest
  such instance for the common cases."""
  lib = _abi.library()
  S = _abi.structs()

  def mask(**kw):
    s = S["MjpbrSettings"]()
    s.samples_per_axis = 1
    s.shadow_samples = 1
    s.use_shadows = 1
    for k, v in kw.items():
      setattr(s, k, v)
    return lib.mjpbr_rt_kernel_features(ctypes.addressof(s))

  assert mask() == 0
  assert mask(textures=1) == 32
  assert mask(textures=1, has_reflection=1) == 40
  assert mask(has_meshes=1) == 1
  assert mask(has_meshes=1, textures=1) == 33
  full = dict(has_meshes=1, samples_per_axis=2, shadow_samples=4, has_reflection=1, textures=1)
  assert mask(**full) == 47 | 128
  assert mask(**full, has_transparency=1) == 63 | 128
  assert mask(**{**full, "samples_per_axis": 3}) == 47
  assert mask(**{**full, "samples_per_axis": 3}, has_transparency=1) == 63
  assert mask(samples_per_axis=2) == 47 | 128  # the smallest one-lane-a-sample instance
  # soft shadows only count with shadows on; any needed bit is covered
  assert mask(shadow_samples=4, use_shadows=0) == 0
  for bits in range(64):
    kw = {}
    if bits & 1: kw["has_meshes"] = 1
    if bits & 2: kw["samples_per_axis"] = 3
    if bits & 4: kw["shadow_samples"] = 2
    if bits & 8: kw["has_reflection"] = 1
    if bits & 16: kw["has_transparency"] = 1
    if bits & 32: kw["textures"] = 1
    assert mask(**kw) & bits == bits
    assert not mask(**kw) & 128  # 3x3 supersampling runs the per-pixel sample loop


def test_library_built_against_another_header_is_refused(tmp_path, monkeypatch):
  """A library whose compiled-in header hash differs from the header the bindings read is
  refused before any call: its structs would be read at other offsets (host-side check)."""
  other = tmp_path / "mjpbr.h"
  other.write_text(_abi.HEADER.read_text() + "\n/* another version */\n")
  monkeypatch.setattr(_abi, "HEADER", other)
  monkeypatch.setattr(_abi, "_LIB", None)
  import pytest

  with pytest.raises(_abi.MjpbrError, match="another include/mjpbr.h"):
    _abi.library()
