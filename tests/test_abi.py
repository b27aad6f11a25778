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
