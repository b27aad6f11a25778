"""ctypes bindings of the C ABI in include/mjpbr.h (libmjpbr_hip.so).

The ctypes.Structure mirrors are generated from the header's typedefs, so the two cannot drift;
`check_struct_sizes` compares them with the library's own sizeof values.
"""

from __future__ import annotations

import ctypes
import hashlib
import os
import re
from functools import lru_cache
from pathlib import Path

_PACKAGE = Path(__file__).resolve().parent
HEADER = _PACKAGE.parent / "include" / "mjpbr.h"
LIBRARY = _PACKAGE / "libmjpbr_hip.so"

# Struct names in the order of mjpbr_struct_sizes.
STRUCT_ORDER = (
  "MjpbrGeometry",
  "MjpbrMaterials",
  "MjpbrLights",
  "MjpbrCameras",
  "MjpbrEnvironment",
  "MjpbrSettings",
  "MjpbrOutputs",
  "MjpbrScene",
  "MjpbrAtrousParams",
  "MjpbrWavefrontConfig",
  "MjpbrWavefrontPlan",
)

# Exported functions: name -> (restype, argtypes).
_VP = ctypes.c_void_p
FUNCTIONS = {
  "mjpbr_abi_version": (ctypes.c_int, []),
  "mjpbr_header_hash": (ctypes.c_char_p, []),
  "mjpbr_last_error": (ctypes.c_char_p, []),
  "mjpbr_struct_sizes": (ctypes.c_int, [_VP, ctypes.c_int]),
  "mjpbr_scene_bvh": (ctypes.c_int, [_VP, ctypes.c_int, _VP]),
  "mjpbr_rt_render": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int32, _VP]),
  "mjpbr_rt_render_noalloc": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int32, _VP]),
  "mjpbr_rt_kernel_features": (ctypes.c_int, [_VP]),
  "mjpbr_world_stage_floats": (ctypes.c_int64, [_VP]),
  "mjpbr_pt_render": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int32, _VP]),
  "mjpbr_atrous": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
  "mjpbr_dfg_lut": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _VP]),
  "mjpbr_env_prefilter": (
    ctypes.c_int,
    [_VP, _VP, _VP, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP],
  ),
  "mjpbr_extract_rgb_image": (
    ctypes.c_int,
    [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _VP],
  ),
}
_I, _L, _F = ctypes.c_int, ctypes.c_int64, ctypes.c_float
FUNCTIONS.update(
  {
    "mjpbr_unet_conv3x3": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _I, _VP]),
    "mjpbr_unet_conv3x3_pool": (_I, [_VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _VP]),
    "mjpbr_unet_conv3x3_pair": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _VP]),
    "mjpbr_unet_conv3x3_upcat": (_I, [_VP, _VP, _I, _VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _VP]),
    "mjpbr_unet_wgrad_workspace": (_L, [_I, _I, _I, _I, _I, _I]),
    "mjpbr_unet_bias_workspace": (_L, [_I, _I, _I, _I]),
    "mjpbr_unet_conv3x3_wgrad": (_I, [_VP, _VP, _VP, _VP, _I, _VP, _VP, _VP, _I, _I, _I, _I, _I, _VP]),
    "mjpbr_unet_avgpool2": (_I, [_VP, _VP, _I, _I, _I, _I, _I, _VP]),
    "mjpbr_unet_upcat": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _I, _VP]),
    "mjpbr_unet_upcat_low_grad": (_I, [_VP, _VP, _I, _I, _I, _I, _I, _VP]),
    "mjpbr_unet_skip_grad": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _I, _VP]),
    "mjpbr_unet_l1": (_I, [_VP, _VP, _I, _VP, _L, _I, _VP, _VP, _VP]),
    "mjpbr_unet_adam": (_I, [_VP, _VP, _VP, _VP, _L, _F, _F, _F, _F, _I, _VP, _VP, _I, _I, _VP]),
    "mjpbr_unet_weight_layouts": (_I, [_VP, _I, _I, _VP, _VP, _VP]),
    "mjpbr_unet_pack_input": (_I, [_VP, _I, _VP, _VP, _VP, _I, _I, _VP, _I, _I, _I, _I, _I, _VP, _VP]),
    "mjpbr_unet_unpack_output": (_I, [_VP, _VP, _I, _I, _I, _I, _VP, _I, _I, _VP, _I, _VP, _I, _I, _VP, _VP]),
    "mjpbr_unet_guide_stats": (_I, [_VP, _I, _I, _I, _I, _VP, _VP, _VP, _I, _VP]),
    "mjpbr_unet_mfma_probe": (_I, [_VP, _VP, _VP, _VP]),
    "mjpbr_mesh_bvh_workspace": (_L, [_I, _I]),
    "mjpbr_mesh_bvh_build": (_I, [_VP, _VP, _I, _I, _VP, _VP, _VP, _L, _VP]),
    "mjpbr_texture_mips": (_I, [_VP, _I, _I, _I, _VP, _VP, _VP, _I, _VP, _L, _VP]),
    "mjpbr_radiance_mips": (_I, [_VP, _I, _I, _VP, _VP, _I, _VP, _L, _VP]),
    "mjpbr_env_sh9": (_I, [_VP, _I, _I, _VP, _L, _VP, _VP]),
    "mjpbr_env_sampling_tables": (_I, [_VP, _I, _I, _I, _VP, _L, _VP, _VP, _VP, _VP]),
    "mjpbr_pt_wavefront_workspace": (_L, [_VP, _I, _VP, _I]),
    "mjpbr_pt_wavefront_prepare": (_I, [_VP, _VP, _I, _VP, _VP, _L, _VP, _VP]),
    "mjpbr_pt_wavefront_launch": (_I, [_VP, _VP, _VP, ctypes.c_int32, _VP]),
  }
)

_SCALARS = {
  "int32_t": ctypes.c_int32,
  "int64_t": ctypes.c_int64,
  "uint32_t": ctypes.c_uint32,
  "uint16_t": ctypes.c_uint16,
  "uint8_t": ctypes.c_uint8,
  "float": ctypes.c_float,
  "int": ctypes.c_int,
}


def _header_text() -> str:
  text = HEADER.read_text()
  text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
  return re.sub(r"//[^\n]*", "", text)


def _defines(text: str) -> dict[str, int]:
  out = {}
  for name, value in re.findall(r"#define\s+(\w+)\s+\(?(-?\d+)\)?", text):
    out[name] = int(value)
  return out


@lru_cache(maxsize=None)
def structs() -> dict[str, type[ctypes.Structure]]:
  """ctypes mirrors of every typedef struct in mjpbr.h."""
  text = _header_text()
  defines = _defines(text)
  out: dict[str, type[ctypes.Structure]] = {}
  for name, body in re.findall(r"typedef struct (\w+) \{(.*?)\}\s*\w+;", text, flags=re.S):
    fields = []
    for decl in body.split(";"):
      decl = " ".join(decl.split())
      if not decl:
        continue
      decl = decl.replace("const ", "")
      m = re.match(r"(\w+)\s*(\*?)\s*(.*)", decl)
      assert m, decl
      base, star, rest = m.groups()
      for item in rest.split(","):
        item = item.strip()
        ptr = bool(star)
        if item.startswith("*"):
          ptr = True
          item = item[1:].strip()
        dims = [d for d in re.findall(r"\[(\w+)\]", item)]
        fname = item.split("[")[0].strip()
        if ptr:
          ctype = ctypes.c_void_p
        elif base in _SCALARS:
          ctype = _SCALARS[base]
        elif base in out:
          ctype = out[base]
        else:
          raise ValueError(f"unknown type {base} in {name}")
        for d in reversed(dims):
          ctype = ctype * (defines[d] if d in defines else int(d))
        fields.append((fname, ctype))
    out[name] = type(name, (ctypes.Structure,), {"_fields_": fields})
  return out


class MjpbrError(RuntimeError):
  """A nonzero return of a libmjpbr_hip entry point."""


_LIB = None


def library() -> ctypes.CDLL:
  """The loaded libmjpbr_hip.so. Raises when it was not built: there is no fallback."""
  global _LIB
  if _LIB is None:
    path = Path(os.environ.get("MJPBR_LIBRARY", LIBRARY))
    if not path.exists():
      raise MjpbrError(
        f"{path} is missing: build the HIP library first (python -m mujoco_warp_pbr_amd.csrc.build or __graft_entry__.build())"
      )
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in FUNCTIONS.items():
      fn = getattr(lib, name, None)
      if fn is None:
        continue
      fn.restype = res
      fn.argtypes = args
    _check_build(lib, path)
    _LIB = lib
  return _LIB


def _check_build(lib: ctypes.CDLL, path: Path) -> None:
  """Refuses a library built against another version of include/mjpbr.h: its structs would be
  read at other offsets than the ctypes mirrors write (host-side checks, no device call)."""
  want = hashlib.sha256(HEADER.read_bytes()).hexdigest()[:16]
  got = lib.mjpbr_header_hash().decode() if hasattr(lib, "mjpbr_header_hash") else "none"
  if got != want:
    raise MjpbrError(f"{path} was built against another include/mjpbr.h (header hash {got}, expected {want}): rebuild it")
  sizes = (ctypes.c_int64 * len(STRUCT_ORDER))()
  lib.mjpbr_struct_sizes(ctypes.cast(sizes, ctypes.c_void_p), len(STRUCT_ORDER))
  st = structs()
  for i, name in enumerate(STRUCT_ORDER):
    if ctypes.sizeof(st[name]) != int(sizes[i]):
      raise MjpbrError(f"{path}: {name} is {int(sizes[i])} bytes in the library, {ctypes.sizeof(st[name])} in the bindings: rebuild it")


def call(name: str, *args) -> None:
  """Calls an entry point and raises MjpbrError with the library's message on failure."""
  lib = library()
  rc = getattr(lib, name)(*args)
  if rc != 0:
    msg = lib.mjpbr_last_error().decode(errors="replace")
    raise MjpbrError(f"{name} failed ({rc}): {msg}")


def check_struct_sizes() -> dict[str, tuple[int, int]]:
  """{struct: (ctypes size, C size)} for every ABI struct; raises on a mismatch."""
  lib = library()
  sizes = (ctypes.c_int64 * len(STRUCT_ORDER))()
  n = lib.mjpbr_struct_sizes(ctypes.cast(sizes, ctypes.c_void_p), len(STRUCT_ORDER))
  if n != len(STRUCT_ORDER):
    raise MjpbrError(f"library reports {n} structs, bindings know {len(STRUCT_ORDER)}")
  st = structs()
  result = {}
  for i, name in enumerate(STRUCT_ORDER):
    result[name] = (ctypes.sizeof(st[name]), int(sizes[i]))
    if result[name][0] != result[name][1]:
      raise MjpbrError(f"{name}: ctypes size {result[name][0]} != C size {result[name][1]}")
  return result


def ptr(t) -> int:
  """Device (or host) address of a torch tensor, 0 for None."""
  return 0 if t is None else int(t.data_ptr())


def byref(obj) -> ctypes.c_void_p:
  return ctypes.c_void_p(ctypes.addressof(obj))
