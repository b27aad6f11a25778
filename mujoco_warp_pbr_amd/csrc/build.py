"""Builds libmjpbr_hip.so (every .hip source of this folder) for gfx950 with hipcc.

Usage: python -m mujoco_warp_pbr_amd.csrc.build [--jobs N]

Objects go to csrc/_build/, the library to mujoco_warp_pbr_amd/libmjpbr_hip.so (in-tree, so it
travels with the source tree).
"""

from __future__ import annotations

import argparse
import concurrent.futures
import hashlib
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
PACKAGE = HERE.parent
ROOT = PACKAGE.parent
LIB = PACKAGE / "libmjpbr_hip.so"
BUILD = HERE / "_build"

ARCH = os.environ.get("MJPBR_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled: hipcc forms v_pk_* FP32 ops from scalar code otherwise.
CFLAGS = [
  f"--offload-arch={ARCH}",
  "-O3",
  "-fPIC",
  "-std=c++17",
  "-Xclang",
  "-target-feature",
  "-Xclang",
  "-packed-fp32-ops",
  "-Wno-unused-parameter",
  f"-I{ROOT / 'include'}",
]


def header_hash() -> str:
  """First 16 hex digits of the SHA-256 of include/mjpbr.h: compiled into the library so the
  bindings refuse a library built against another version of the ABI structs."""
  return hashlib.sha256((ROOT / "include" / "mjpbr.h").read_bytes()).hexdigest()[:16]


def sources() -> list[Path]:
  return sorted(HERE.glob("*.hip"))


def _compile(src: Path) -> Path:
  obj = BUILD / (src.stem + ".o")
  cmd = [HIPCC, *CFLAGS, f'-DMJPBR_HEADER_HASH="{header_hash()}"', "-c", str(src), "-o", str(obj)]
  proc = subprocess.run(cmd, capture_output=True, text=True)
  if proc.returncode != 0:
    raise RuntimeError(f"hipcc failed on {src.name}:\n{proc.stderr}")
  return obj


def build(jobs: int | None = None, verbose: bool = True) -> Path:
  BUILD.mkdir(exist_ok=True)
  srcs = sources()
  jobs = jobs or min(len(srcs), os.cpu_count() or 1, 8)
  with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
    objs = list(pool.map(_compile, srcs))
  tmp = LIB.with_suffix(".so.tmp")
  cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
  proc = subprocess.run(cmd, capture_output=True, text=True)
  if proc.returncode != 0:
    raise RuntimeError(f"link failed:\n{proc.stderr}")
  os.replace(tmp, LIB)
  if verbose:
    print(f"built {LIB} from {len(srcs)} sources")
  return LIB


def main(argv: list[str] | None = None) -> int:
  p = argparse.ArgumentParser()
  p.add_argument("--jobs", type=int, default=None)
  args = p.parse_args(argv)
  build(args.jobs)
  return 0


if __name__ == "__main__":
  sys.exit(main())
