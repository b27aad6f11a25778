"""Benchmark: real-time tier, full-feature scene, 1024 worlds x 128x128 on one GPU.

Workload (one step = refit_bvh + render of every world): mirror plane (reflectance 0.3), 12
primitives, 4 convex-hull meshes (500-2000 triangles), 256^2 base / ORM / normal textures with
mips, 512x256 procedural HDR sky (split-sum IBL), directional + point light with bulb radius
0.05 and 4 shadow rays each, 2x2 supersampling, per-world roughness / metallic / colour and
per-world jittered cameras. Outputs rgb (packed sRGB8), depth, segmentation. Inputs are
resident in HBM before the timed region.

Prints one JSON line: Mpixels/s, the roofline of the render kernel (algorithmic HBM bytes of
one launch / its average duration measured with HIP events on the launch stream) and the CPU
oracle (scalar C restatement, threads over worlds) timed on a bounded sample of the same
workload on this host.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "Mpixels/s at 1024 worlds×128² RT tier; PT PSNR@equal-time learned vs à-trous"


def log(msg: str) -> None:
  print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(nworld: int, res: int, device):
  import numpy as np
  import torch

  import mujoco_warp_pbr_amd as mj
  from mujoco_warp_pbr_amd import scenes

  mjm = scenes.full_scene(seed=0, res=res)
  m = mj.put_model(mjm, device)
  d = mj.put_data(mjm, nworld, device)
  rc = mj.create_render_context(
    mjm,
    nworld=nworld,
    cam_res=(res, res),
    envmap=scenes.procedural_sky(0, 256, 512),
    samples_per_pixel=2,
    shadow_samples=4,
    ev100=math.log2(1.0 / 1.2),
    device=device,
  )
  scenes.randomize_worlds(m, rc, nworld, seed=1)
  scenes.jitter_cameras(d, seed=2)
  rc.light_bulbradius = torch.full_like(rc.light_bulbradius, 0.05)
  return mjm, m, d, rc


def algorithmic_bytes(mjm, rc, nworld: int, npix: int) -> tuple[float, float]:
  """(bytes one launch must move, bytes per pixel): outputs (rgb 4 + depth 4 + seg 8 per pixel),
  per-world scene data (geom pose 48 B, colour / metallic / roughness 24 B per geom, scene BVH
  48 B per inner node + 4 B per leaf, camera 48 B) and the shared assets once (mesh BVHs,
  textures, environment, DFG LUT)."""
  ngeom = mjm.ngeom
  nleaf = 1 << rc.base.tlas_depth
  per_world = ngeom * (48 + 24) + (nleaf - 1) * 48 + nleaf * 4 + 48
  shared = (
    rc.base.blas_box.numel() * 4
    + rc.base.blas_tri.numel() * 4
    + rc.textures.data.numel()
    + rc.env.radiance.numel() * 2
    + rc.env.prefiltered.numel() * 2
    + rc.dfg_lut.numel() * 4
  )
  total = npix * 16.0 + nworld * per_world + shared
  return total, total / npix


def cpu_baseline(mjm, m, d, rc, target_s: float = 12.0) -> dict:
  """The CPU oracle on a sample of the workload's worlds, one thread per available core."""
  import oracle

  cores = len(os.sched_getaffinity(0))
  res = int(rc.base.cam_res[0][0])
  # size the sample from one world, then run about target_s seconds of work
  t0 = time.perf_counter()
  oracle.rt_render(m, d, rc, seed=0, threads=1, worlds=(0, 1))
  one = time.perf_counter() - t0
  nsample = max(cores, min(d.nworld, int(target_s * cores / max(one, 1e-3))))
  nsample = min(nsample, d.nworld)
  t0 = time.perf_counter()
  oracle.rt_render(m, d, rc, seed=0, threads=cores, worlds=(0, nsample))
  dt = time.perf_counter() - t0
  mpx = nsample * res * res / dt / 1e6
  return {
    "value": round(mpx, 4),
    "unit": "Mpixels/s",
    "cores": cores,
    "kind": "port",
    "sample": f"{nsample} of the {d.nworld} worlds ({res}x{res}, same scene and settings), {dt:.1f} s on {cores} threads",
  }


def main(argv=None) -> int:
  p = argparse.ArgumentParser()
  p.add_argument("--gpus", type=int, default=1)
  p.add_argument("--steps", type=int, default=20)
  p.add_argument("--warmup", type=int, default=3)
  p.add_argument("--nworld", type=int, default=1024)
  p.add_argument("--res", type=int, default=128)
  p.add_argument("--no-cpu-baseline", action="store_true")
  args = p.parse_args(argv)
  if args.gpus != 1:
    print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (worlds shard across GPUs with no exchange)", file=sys.stderr)
    return 2

  import torch

  import mujoco_warp_pbr_amd as mj

  device = torch.device("cuda:0")
  log("building workload")
  mjm, m, d, rc = build_workload(args.nworld, args.res, device)
  npix = args.nworld * args.res * args.res

  def step():
    mj.refit_bvh(m, d, rc)
    mj.render(m, d, rc, seed=0)

  log("warmup")
  for _ in range(args.warmup):
    step()
  torch.cuda.synchronize()
  stream = torch.cuda.current_stream()
  events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
  log(f"timing {args.steps} steps")
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for i in range(args.steps):
    mj.refit_bvh(m, d, rc)
    events[i][0].record(stream)
    mj.render(m, d, rc, seed=0)
    events[i][1].record(stream)
  torch.cuda.synchronize()
  elapsed = time.perf_counter() - t0
  kernel_ms = sum(a.elapsed_time(b) for a, b in events) / args.steps
  ms_per_step = elapsed * 1e3 / args.steps
  value = npix / (ms_per_step * 1e-3) / 1e6
  total_bytes, bpp = algorithmic_bytes(mjm, rc, args.nworld, npix)
  achieved = total_bytes / (kernel_ms * 1e-3) / 1e9
  log(f"{value:.1f} Mpx/s, render kernel {kernel_ms:.3f} ms")

  baseline = None
  if not args.no_cpu_baseline:
    log("cpu baseline (oracle)")
    baseline = cpu_baseline(mjm, m, d, rc)

  line = {
    "metric": METRIC,
    "value": round(value, 2),
    "unit": "Mpixels/s",
    "n_gpus": 1,
    "steps": args.steps,
    "warmup": args.warmup,
    "ms_per_step": round(ms_per_step, 4),
    "higher_is_better": True,
    "scaling": "weak",
    "vs_baseline": None,
    "dtype": "f32",
    "data": "synthetic (seeded procedural scene, random per-world materials and cameras)",
    "config": {
      "workload": f"real-time tier, full-feature scene, {args.nworld} worlds x {args.res}x{args.res}, 2x2 supersampling, 4 soft-shadow rays/light, mirror plane, 4 meshes, textures, HDR sky IBL",
      "nworld": args.nworld,
      "resolution": args.res,
      "kernel_ms": round(kernel_ms, 4),
    },
    "roofline": {
      "bound": "hbm",
      "achieved": round(achieved, 3),
      "peak": HBM_PEAK_GBS,
      "unit": "GB/s",
      "frac": round(achieved / HBM_PEAK_GBS, 6),
      "traffic": None,
      "bytes_per_pixel": round(bpp, 3),
    },
    "cpu_baseline": baseline,
  }
  print(json.dumps(line), flush=True)
  return 0


if __name__ == "__main__":
  sys.exit(main())
