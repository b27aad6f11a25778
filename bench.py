"""Benchmark: real-time tier, full-feature scene, 1024 worlds x 128x128 on one GPU.

Workload (one step = refit_bvh + render of every world): mirror plane (reflectance 0.3), 12
primitives, 4 convex-hull meshes (500-2000 triangles), 256^2 base / ORM / normal textures with
mips, 512x256 procedural HDR sky (split-sum IBL), directional + point light with bulb radius
0.05 and 4 shadow rays each, 2x2 supersampling, per-world roughness / metallic / colour and
per-world jittered cameras. Outputs rgb (packed sRGB8), depth, segmentation. Inputs are
resident in HBM before the timed region.

Prints one JSON line: Mpixels/s, the roofline of the render kernel (algorithmic HBM bytes of
one launch / its average duration measured with HIP events on the launch stream; `traffic` from
the committed rocprofv3 counter runs), `valu_roofline` (the kernel's actual limit: its VALU
instructions from the committed SQ_INSTS_VALU run over the measured launch time, against the
GPU's VALU issue rate), the CPU oracle (scalar C restatement, threads over
worlds) timed on a bounded sample of the same workload on this host, `parity` (the sampled
worlds of the timed render against the oracle's images: segmentation bit-exact, depth 1e-5
relative, sRGB8 within 1), `small_case` (16 worlds x 64^2 on the GPU and on every host core the
way the reference's CPU benchmark runs), and `reference_suite` / `reference_suite_path_traced`
(the reference suite's procedural scenes next to its published rates). Path-tracing sections
follow in the same line (skip with --no-pt); their Mrays/s count every ray a render traced
(path segments, light and environment shadow rays: one untimed render through the kernels'
counting instances) over the timed render:
  pt_wavefront      4096 worlds at mixed resolutions (2048 x 64^2, 1536 x 128^2, 512 x 256^2),
                    4 spp, depth 8: the persistent wavefront tracer (autotuned, graph replay)
                    against the per-pixel megakernel, images identical;
  learned_denoiser  128x128 path tracing, 1 spp + learned U-Net against a-trous at the sample
                    count of equal GPU time (timed on 1024 worlds), PSNR against a 4096-spp
                    reference, and the U-Net's matrix-core throughput against the BF16 peak;
  learned_denoiser_mesh_256  the same on the mesh-heavy scene at 256x256, 1024 worlds, with SSIM.
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


TRAFFIC_PROFILE = ROOT / "profiles" / "rt_kernel_full_scene_1024x128_traffic.md"


RT_PROFILE = ROOT / "profiles" / "rt_kernel_full_scene_1024x128.md"


def valu_roofline(insts, kernel_ms: float):
  if insts is None:
    return None
  achieved = insts / (kernel_ms * 1e-3)
  return {"bound": "valu", "instructions_per_launch": insts, "achieved": round(achieved, -6), "peak": VALU_PEAK_WAVE_INSTS,
          "unit": "wave64 VALU instructions/s", "frac": round(achieved / VALU_PEAK_WAVE_INSTS, 4)}
# VALU issue of one MI355X: 256 compute units x 4 SIMDs, a wave64 FP32 instruction every 4 cycles
# of a 16-lane SIMD, at the 2400 MHz peak engine clock (wave instructions per second).
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4


def measured_valu(nworld: int, res: int):
  """VALU wave instructions per launch of the render kernel from the committed rocprofv3
  SQ_INSTS_VALU run, for the workload it was taken on; None for any other shape."""
  if (nworld, res) != (1024, 128) or not RT_PROFILE.exists():
    return None
  for line in RT_PROFILE.read_text().splitlines():
    if line.startswith("valu_instructions_per_launch:"):
      return float(line.split(":", 1)[1])
  return None


def measured_traffic(nworld: int, res: int):
  """HBM bytes per launch of the render kernel from the committed rocprofv3 FETCH_SIZE /
  WRITE_SIZE runs (scaled by a 1 GiB copy kernel's counters), for the workload they were taken
  on; None for any other shape."""
  if (nworld, res) != (1024, 128) or not TRAFFIC_PROFILE.exists():
    return None
  for line in TRAFFIC_PROFILE.read_text().splitlines():
    if line.startswith("traffic_bytes_per_launch:"):
      return int(line.split(":", 1)[1])
  return None


def cpu_baseline(mjm, m, d, rc, target_s: float = 12.0) -> tuple[dict, dict, int]:
  """The CPU oracle on a sample of the workload's worlds, one thread per available core.
  Returns the baseline record, the oracle's output images and the number of sampled worlds."""
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
  images = oracle.rt_render(m, d, rc, seed=0, threads=cores, worlds=(0, nsample))
  dt = time.perf_counter() - t0
  mpx = nsample * res * res / dt / 1e6
  return {
    "value": round(mpx, 4),
    "unit": "Mpixels/s",
    "cores": cores,
    "kind": "port",
    "sample": f"{nsample} of the {d.nworld} worlds ({res}x{res}, same scene and settings), {dt:.1f} s on {cores} threads; "
    "CPU oracle restatement (scalar C, brute-force ray queries, not Warp)",
  }, images, nsample


def full_size_parity(gpu: dict, ref: dict, nworld: int) -> dict:
  """The bench render's first `nworld` worlds against the oracle's images of the same worlds:
  segmentation bit-exact, depth within 1e-5 relative, packed sRGB8 within 1 per channel."""
  import numpy as np

  seg_g, seg_r = gpu["seg"][:nworld], ref["seg"][:nworld]
  dg, dr = gpu["depth"][:nworld].astype(np.float64), ref["depth"][:nworld].astype(np.float64)
  rel = np.abs(dg - dr) / np.maximum(np.abs(dr), 1e-30)
  rel[(dg == 0) & (dr == 0)] = 0.0
  a = gpu["rgb"][:nworld].view(np.uint32)
  b = ref["rgb"][:nworld].view(np.uint32)
  lsb = np.max([np.abs(((a >> sh) & 0xFF).astype(np.int32) - ((b >> sh) & 0xFF).astype(np.int32)) for sh in (16, 8, 0)], axis=0)
  seg_bad = int(np.sum(np.any(seg_g != seg_r, axis=-1)))
  npix = int(seg_g.shape[0] * seg_g.shape[1])
  over = int(np.sum(lsb > 1))
  out = {
    "worlds": nworld,
    "pixels": npix,
    "seg_mismatched_pixels": seg_bad,
    "depth_max_rel": float(rel.max()),
    "rgb_max_lsb": int(lsb.max()),
    "rgb_pixels_over_1_lsb": over,
    "criterion": "segmentation bit-exact, depth <= 1e-5 relative, sRGB8 within 1 on all but 1e-5 of the pixels",
  }
  out["pass"] = seg_bad == 0 and out["depth_max_rel"] <= 1e-5 and over <= 1e-5 * npix
  return out


# The reference's published real-time rates for its suite at 1024 worlds x 128^2 (frames/s x
# 16384 = Mpixels/s): results/suite/suite-v2/README.md:54-55 (RTX PRO 6000 Blackwell), :129-130
# (H200), :204-205 (A100-SXM4-80GB).
SUITE_PUBLISHED_MPX = {
  "material_grid": {"rtx_pro_6000": 9635.4, "h200": 4614.0, "a100": 2718.0},
  "primitives": {"rtx_pro_6000": 5605.0, "h200": 1976.0, "a100": 1201.0},
}


def reference_suite_section(device, nworld: int = 1024, res: int = 128, reps: int = 20) -> dict:
  """The reference suite's two procedural scenes (material_grid, primitives) at its batch shape:
  1024 worlds x 128^2, one camera per world, identical cameras, rgb only, 1 spp, default
  settings (shadows on, PBR neutral tone mapping, default EV100), a 1024 x 512 HDR environment
  (procedural here; the suite's studio HDR file is not shipped). The render is timed with HIP
  events over `reps` back-to-back launches, as the suite times graph replays of renders."""
  import torch

  import mujoco_warp_pbr_amd as mj
  from mujoco_warp_pbr_amd import scenes

  out = {}
  stream = torch.cuda.current_stream()
  for name, fn in (("material_grid", scenes.suite_material_grid), ("primitives", scenes.suite_primitives)):
    mjm = fn(res)
    m = mj.put_model(mjm, device)
    d = mj.put_data(mjm, nworld, device)
    rc = mj.create_render_context(mjm, nworld=nworld, cam_res=(res, res), render_rgb=True, render_depth=False, render_seg=False,
                                  envmap=scenes.procedural_sky(3, 512, 1024), device=device)
    mj.refit_bvh(m, d, rc)
    for _ in range(3):
      mj.render(m, d, rc)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
      mj.render(m, d, rc)
    b.record(stream)
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    mpx = nworld * res * res / ms / 1e3
    pub = SUITE_PUBLISHED_MPX[name]
    log(f"suite {name}: {ms:.3f} ms = {mpx:.0f} Mpx/s (reference H200 {pub['h200']:.0f}, RTX PRO 6000 {pub['rtx_pro_6000']:.0f})")
    out[name] = {
      "ms": round(ms, 4),
      "Mpixels_s": round(mpx, 1),
      "frames_s": round(nworld / ms * 1e3, 1),
      "reference_published_Mpixels_s": pub,
      "vs_h200": round(mpx / pub["h200"], 3),
      "vs_rtx_pro_6000": round(mpx / pub["rtx_pro_6000"], 3),
    }
    del rc, m, d
    torch.cuda.empty_cache()
  out["workload"] = f"reference suite scenes, {nworld} worlds x {res}^2, rgb only, 1 spp, shadows, 1024x512 HDR environment"
  return out


# The reference's published Warp CPU rates (results/cpu.jsonl:201-208; AMD EPYC 7R13, 16 cores /
# 32 threads, real-time tier at 128^2): Mpixels/s on 1 thread and on 32 threads.
WARP_CPU_PUBLISHED_MPX = {
  "material_grid": {"1_thread": 1.44, "32_threads": 27.5},
  "primitives": {"1_thread": 0.67, "32_threads": 12.6},
}


def small_case_section(device, nworld: int = 16, res: int = 64, seconds: float = 3.0) -> dict:
  """The small real-time case (plane + 8 primitives, a directional and a point light, constant grey
  sky, 16 worlds x 64^2, 1 spp): the GPU render timed with HIP events, and the CPU oracle run the
  way the reference's CPU benchmark runs Warp's CPU device (benchmarks/cpu_bench.py:16-46,
  :308-333): one worker thread per available core, each rendering its own copy of the 16 worlds
  back to back for `seconds` after a common start, aggregate frames/s."""
  import ctypes
  import threading

  import torch

  import mujoco_warp_pbr_amd as mj
  import oracle
  from mujoco_warp_pbr_amd import scenes
  from mujoco_warp_pbr_amd.render import outputs_struct
  from mujoco_warp_pbr_amd.render import scene_struct
  from mujoco_warp_pbr_amd.render import settings_struct

  mjm = scenes.primitives_scene(0, res)
  m = mj.put_model(mjm, device)
  d = mj.put_data(mjm, nworld, device)
  rc = mj.create_render_context(mjm, nworld=nworld, cam_res=(res, res), envmap=scenes.constant_sky(), render_rgb=True,
                                render_depth=True, render_seg=True, device=device)
  mj.refit_bvh(m, d, rc)
  for _ in range(5):
    mj.render(m, d, rc)
  torch.cuda.synchronize()
  stream = torch.cuda.current_stream()
  reps = 200
  a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  a.record(stream)
  for _ in range(reps):
    mj.render(m, d, rc)
  b.record(stream)
  b.synchronize()
  gpu_ms = a.elapsed_time(b) / reps

  cores = len(os.sched_getaffinity(0))
  lib = oracle.library()
  workers = []
  for _ in range(cores):
    hm, hd, hrc = oracle.to_host(m, d, rc)
    structs = (scene_struct(hm, hd, hrc), settings_struct(hrc), outputs_struct(hrc))
    workers.append((structs, (hm, hd, hrc)))
  frames = [0] * cores
  start = threading.Barrier(cores + 1)
  stop = [0.0]

  def run(i):
    (sc, st, out), _keep = workers[i]
    args = (ctypes.c_void_p(ctypes.addressof(sc)), ctypes.c_void_p(ctypes.addressof(st)), ctypes.c_void_p(ctypes.addressof(out)))
    start.wait()
    while time.perf_counter() < stop[0]:
      lib.oracle_rt_render(*args, 0, 0, nworld)
      frames[i] += nworld

  pool = [threading.Thread(target=run, args=(i,)) for i in range(cores)]
  for t in pool:
    t.start()
  t0 = time.perf_counter()
  stop[0] = t0 + seconds
  start.wait()
  for t in pool:
    t.join()
  dt = time.perf_counter() - t0
  cpu_fps = sum(frames) / dt
  gpu_fps = nworld / gpu_ms * 1e3
  log(f"small case: GPU {gpu_fps:.0f} frames/s, CPU oracle {cpu_fps:.0f} frames/s on {cores} threads")
  return {
    "workload": f"real-time tier, plane + 8 primitives, 2 lights, constant grey sky, {nworld} worlds x {res}^2, 1 spp, rgb + depth + seg",
    "gpu_ms": round(gpu_ms, 4),
    "gpu_frames_s": round(gpu_fps, 1),
    "gpu_Mpixels_s": round(gpu_fps * res * res / 1e6, 2),
    "cpu_oracle": {
      "label": "CPU oracle restatement (scalar C, brute-force ray queries, not Warp)",
      "threads": cores,
      "frames_s": round(cpu_fps, 1),
      "Mpixels_s": round(cpu_fps * res * res / 1e6, 3),
      "Mpixels_s_per_thread": round(cpu_fps * res * res / 1e6 / cores, 4),
      "seconds": round(dt, 2),
    },
    "gpu_vs_cpu_oracle": round(gpu_fps / cpu_fps, 1),
    "reference_warp_cpu_published_Mpixels_s_128px": WARP_CPU_PUBLISHED_MPX,
  }


# The reference's published path-traced suite rates (16 spp + a-trous, 1024 worlds x 128^2;
# results/suite/suite-v2/README.md:54-55, :129-130, :204-205), frames/s x 16384 = Mpixels/s.
PT_SUITE_PUBLISHED_MPX = {
  "material_grid": {"rtx_pro_6000": 3.6e3 * 16384 / 1e6, "h200": 1.4e3 * 16384 / 1e6, "a100": 1.0e3 * 16384 / 1e6},
  "primitives": {"rtx_pro_6000": 2.9e3 * 16384 / 1e6, "h200": 658 * 16384 / 1e6, "a100": 500 * 16384 / 1e6},
}


def pt_suite_section(device, nworld: int = 1024, res: int = 128, reps: int = 3) -> dict:
  """The reference suite's path-traced setting on its two procedural scenes: 16 paths a pixel,
  depth 4, Russian roulette from 3, a-trous denoising (3 iterations), 1024 worlds x 128^2 (the
  reference suite's `pbr_path`: tier=1, path_samples=16, denoise=true, other settings default)."""
  import torch

  import mujoco_warp_pbr_amd as mj
  from mujoco_warp_pbr_amd import scenes

  out = {}
  stream = torch.cuda.current_stream()
  for name, fn in (("material_grid", scenes.suite_material_grid), ("primitives", scenes.suite_primitives)):
    mjm = fn(res)
    m = mj.put_model(mjm, device)
    d = mj.put_data(mjm, nworld, device)
    rc = mj.create_render_context(mjm, nworld=nworld, cam_res=(res, res), render_rgb=True, render_depth=False, render_seg=False,
                                  envmap=scenes.procedural_sky(3, 512, 1024), tier=mj.Tier.PATH, path_samples=16, denoise=True,
                                  device=device)
    mj.refit_bvh(m, d, rc)

    def timed():
      mj.render(m, d, rc)
      torch.cuda.synchronize()
      a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
      a.record(stream)
      for _ in range(reps):
        mj.render(m, d, rc)
      b.record(stream)
      b.synchronize()
      return a.elapsed_time(b) / reps

    rc.path_kernel = "megakernel"  # the per-pixel kernel (the reference's scheme), for comparison
    mega_ms = timed()
    mega_rgb = rc.base.rgb_data.clone()
    rc.path_kernel = "wavefront"  # render()'s default
    ms = timed()
    identical = bool(torch.equal(mega_rgb, rc.base.rgb_data))
    del mega_rgb
    mpx = nworld * res * res / ms / 1e3
    rays = mj.count_path_rays(m, d, rc)  # one untimed render through the counting kernel instance
    pub = PT_SUITE_PUBLISHED_MPX[name]
    log(f"suite PT {name}: {ms:.1f} ms = {mpx:.1f} Mpx/s (megakernel {mega_ms:.1f} ms, identical: {identical}), "
        f"{rays['total'] / ms / 1e3:.0f} Mrays/s (reference RTX PRO 6000 {pub['rtx_pro_6000']:.1f} Mpx/s)")
    out[name] = {
      "ms": round(ms, 3),
      "megakernel_ms": round(mega_ms, 3),
      "identical_to_megakernel": identical,
      "Mpixels_s": round(mpx, 2),
      "Mpaths_s": round(mpx * 16, 1),
      "rays_per_render": rays,
      "Mrays_s": round(rays["total"] / ms / 1e3, 1),
      "reference_published_Mpixels_s": {k: round(v, 2) for k, v in pub.items()},
      "vs_rtx_pro_6000": round(mpx / pub["rtx_pro_6000"], 3),
      "vs_h200": round(mpx / pub["h200"], 3),
    }
    del rc, m, d
    torch.cuda.empty_cache()
  out["workload"] = (f"reference suite scenes, path-traced tier, {nworld} worlds x {res}^2, 16 spp, depth 4, a-trous (3 iterations), rgb; "
                     "render() through the wavefront kernel, the per-pixel megakernel timed beside it")
  return out


C2_GROUPS = ((2048, 64), (1536, 128), (512, 256))


def pt_wavefront_section(device) -> dict:
  """Batched path tracing: 4096 worlds of the full-feature scene, 2048 at 64x64, 1536 at 128x128
  and 512 at 256x256, 4 samples, depth 8, Russian roulette from 3, per-world materials, cameras,
  sky rotation and light intensities. The persistent wavefront tracer (one launch over the three
  groups, autotuned shape, graph replay) against the per-pixel megakernel (one render() a group,
  the contexts created with path_kernel="megakernel"); images must be identical."""
  import numpy as np
  import torch

  import mujoco_warp_pbr_amd as mj
  from mujoco_warp_pbr_amd import autotune
  from mujoco_warp_pbr_amd import scenes
  from mujoco_warp_pbr_amd.wavefront import WavefrontBatch

  groups = []
  for i, (nworld, res) in enumerate(C2_GROUPS):
    mjm = scenes.full_scene(seed=0, res=res)
    m = mj.put_model(mjm, device)
    d = mj.put_data(mjm, nworld, device)
    rc = mj.create_render_context(
      mjm, nworld=nworld, cam_res=(res, res), envmap=scenes.procedural_sky(0, 256, 512), tier=mj.Tier.PATH, path_samples=4,
      max_depth=8, rr_depth=3, ev100=math.log2(1.0 / 1.2), device=device, path_kernel="megakernel",
    )
    scenes.randomize_worlds(m, rc, nworld, seed=10 * i + 1)
    scenes.jitter_cameras(d, seed=10 * i + 2)
    rng = np.random.default_rng(10 * i + 3)
    rc.env_yaw = torch.as_tensor(rng.uniform(0, 2 * math.pi, nworld).astype(np.float32), device=device)
    li = rc.light_intensity.cpu().numpy()
    li = np.broadcast_to(li[:1], (nworld, li.shape[1])) * np.exp(rng.normal(0.0, 0.5, (nworld, li.shape[1])))
    rc.light_intensity = torch.as_tensor(li.astype(np.float32), device=device)
    mj.refit_bvh(m, d, rc)
    groups.append((m, d, rc))
  npix = sum(rc.base.nworld * rc.base.total_rays for _, _, rc in groups)
  paths = npix * 4
  stream = torch.cuda.current_stream()

  def timed(fn, reps=3):
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
      fn()
    b.record(stream)
    b.synchronize()
    return a.elapsed_time(b) / reps

  mega_ms = timed(lambda: [mj.render(m, d, rc, seed=3) for m, d, rc in groups])
  ref = [rc.base.rgb_data.clone() for _, _, rc in groups]
  batch = WavefrontBatch(groups)
  cfg, table = autotune.tune(batch, seed=3, log=log)
  graph = batch.capture(seed=3, config=cfg)
  wf_ms = timed(graph.replay)
  identical = all(torch.equal(r, rc.base.rgb_data) for r, (_, _, rc) in zip(ref, groups))
  rays = batch.count_rays(seed=3, config=cfg)  # untimed; the megakernel traces the same rays
  log(f"pt wavefront {wf_ms:.1f} ms vs megakernel {mega_ms:.1f} ms, identical images: {identical}; "
      f"{rays['total'] / 1e6:.0f} M rays: {rays['total'] / wf_ms / 1e3:.0f} vs {rays['total'] / mega_ms / 1e3:.0f} Mrays/s")
  return {
    "workload": "path tracing, full-feature scene, 4096 worlds: 2048 x 64^2 + 1536 x 128^2 + 512 x 256^2, 4 spp, depth 8, RR from 3",
    "pixels": npix,
    "paths": paths,
    "megakernel_ms": round(mega_ms, 3),
    "megakernel_Mpaths_s": round(paths / mega_ms / 1e3, 1),
    "wavefront_ms": round(wf_ms, 3),
    "wavefront_Mpaths_s": round(paths / wf_ms / 1e3, 1),
    "wavefront_Mpixels_s": round(npix / wf_ms / 1e3, 1),
    "speedup": round(mega_ms / wf_ms, 3),
    "rays_per_render": rays,
    "wavefront_Mrays_s": round(rays["total"] / wf_ms / 1e3, 1),
    "megakernel_Mrays_s": round(rays["total"] / mega_ms / 1e3, 1),
    "identical_to_megakernel": identical,
    "config": {k: getattr(cfg, k) for k in ("waves_per_eu", "groups_per_cu", "refill_threshold")},
    "autotune_table": table,
  }


def learned_section(device, time_worlds: int = 1024, quality_worlds: int = 128, ref_spp: int = 4096) -> dict:
  """Path tracing at 128x128, 1 sample + learned denoiser against a-trous (5 iterations) at k
  samples, on a held-out scene seed; GPU time of each pipeline on `time_worlds` worlds, PSNR on
  `quality_worlds` worlds against a `ref_spp` reference. Equal time: the smallest k whose a-trous
  pipeline is no faster than the learned one."""
  from mujoco_warp_pbr_amd import train_denoiser as td

  res = td.evaluate([900], quality_worlds, 128, ref_spp, device, atrous_spp=(1, 2, 3, 4, 6, 8), time_worlds=time_worlds)
  log(f"learned {res['learned_psnr']:.2f} dB vs a-trous@{res['equal_time_atrous_spp']} {res['equal_time_atrous_psnr']:.2f} dB "
      f"at {res['learned_ms']:.1f} / {res['equal_time_atrous_ms']:.1f} ms")
  peak = 2500.0
  return {
    "workload": f"path tracing 128x128, full-feature scene (held-out seed), time on {time_worlds} worlds, PSNR on {quality_worlds} worlds vs {ref_spp}-spp reference",
    "learned_psnr_db": round(res["learned_psnr"], 3),
    "equal_time_atrous_psnr_db": round(res["equal_time_atrous_psnr"], 3),
    "gain_db": round(res["gain_db"], 3),
    "learned_ms": round(res["learned_ms"], 3),
    "equal_time_atrous_spp": res["equal_time_atrous_spp"],
    "equal_time_atrous_ms": round(res["equal_time_atrous_ms"], 3),
    "noisy_1spp_psnr_db": round(res["noisy_1spp_psnr"], 3),
    "atrous": res["atrous"],
    "fallback_worlds": res["fallback_worlds"],
    "unet_roofline": {
      "bound": "mfma",
      "achieved": round(res["unet_tflops_algorithmic"], 2),
      "achieved_as_executed": round(res["unet_tflops"], 2),
      "peak": peak,
      "unit": "TFLOP/s",
      "frac": round(res["unet_tflops_algorithmic"] / peak, 4),
      "ms": round(res["unet_ms"], 3),
    },
  }


def learned_mesh_section(device, time_worlds: int = 1024, quality_worlds: int = 16, ref_spp: int = 4096) -> dict:
  """Path tracing the mesh-heavy scene (64 instances of 16 procedural meshes, ~200k
  triangles, textures, HDR sky) at 256x256, 1 spp + learned U-Net against a-trous (5 iterations)
  at k spp; times on `time_worlds` worlds, PSNR / SSIM on `quality_worlds` worlds against a
  `ref_spp` reference."""
  from mujoco_warp_pbr_amd import train_denoiser as td

  res = td.evaluate([950], quality_worlds, 256, ref_spp, device, atrous_spp=(1, 2, 3, 4), time_worlds=time_worlds, kind="mesh")
  log(f"mesh-heavy 256^2: learned {res['learned_psnr']:.2f} dB vs a-trous@{res['equal_time_atrous_spp']} "
      f"{res['equal_time_atrous_psnr']:.2f} dB at {res['learned_ms']:.1f} / {res['equal_time_atrous_ms']:.1f} ms")
  return {
    "workload": f"path tracing the mesh-heavy scene at 256x256 (held-out seed), time on {time_worlds} worlds, PSNR/SSIM on {quality_worlds} worlds vs {ref_spp}-spp reference",
    "learned_psnr_db": round(res["learned_psnr"], 3),
    "learned_ssim": round(res["learned_ssim"], 4),
    "equal_time_atrous_psnr_db": round(res["equal_time_atrous_psnr"], 3),
    "equal_time_atrous_ssim": round(res["equal_time_atrous_ssim"], 4),
    "gain_db": round(res["gain_db"], 3),
    "learned_ms": round(res["learned_ms"], 3),
    "equal_time_atrous_spp": res["equal_time_atrous_spp"],
    "equal_time_atrous_ms": round(res["equal_time_atrous_ms"], 3),
    "atrous": res["atrous"],
    "fallback_worlds": res["fallback_worlds"],
    "unet_ms": round(res["unet_ms"], 3),
    "unet_tflops": round(res["unet_tflops"], 2),
  }


def kernels_only(nworld: int, res: int) -> int:
  """The U-Net forward (packaged weights, random BF16 input) and a 1-spp path-traced render of the
  full-feature scene with a-trous denoising (3 iterations), nworld x res^2, three times each;
  prints one JSON line of mean times."""
  import torch

  import mujoco_warp_pbr_amd as mj
  from mujoco_warp_pbr_amd import learned
  from mujoco_warp_pbr_amd import scenes

  device = torch.device("cuda:0")
  stream = torch.cuda.current_stream()

  def timed(fn, reps=3):
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
      fn()
    b.record(stream)
    b.synchronize()
    return a.elapsed_time(b) / reps

  net, _ = learned.load(learned.DEFAULT_WEIGHTS, device)
  x = torch.randint(-16000, 16000, (nworld, res, res, learned.IN_CHANNELS), dtype=torch.int16, device=device)
  unet_ms = timed(lambda: net.forward(x))
  del x
  net.release()
  mjm = scenes.full_scene(seed=0, res=res)
  m, d = mj.put_model(mjm, device), mj.put_data(mjm, nworld, device)
  rc = mj.create_render_context(mjm, nworld=nworld, cam_res=(res, res), envmap=scenes.procedural_sky(0, 256, 512), tier=mj.Tier.PATH,
                                path_samples=1, denoise=True, device=device)
  scenes.randomize_worlds(m, rc, nworld, seed=1)
  mj.refit_bvh(m, d, rc)
  pt_ms = timed(lambda: mj.render(m, d, rc))
  print(json.dumps({"kernels_only": f"{nworld} worlds x {res}^2", "unet_forward_ms": round(unet_ms, 3),
                    "pt_1spp_atrous_ms": round(pt_ms, 3)}), flush=True)
  return 0


def main(argv=None) -> int:
  p = argparse.ArgumentParser()
  p.add_argument("--gpus", type=int, default=1)
  p.add_argument("--steps", type=int, default=20)
  p.add_argument("--warmup", type=int, default=3)
  p.add_argument("--nworld", type=int, default=1024)
  p.add_argument("--res", type=int, default=128)
  p.add_argument("--no-cpu-baseline", action="store_true")
  p.add_argument("--no-pt", action="store_true", help="skip the path-tracing sections (wavefront batch, learned denoiser)")
  p.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                 help="only copy BYTES between two device buffers 3 times (the reference kernel that scales HBM counters)")
  p.add_argument("--kernels-only", action="store_true",
                 help="only the U-Net forward and a 1-spp path-traced render with a-trous, --nworld x --res^2, 3 times each "
                      "(the workloads the profiles of those kernels are taken on)")
  args = p.parse_args(argv)
  if args.kernels_only:
    return kernels_only(args.nworld, args.res)
  if args.copy_calibration:
    import torch

    src = torch.ones(args.copy_calibration, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
      dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copied {args.copy_calibration} bytes 3 times")
    return 0
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

  baseline = parity = configs0 = None
  if not args.no_cpu_baseline:
    gpu_images = {
      "rgb": rc.base.rgb_data.cpu().numpy(),
      "depth": rc.base.depth_data.cpu().numpy(),
      "seg": rc.base.seg_data.cpu().numpy(),
    }
    log("cpu baseline (oracle)")
    baseline, ref_images, nsample = cpu_baseline(mjm, m, d, rc)
    parity = full_size_parity(gpu_images, ref_images, nsample)
    log(f"full-size parity on {nsample} worlds: {parity}")
    del ref_images, gpu_images

  del rc, m, d
  torch.cuda.empty_cache()
  if not args.no_cpu_baseline:
    log("small real-time case (16 worlds x 64^2): GPU and CPU oracle")
    configs0 = small_case_section(device)
  log("reference suite scenes (material_grid, primitives)")
  suite = reference_suite_section(device)
  pt_suite = None
  if not args.no_pt:
    log("reference suite scenes, path-traced tier (16 spp + a-trous)")
    pt_suite = pt_suite_section(device)

  pt_wavefront = learned_denoiser = learned_denoiser_mesh = None
  if not args.no_pt:
    log("path tracing: wavefront batch (4096 mixed-resolution worlds)")
    pt_wavefront = pt_wavefront_section(device)
    torch.cuda.empty_cache()
    log("path tracing: learned denoiser vs a-trous at equal time")
    learned_denoiser = learned_section(device)
    torch.cuda.empty_cache()
    log("path tracing: learned denoiser vs a-trous, mesh-heavy scene at 256x256")
    learned_denoiser_mesh = learned_mesh_section(device)

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
      "traffic": measured_traffic(args.nworld, args.res),
      "bytes_per_pixel": round(bpp, 3),
    },
    # the kernel's actual limit: VALU issue (SQ_INSTS_VALU of the committed profile over the
    # measured launch time, against one wave64 instruction every 4 cycles on every SIMD)
    "valu_roofline": valu_roofline(measured_valu(args.nworld, args.res), kernel_ms),
    "cpu_baseline": baseline,
    "parity": parity,
    "small_case": configs0,
    "reference_suite": suite,
    "reference_suite_path_traced": pt_suite,
    "pt_wavefront": pt_wavefront,
    "learned_denoiser": learned_denoiser,
    "learned_denoiser_mesh_256": learned_denoiser_mesh,
  }
  print(json.dumps(line), flush=True)
  return 0


if __name__ == "__main__":
  sys.exit(main())
