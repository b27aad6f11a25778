"""Autotuner of the persistent wavefront path tracer's launch shape.

Coordinate descent over the kernel's occupancy variant (waves per SIMD), the persistent grid
(workgroups per compute unit) and the lane-refill threshold, each candidate timed with HIP events
on the workload itself (the batch to be rendered, its own seed). The winner is cached per
workload signature (group shapes, samples, depth, device) in a JSON file, so later processes
skip the search.

  from mujoco_warp_pbr_amd import autotune
  config, table = autotune.tune(batch)
"""

from __future__ import annotations

import json
import os
import time
from pathlib import Path

import torch

from mujoco_warp_pbr_amd.wavefront import WavefrontBatch
from mujoco_warp_pbr_amd.wavefront import WavefrontConfig

SPACE = {
  "waves_per_eu": (8, 4),
  "groups_per_cu": (2, 4, 8),
  "refill_threshold": (1, 8, 16, 32, 64),
}
ORDER = ("groups_per_cu", "waves_per_eu", "refill_threshold")


def cache_path() -> Path:
  root = os.environ.get("MJPBR_CACHE_DIR") or os.path.join(os.path.expanduser("~"), ".cache", "mujoco_warp_pbr_amd")
  return Path(root) / "autotune.json"


def signature(batch: WavefrontBatch) -> str:
  """Workload key: per group worlds and camera sizes, samples, depth, lights, device name."""
  parts = []
  for _, _, rc in batch.groups:
    res = ",".join(f"{int(w)}x{int(h)}" for w, h in rc.base.cam_res)
    parts.append(f"{rc.base.nworld}@{res}")
    rc0 = rc
  dev = torch.cuda.get_device_name(batch.device)
  return f"pt-wavefront|{'+'.join(parts)}|spp{rc0.path_samples}|d{rc0.max_depth}|rr{rc0.rr_depth}|{dev}"


def _load_cache() -> dict:
  try:
    return json.loads(cache_path().read_text())
  except (OSError, ValueError):
    return {}


def _store_cache(key: str, cfg: WavefrontConfig) -> None:
  try:
    data = _load_cache()
    data[key] = {k: getattr(cfg, k) for k in ("waves_per_eu", "groups_per_cu", "refill_threshold")}
    p = cache_path()
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(json.dumps(data, indent=1))
  except OSError:
    pass  # a read-only home only costs a re-tune


def time_config(batch: WavefrontBatch, cfg: WavefrontConfig, seed: int = 0, reps: int = 1) -> float:
  """Milliseconds of one batch render with `cfg` (HIP events on the launch stream)."""
  stream = torch.cuda.current_stream(batch.device)
  batch.render(seed, cfg)  # warm (and prepare)
  a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
  a.record(stream)
  for _ in range(reps):
    batch.render(seed, cfg)
  b.record(stream)
  b.synchronize()
  return a.elapsed_time(b) / reps


def tune(batch: WavefrontBatch, seed: int = 0, reps: int = 1, use_cache: bool = True, log=None) -> tuple[WavefrontConfig, list]:
  """(best config, [(config dict, ms), ...]) for `batch`; also sets batch.config."""
  key = signature(batch)
  if use_cache:
    hit = _load_cache().get(key)
    if hit:
      cfg = WavefrontConfig(**hit)
      batch.config = cfg
      return cfg, [("cached", hit)]
  best = WavefrontConfig()
  table = []
  timed: dict[WavefrontConfig, float] = {}

  def measure(cfg):
    if cfg not in timed:
      t0 = time.perf_counter()
      timed[cfg] = time_config(batch, cfg, seed, reps)
      table.append(({k: getattr(cfg, k) for k in ORDER}, round(timed[cfg], 4)))
      if log:
        log(f"autotune {table[-1][0]}: {timed[cfg]:.3f} ms ({time.perf_counter() - t0:.1f} s)")
    return timed[cfg]

  best_ms = measure(best)
  for knob in ORDER:
    for value in SPACE[knob]:
      cfg = WavefrontConfig(**{**{k: getattr(best, k) for k in ORDER}, knob: value})
      ms = measure(cfg)
      if ms < best_ms:
        best, best_ms = cfg, ms
  batch.config = best
  if use_cache:
    _store_cache(key, best)
  return best, table
