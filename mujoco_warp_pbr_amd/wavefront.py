"""Persistent wavefront path tracer over batches of world groups (csrc/wavefront.hip).

A batch is a list of path-traced (model, data, context) groups that share the render settings;
each group has its own worlds, camera resolutions and output buffers, so one launch renders,
say, 2048 worlds at 64x64, 1536 at 128x128 and 512 at 256x256. The images are bit-identical to
`render` of each group with the same seed: only the scheduling differs (dynamic tile queue,
lanes refilled with new paths as old ones end, one resolve pass per pixel).

No counterpart in the reference (mujoco_warp_pbr renders one context per `render` call with the
per-pixel megakernel, pathtrace.py:426-438); the API mirrors `render(m, d, rc, seed)`.
"""

from __future__ import annotations

import ctypes
import dataclasses

import torch

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd.render import _max_cams
from mujoco_warp_pbr_amd.render import outputs_struct
from mujoco_warp_pbr_amd.render import scene_struct
from mujoco_warp_pbr_amd.render import settings_struct
from mujoco_warp_pbr_amd.types import Tier

# world groups a batch may hold (MJPBR_WF_MAX_GROUPS of include/mjpbr.h)
MAX_GROUPS = _abi._defines(_abi._header_text())["MJPBR_WF_MAX_GROUPS"]


@dataclasses.dataclass(frozen=True)
class WavefrontConfig:
  """Launch shape of the persistent kernel (see include/mjpbr.h MjpbrWavefrontConfig)."""

  # defaults: the autotuner's pick on 4096 worlds (64^2 / 128^2 / 256^2 mixed), 4 spp, depth 8
  waves_per_eu: int = 4
  groups_per_cu: int = 8
  refill_threshold: int = 32
  item_limit: int = 0
  # int64 device tensor of 3: the launch adds the rays it traced (a counting kernel instance)
  ray_counts: object = dataclasses.field(default=None, compare=False, hash=False)

  def to_c(self):
    c = _abi.structs()["MjpbrWavefrontConfig"]()
    c.waves_per_eu = self.waves_per_eu
    c.groups_per_cu = self.groups_per_cu
    c.refill_threshold = self.refill_threshold
    c.item_limit = self.item_limit
    c.ray_counts = _abi.ptr(self.ray_counts)
    return c


class WavefrontBatch:
  """One persistent launch over several path-traced world groups."""

  def __init__(self, groups, config: WavefrontConfig | None = None):
    groups = list(groups)
    if not 1 <= len(groups) <= MAX_GROUPS:
      raise ValueError(f"a wavefront batch holds 1 to {MAX_GROUPS} world groups, got {len(groups)}")
    for i, (m, d, rc) in enumerate(groups):
      if rc.tier != Tier.PATH:
        raise ValueError(f"group {i}: the wavefront tracer renders path-traced contexts (Tier.PATH)")
      if rc.denoise:
        raise ValueError(f"group {i}: denoising is done by render(); create the context with denoise=False")
      if d.nworld != rc.base.nworld:
        raise ValueError(f"group {i}: data has {d.nworld} worlds, the context {rc.base.nworld}")
      if int(rc.base.nrender) > _max_cams():
        raise ValueError(f"group {i}: a wavefront group renders at most MJPBR_MAX_CAMS cameras, got {int(rc.base.nrender)} (use render())")
    s0 = bytes(settings_struct(groups[0][2]))
    for i, (_, _, rc) in enumerate(groups[1:], 1):
      if bytes(settings_struct(rc)) != s0:
        raise ValueError(f"group {i}: render settings differ from group 0 (samples, depth, lights, features must match)")
    devs = {rc.hdr_data.device for _, _, rc in groups}
    if len(devs) != 1:
      raise ValueError("every group of a batch must live on one device")
    dev = devs.pop()
    if dev.type != "cuda":
      raise ValueError(f"the wavefront tracer runs on a GPU, got {dev}")
    self.groups = groups
    self.device = dev
    self.config = config or WavefrontConfig()
    self._workspace = None
    self._plan = None
    self._key = None

  @property
  def npixels(self) -> int:
    return sum(int(rc.base.nworld) * int(rc.base.total_rays) for _, _, rc in self.groups)

  @property
  def path_samples(self) -> int:
    return int(self.groups[0][2].path_samples)

  def _structs(self):
    S = _abi.structs()
    n = len(self.groups)
    scenes = (S["MjpbrScene"] * n)()
    outs = (S["MjpbrOutputs"] * n)()
    for i, (m, d, rc) in enumerate(self.groups):
      scenes[i] = scene_struct(m, d, rc)
      outs[i] = outputs_struct(rc)
    settings = settings_struct(self.groups[0][2])
    return scenes, outs, settings

  def prepare(self) -> None:
    """(Re)writes the group descriptors into the workspace; synchronises the current stream.
    render() calls it whenever a group's buffers changed (a replaced per-world array)."""
    scenes, outs, settings = self._structs()
    lib = _abi.library()
    n = len(self.groups)
    need = lib.mjpbr_pt_wavefront_workspace(_abi.byref(scenes), n, _abi.byref(settings))
    if need < 0:
      raise ValueError("invalid wavefront batch (groups or samples)")
    if self._workspace is None or self._workspace.numel() < need:
      self._workspace = torch.empty(int(need), dtype=torch.uint8, device=self.device)
    plan = _abi.structs()["MjpbrWavefrontPlan"]()
    stream = torch.cuda.current_stream(self.device).cuda_stream
    _abi.call(
      "mjpbr_pt_wavefront_prepare", _abi.byref(scenes), _abi.byref(outs), n, _abi.byref(settings), _abi.ptr(self._workspace),
      int(self._workspace.numel()), _abi.byref(plan), stream,
    )
    self._plan = plan
    self._key = bytes(scenes) + bytes(outs) + bytes(settings)

  def _launch(self, seed: int, config: WavefrontConfig) -> None:
    cfg = config.to_c()
    stream = torch.cuda.current_stream(self.device).cuda_stream
    _abi.call("mjpbr_pt_wavefront_launch", _abi.ptr(self._workspace), _abi.byref(self._plan), _abi.byref(cfg), ctypes.c_int32(int(seed)), stream)

  def render(self, seed: int = 0, config: WavefrontConfig | None = None) -> None:
    """Renders every world of every group (as `render(m, d, rc, seed)` of each group would)."""
    from mujoco_warp_pbr_amd.render import check_env_id

    for _, _, rc in self.groups:
      check_env_id(rc)
    scenes, outs, settings = self._structs()
    if self._key != bytes(scenes) + bytes(outs) + bytes(settings):
      self.prepare()
    self._launch(seed, config or self.config)

  def count_rays(self, seed: int = 0, config: WavefrontConfig | None = None) -> dict:
    """Renders once with the counting kernel instance (the same images as render()) and returns
    the rays traced: path segments, light and environment shadow rays, and their total."""
    from mujoco_warp_pbr_amd.render import RAY_KINDS

    counts = torch.zeros(3, dtype=torch.int64, device=self.device)
    self.render(seed, dataclasses.replace(config or self.config, ray_counts=counts))
    out = dict(zip(RAY_KINDS, (int(x) for x in counts.cpu().tolist())))
    out["total"] = sum(out[k] for k in RAY_KINDS)
    return out

  def capture(self, seed: int = 0, config: WavefrontConfig | None = None) -> torch.cuda.CUDAGraph:
    """A graph of one render (counter reset, persistent kernel, resolve); replay() re-renders with
    the same seed and buffers."""
    self.render(seed, config)  # prepares and warms up outside the capture
    torch.cuda.synchronize(self.device)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
      self._launch(seed, config or self.config)
    return graph


def render_batch(groups, seed: int = 0, config: WavefrontConfig | None = None) -> WavefrontBatch:
  """One-shot wavefront render of `groups` [(m, d, rc), ...]; returns the batch for reuse."""
  batch = WavefrontBatch(groups, config)
  batch.render(seed)
  return batch
