"""Persistent wavefront path tracer over batches of world groups (csrc/wavefront.hip).

A batch is a list of path-traced (model, data, context) groups that share the render settings;
each group has its own worlds, camera resolutions and output buffers, so one launch renders,
say, 2048 worlds at 64x64, 1536 at 128x128 and 512 at 256x256. The images are bit-identical to
`render` of each group with the same seed: only the scheduling differs (dynamic tile queue,
lanes refilled with new paths as old ones end, each tile resolved by its workgroup). render()
itself takes this kernel for path-traced contexts (render_context), as a batch of one group.

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
  waves_per_eu: int = 4  # 4 or 8; at 1 sample a pixel (work items of four tiles) the kernel runs 4
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


class _Prepared:
  """A workspace holding the descriptors of one set of groups (mjpbr_pt_wavefront_prepare) and the
  plan of its launches; `key` is the descriptor bytes it was written from.

  A launch captured into a graph keeps reading its workspace's descriptors at every replay, so a
  captured workspace is never rewritten or freed: the next prepare moves to a new workspace and
  keeps the captured one alive (`retired`) for as long as this object lives."""

  def __init__(self, device, groups_per_cu: int):
    self.device = device
    self.groups_per_cu = groups_per_cu
    self.workspace = None
    self.plan = None
    self.key = None
    self.captured = False
    self.retired = []

  @staticmethod
  def descriptor_key(scenes, outs, settings) -> bytes:
    return bytes(scenes) + bytes(outs) + bytes(settings)

  def workspace_bytes(self, scenes, settings) -> int:
    """Workspace for launches of up to `groups_per_cu` workgroups per CU."""
    need = _abi.library().mjpbr_pt_wavefront_workspace(_abi.byref(scenes), len(scenes), _abi.byref(settings), self.groups_per_cu)
    if need < 0:
      raise ValueError("invalid wavefront batch (groups, cameras or samples)")
    return int(need)

  def prepare(self, scenes, outs, settings, stream: int) -> None:
    """Sizes the workspace for launches of up to `groups_per_cu` workgroups per CU and writes the
    descriptors into it; synchronises `stream` (and the device, before it rewrites or frees a
    workspace that launches on another stream may still read)."""
    n = len(scenes)
    need = self.workspace_bytes(scenes, settings)
    if self.captured:
      self.retired.append(self.workspace)
      self.workspace, self.captured = None, False
    elif self.workspace is not None:
      torch.cuda.synchronize(self.device)
      if self.workspace.numel() < need:
        self.workspace = None  # released before the larger one is allocated
    if self.workspace is None:
      self.workspace = torch.empty(int(need), dtype=torch.uint8, device=self.device)
    plan = _abi.structs()["MjpbrWavefrontPlan"]()
    _abi.call(
      "mjpbr_pt_wavefront_prepare", _abi.byref(scenes), _abi.byref(outs), n, _abi.byref(settings), _abi.ptr(self.workspace),
      int(self.workspace.numel()), _abi.byref(plan), stream,
    )
    self.plan = plan
    self.key = self.descriptor_key(scenes, outs, settings)

  def launch(self, seed: int, config: WavefrontConfig, stream: int) -> None:
    if torch.cuda.is_current_stream_capturing():
      self.captured = True
    cfg = config.to_c()
    _abi.call("mjpbr_pt_wavefront_launch", _abi.ptr(self.workspace), _abi.byref(self.plan), _abi.byref(cfg), ctypes.c_int32(int(seed)), stream)


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
    # sized for the largest grid, so every launch shape the autotuner tries gets its full grid
    self._state = _Prepared(dev, 16)

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

  def _stream(self) -> int:
    return torch.cuda.current_stream(self.device).cuda_stream

  def prepare(self) -> None:
    """(Re)writes the group descriptors into the workspace; synchronises the current stream.
    render() calls it whenever a group's buffers changed (a replaced per-world array)."""
    self._state.prepare(*self._structs(), self._stream())

  def _launch(self, seed: int, config: WavefrontConfig) -> None:
    self._state.launch(seed, config, self._stream())

  def render(self, seed: int = 0, config: WavefrontConfig | None = None) -> None:
    """Renders every world of every group (as `render(m, d, rc, seed)` of each group would)."""
    from mujoco_warp_pbr_amd.render import check_env_id

    for _, _, rc in self.groups:
      check_env_id(rc)
    scenes, outs, settings = self._structs()
    if self._state.key != _Prepared.descriptor_key(scenes, outs, settings):
      self._state.prepare(scenes, outs, settings, self._stream())
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
    """A graph of one render (counter reset, staging of oversized worlds, persistent kernel);
    replay() re-renders with the same seed and buffers."""
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


# Launch shapes of render()'s wavefront launches (a path-traced context as a batch of one group),
# the best of sweeps at 1024 worlds: RENDER_CONFIG on the reference suite's scenes and the
# full-feature scene (4k mesh triangles), MESH_HEAVY_CONFIG on scenes of more than
# MESH_HEAVY_TRIANGLES rendered mesh triangles, whose mesh walks hide their load latency better
# at 8 waves a SIMD (the 200k-triangle scene at 4 spp 128^2: 173.9 ms against 193.4 at 4 waves).
RENDER_CONFIG = WavefrontConfig(waves_per_eu=4, groups_per_cu=4, refill_threshold=32)
MESH_HEAVY_CONFIG = WavefrontConfig(waves_per_eu=8, groups_per_cu=8, refill_threshold=32)
MESH_HEAVY_TRIANGLES = 1 << 16


def render_config(rc) -> WavefrontConfig:
  """The launch shape render() gives the context's wavefront launches."""
  return MESH_HEAVY_CONFIG if int(rc.base.mesh_triangles) > MESH_HEAVY_TRIANGLES else RENDER_CONFIG

# Largest workspace render() gives the wavefront kernel (its sample regions grow with the samples
# a pixel: 192 KiB a workgroup at 16 samples with the denoiser features); beyond it, hundreds of
# samples a pixel, render() launches the per-pixel megakernel.
WORKSPACE_BUDGET = 2 << 30


def render_context(m, d, rc, settings, seed: int, stream: int) -> bool:
  """The path tracer of render(m, d, rc, seed) through the wavefront kernel: the noisy HDR and
  features with `rc.denoise`, else the images (bit-identical to the per-pixel kernel's). The
  workspace and descriptors are kept in the context and rewritten when a buffer changed.
  Returns False, having done nothing, under graph capture before the context's current buffers
  were prepared (preparing synchronises), or when the workspace would exceed WORKSPACE_BUDGET:
  the caller then launches the per-pixel kernel."""
  S = _abi.structs()
  scenes = (S["MjpbrScene"] * 1)()
  outs = (S["MjpbrOutputs"] * 1)()
  scenes[0] = scene_struct(m, d, rc)
  outs[0] = outputs_struct(rc)
  counts = rc._host_cache.get("ray_counts")
  outs[0].ray_counts = None  # the launch counts through its config
  config = render_config(rc)
  state = rc._host_cache.get("wavefront")
  if state is None or state.key != _Prepared.descriptor_key(scenes, outs, settings):
    if torch.cuda.is_current_stream_capturing():
      return False
    fresh = _Prepared(rc.hdr_data.device, config.groups_per_cu)
    if fresh.workspace_bytes(scenes, settings) > WORKSPACE_BUDGET:
      return False
    if state is None or state.device != fresh.device or state.groups_per_cu != fresh.groups_per_cu:
      state = rc._host_cache["wavefront"] = fresh
    state.prepare(scenes, outs, settings, stream)
  state.launch(seed, dataclasses.replace(config, ray_counts=counts), stream)
  return True
