"""Learned denoiser of the path-traced tier: a 3-level U-Net (32 / 64 / 128 channels, 3x3
convolutions, ReLU, 2x2 average pooling, nearest upsampling with skip concatenation) whose
forward and backward passes run in the hand-written matrix-core kernels of csrc/unet.hip (BF16
activations and weights, FP32 accumulation, FP32 master weights with Adam).

Input (10 channels, padded to 32): log1p of the exposed 1-spp radiance, first-hit albedo,
normal, depth / (1 + depth). The network predicts a residual in log space: denoised =
expm1(out + log1p(noisy)). Training pairs come from the renderer itself (low- and high-sample
renders of the same seeded scenes, see train_denoiser.py). The a-trous filter stays as the
fallback: worlds whose guide statistics leave the range seen in training are filtered by it.

The reference has no counterpart (its only denoiser is the a-trous filter,
mujoco_warp_pbr/_src/denoise.py:204-248, selected with `denoise=True`); this module is selected
with `create_render_context(..., denoise=True, denoiser="learned")`.
"""

from __future__ import annotations

import dataclasses
import json
import math
from pathlib import Path

import numpy as np
import torch

from mujoco_warp_pbr_amd import _abi

IN_CHANNELS = 32  # 10 used
OUT_CHANNELS = 32  # 3 used
USED_IN = 10
NSTATS = 4
# (name, cin, cout, relu, level)
LAYERS = (
  ("e0a", IN_CHANNELS, 32, True, 0),
  ("e0b", 32, 32, True, 0),
  ("e1a", 32, 64, True, 1),
  ("e1b", 64, 64, True, 1),
  ("b0", 64, 128, True, 2),
  ("b1", 128, 128, True, 2),
  ("d1a", 192, 64, True, 1),
  ("d1b", 64, 64, True, 1),
  ("d0a", 96, 32, True, 0),
  ("d0b", 32, 32, True, 0),
  ("out", 32, OUT_CHANNELS, False, 0),
)
# Image sides must be multiples of this (16x16 convolution blocks at the quarter resolution).
SIZE_MULTIPLE = 64
DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "weights" / "unet_denoiser.safetensors"


def flops_per_pixel(padded: bool = True) -> float:
  """Forward multiply-adds x 2 per full-resolution pixel: as executed (channel padding included)
  or algorithmic (the 10 used input and 3 used output channels)."""
  total = 0.0
  for name, cin, cout, _, level in LAYERS:
    if not padded:
      cin = USED_IN if name == "e0a" else cin
      cout = 3 if name == "out" else cout
    total += 2.0 * 9 * cin * cout / (4**level)
  return total


def _i16(shape, device):
  return torch.zeros(shape, dtype=torch.int16, device=device)


def _stream(t: torch.Tensor) -> int:
  return int(torch.cuda.current_stream(t.device).cuda_stream) if t.is_cuda else 0


@dataclasses.dataclass
class Param:
  cin: int
  cout: int
  relu: bool
  w: torch.Tensor  # FP32 [cout][9][cin]
  b: torch.Tensor  # FP32 [cout]
  wf: torch.Tensor  # BF16 bits [cout][9][cin]
  wd: torch.Tensor  # BF16 bits [cin][9][cout], taps mirrored
  mw: torch.Tensor
  vw: torch.Tensor
  mb: torch.Tensor
  vb: torch.Tensor
  gw: torch.Tensor
  gb: torch.Tensor


class UNet:
  """Parameters, activations and gradients of the U-Net on one device."""

  def __init__(self, device="cuda", seed: int = 0):
    self.device = torch.device(device)
    gen = torch.Generator().manual_seed(seed)
    self.params: dict[str, Param] = {}
    dev = self.device
    for name, cin, cout, relu, _ in LAYERS:
      fan_in = 9 * (USED_IN if name == "e0a" else cin)
      std = math.sqrt(2.0 / fan_in) if relu else 0.1 * math.sqrt(1.0 / fan_in)
      w = (torch.randn((cout, 9, cin), generator=gen) * std).float()
      if name == "e0a":
        w[:, :, USED_IN:] = 0.0
      z = lambda *s: torch.zeros(s, dtype=torch.float32, device=dev)  # noqa: E731
      self.params[name] = Param(
        cin, cout, relu, w.to(dev), z(cout), _i16((cout, 9, cin), dev), _i16((cin, 9, cout), dev),
        z(cout, 9, cin), z(cout, 9, cin), z(cout), z(cout), z(cout, 9, cin), z(cout),
      )
    self.step_count = 0
    self._bufs: dict[tuple, dict] = {}
    self._grads: dict[tuple, dict] = {}
    self._ws: dict[str, torch.Tensor] = {}
    self._shape = None
    self.refresh_layouts()

  # ----------------------------------------------------------------- weights
  def refresh_layouts(self) -> None:
    """BF16 forward / data-gradient layouts from the FP32 master weights."""
    for p in self.params.values():
      _abi.call("mjpbr_unet_weight_layouts", _abi.ptr(p.w), p.cin, p.cout, _abi.ptr(p.wf), _abi.ptr(p.wd), _stream(p.w))

  def state_dict(self) -> dict[str, torch.Tensor]:
    out = {}
    for name, p in self.params.items():
      out[f"{name}.w"] = p.w.detach().cpu().contiguous()
      out[f"{name}.b"] = p.b.detach().cpu().contiguous()
    return out

  def load_state_dict(self, state: dict[str, torch.Tensor]) -> None:
    for name, p in self.params.items():
      w, b = state[f"{name}.w"], state[f"{name}.b"]
      if tuple(w.shape) != tuple(p.w.shape) or tuple(b.shape) != tuple(p.b.shape):
        raise ValueError(f"{name}: weights of shape {tuple(w.shape)} / {tuple(b.shape)}, expected {tuple(p.w.shape)} / {tuple(p.b.shape)}")
      p.w.copy_(w.to(p.w.device, torch.float32))
      p.b.copy_(b.to(p.b.device, torch.float32))
    self.refresh_layouts()

  def parameter_count(self) -> int:
    return sum(p.w.numel() + p.b.numel() for p in self.params.values())

  # ----------------------------------------------------------------- buffers
  def _alloc(self, n: int, h: int, w: int) -> dict:
    key = (n, h, w)
    bufs = self._bufs.get(key)
    if bufs is None:
      if h % SIZE_MULTIPLE or w % SIZE_MULTIPLE:
        raise ValueError(f"image size must be a multiple of {SIZE_MULTIPLE} (three levels of 16x16 blocks), got {h}x{w}")
      dev = self.device
      sizes = {0: (n, h, w), 1: (n, h // 2, w // 2), 2: (n, h // 4, w // 4)}
      act = {name: _i16((*sizes[level], cout), dev) for name, _, cout, _, level in LAYERS}
      act["p0"] = _i16((*sizes[1], 32), dev)
      act["p1"] = _i16((*sizes[2], 64), dev)
      act["u1"] = _i16((*sizes[1], 192), dev)
      act["u0"] = _i16((*sizes[0], 96), dev)
      bufs = {
        "act": act,
        "pred": torch.zeros((n * h * w, 3), dtype=torch.float32, device=dev),
        "loss_part": torch.zeros(((n * h * w + 255) // 256,), dtype=torch.float32, device=dev),
      }
      self._bufs[key] = bufs
    self._shape = key
    return bufs

  def _grad_bufs(self) -> dict:
    key = self._shape
    g = self._grads.get(key)
    if g is None:
      g = {k: torch.zeros_like(v) for k, v in self._bufs[key]["act"].items()}
      self._grads[key] = g
    return g

  def release(self) -> None:
    """Frees the cached activation and gradient buffers."""
    self._bufs.clear()
    self._grads.clear()
    self._ws.clear()

  def _workspace(self, key, numel):
    t = self._ws.get(key)
    if t is None or t.numel() < numel:
      t = torch.zeros(max(int(numel), 1), dtype=torch.float32, device=self.device)
      self._ws[key] = t
    return t

  # ----------------------------------------------------------------- forward
  def _conv(self, name, x, out, y32=None, out_f32=0, cout=None):
    """Convolution `name`; `cout` below the layer's computes only its first output channels (the
    weight rows are [cout][9][cin], so a prefix is a smaller layer); out=None writes only y32."""
    p = self.params[name]
    n, h, w = x.shape[0], x.shape[1], x.shape[2]
    _abi.call(
      "mjpbr_unet_conv3x3", _abi.ptr(x), 0, _abi.ptr(p.wf), _abi.ptr(p.b), _abi.ptr(out), _abi.ptr(y32), out_f32, n,
      h, w, p.cin, p.cout if cout is None else cout, int(p.relu), _stream(x),
    )

  def _conv_pool(self, name, x, out, pooled):
    """Convolution `name` and the 2x2 average pooling of its output, one launch."""
    p = self.params[name]
    n, h, w = x.shape[0], x.shape[1], x.shape[2]
    _abi.call(
      "mjpbr_unet_conv3x3_pool", _abi.ptr(x), _abi.ptr(p.wf), _abi.ptr(p.b), _abi.ptr(out), _abi.ptr(pooled), n, h, w, p.cin, p.cout,
      int(p.relu), _stream(x),
    )

  def _conv_pair(self, first, second, x, y32, out_f32):
    """Layers `first` (32 -> 32, ReLU) and `second` (its first 16 output channels, FP32 into
    y32) in one launch: the intermediate activation stays on chip."""
    p1, p2 = self.params[first], self.params[second]
    n, h, w = x.shape[0], x.shape[1], x.shape[2]
    _abi.call(
      "mjpbr_unet_conv3x3_pair", _abi.ptr(x), _abi.ptr(p1.wf), _abi.ptr(p1.b), _abi.ptr(p2.wf), _abi.ptr(p2.b), _abi.ptr(y32), out_f32, n,
      h, w, _stream(x),
    )

  def _conv_upcat(self, name, low, skip, out, c_split):
    p = self.params[name]
    n, h, w = skip.shape[0], skip.shape[1], skip.shape[2]
    _abi.call(
      "mjpbr_unet_conv3x3_upcat", _abi.ptr(low), _abi.ptr(skip), c_split, _abi.ptr(p.wf), _abi.ptr(p.b), _abi.ptr(out), n, h, w, p.cin,
      p.cout, int(p.relu), _stream(skip),
    )

  def forward(self, x: torch.Tensor, training: bool = False) -> torch.Tensor:
    """x: BF16 bits (n, h, w, 32). Returns the FP32 log-space residual (n * h * w, 3), a buffer
    the next forward of the same shape overwrites. Inference reads the decoders' concatenated
    inputs straight from their two sources; `training` materialises them for the weight
    gradients of backward()."""
    n, h, w, c = x.shape
    if c != IN_CHANNELS or x.dtype != torch.int16 or not x.is_contiguous():
      raise ValueError(f"input must be contiguous int16 (BF16 bits) with {IN_CHANNELS} channels, got {x.dtype} {tuple(x.shape)}")
    bufs = self._alloc(n, h, w)
    self.x = x
    self._trained_forward = training
    A = bufs["act"]
    st = _stream(x)
    self._conv("e0a", x, A["e0a"])
    self._conv_pool("e0b", A["e0a"], A["e0b"], A["p0"])
    self._conv("e1a", A["p0"], A["e1a"])
    self._conv_pool("e1b", A["e1a"], A["e1b"], A["p1"])
    self._conv("b0", A["p1"], A["b0"])
    self._conv("b1", A["b0"], A["b1"])
    if training:
      _abi.call("mjpbr_unet_upcat", _abi.ptr(A["b1"]), _abi.ptr(A["e1b"]), _abi.ptr(A["u1"]), n, h // 2, w // 2, 128, 64, st)
      self._conv("d1a", A["u1"], A["d1a"])
    else:
      self._conv_upcat("d1a", A["b1"], A["e1b"], A["d1a"], 128)
    self._conv("d1b", A["d1a"], A["d1b"])
    if training:
      _abi.call("mjpbr_unet_upcat", _abi.ptr(A["d1b"]), _abi.ptr(A["e0b"]), _abi.ptr(A["u0"]), n, h, w, 64, 32, st)
      self._conv("d0a", A["u0"], A["d0a"])
    else:
      self._conv_upcat("d0a", A["d1b"], A["e0b"], A["d0a"], 64)
    if training:
      self._conv("d0b", A["d0a"], A["d0b"])
      self._conv("out", A["d0b"], A["out"], y32=bufs["pred"], out_f32=3)
    else:  # d0b and the 3 predicted channels (one 16-channel tile) fused, d0b kept on chip
      self._conv_pair("d0b", "out", A["d0a"], bufs["pred"], 3)
    return bufs["pred"]

  # ----------------------------------------------------------------- backward
  def _layer_backward(self, name: str, x: torch.Tensor, g: torch.Tensor, gx: torch.Tensor | None) -> None:
    """Weight and bias gradients of layer `name` (input x, output gradient g) and, if gx is
    given, the input gradient (transposed convolution with the mirrored weights)."""
    p = self.params[name]
    n, h, w = x.shape[0], x.shape[1], x.shape[2]
    st = _stream(x)
    mask = self._bufs[self._shape]["act"][name] if p.relu else None
    nblocks = n * (h // 8) * (w // 8)
    groups = (p.cin // 16) * ((p.cout + 63) // 64)
    nchunk = max(1, min(nblocks, 2048 // max(groups, 1)))
    lib = _abi.library()
    part = self._workspace("wgrad", lib.mjpbr_unet_wgrad_workspace(n, h, w, p.cin, p.cout, nchunk))
    bpart = self._workspace("bgrad", lib.mjpbr_unet_bias_workspace(n, h, w, p.cout))
    _abi.call(
      "mjpbr_unet_conv3x3_wgrad", _abi.ptr(x), _abi.ptr(g), _abi.ptr(mask), _abi.ptr(part), nchunk, _abi.ptr(p.gw), _abi.ptr(bpart),
      _abi.ptr(p.gb), n, h, w, p.cin, p.cout, st,
    )
    if gx is not None:
      _abi.call(
        "mjpbr_unet_conv3x3", _abi.ptr(g), _abi.ptr(mask), _abi.ptr(p.wd), 0, _abi.ptr(gx), 0, 0, n, h, w, p.cout, p.cin, 0, st,
      )

  def backward(self, target: torch.Tensor) -> torch.Tensor:
    """L1 loss of the last forward against FP32 log-space targets (n * h * w, 3); fills every
    parameter gradient. Returns the loss (a device scalar)."""
    if not getattr(self, "_trained_forward", False):
      raise RuntimeError("backward() needs the activations of forward(x, training=True)")
    n, h, w = self._shape
    bufs = self._bufs[self._shape]
    A, G = bufs["act"], self._grad_bufs()
    st = _stream(target)
    npix = n * h * w
    if tuple(target.shape) != (npix, 3) or target.dtype != torch.float32 or not target.is_contiguous():
      raise ValueError(f"target must be contiguous float32 of shape ({npix}, 3), got {target.dtype} {tuple(target.shape)}")
    _abi.call(
      "mjpbr_unet_l1", _abi.ptr(bufs["pred"]), _abi.ptr(self.x), IN_CHANNELS, _abi.ptr(target), npix, OUT_CHANNELS, _abi.ptr(G["out"]),
      _abi.ptr(bufs["loss_part"]), st,
    )
    self._layer_backward("out", A["d0b"], G["out"], G["d0b"])
    self._layer_backward("d0b", A["d0a"], G["d0b"], G["d0a"])
    self._layer_backward("d0a", A["u0"], G["d0a"], G["u0"])
    _abi.call("mjpbr_unet_upcat_low_grad", _abi.ptr(G["u0"]), _abi.ptr(G["d1b"]), n, h, w, 64, 32, st)
    self._layer_backward("d1b", A["d1a"], G["d1b"], G["d1a"])
    self._layer_backward("d1a", A["u1"], G["d1a"], G["u1"])
    _abi.call("mjpbr_unet_upcat_low_grad", _abi.ptr(G["u1"]), _abi.ptr(G["b1"]), n, h // 2, w // 2, 128, 64, st)
    self._layer_backward("b1", A["b0"], G["b1"], G["b0"])
    self._layer_backward("b0", A["p1"], G["b0"], G["p1"])
    _abi.call("mjpbr_unet_skip_grad", _abi.ptr(G["p1"]), _abi.ptr(G["u1"]), _abi.ptr(G["e1b"]), n, h // 2, w // 2, 64, 128, st)
    self._layer_backward("e1b", A["e1a"], G["e1b"], G["e1a"])
    self._layer_backward("e1a", A["p0"], G["e1a"], G["p0"])
    _abi.call("mjpbr_unet_skip_grad", _abi.ptr(G["p0"]), _abi.ptr(G["u0"]), _abi.ptr(G["e0b"]), n, h, w, 32, 64, st)
    self._layer_backward("e0b", A["e0a"], G["e0b"], G["e0a"])
    self._layer_backward("e0a", self.x, G["e0a"], None)
    return bufs["loss_part"].sum() / (npix * 3)

  def adam_step(self, lr: float = 1e-3, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> None:
    self.step_count += 1
    for p in self.params.values():
      st = _stream(p.w)
      _abi.call(
        "mjpbr_unet_adam", _abi.ptr(p.w), _abi.ptr(p.mw), _abi.ptr(p.vw), _abi.ptr(p.gw), p.w.numel(), lr, b1, b2, eps, self.step_count,
        _abi.ptr(p.wf), _abi.ptr(p.wd), p.cin, p.cout, st,
      )
      _abi.call(
        "mjpbr_unet_adam", _abi.ptr(p.b), _abi.ptr(p.mb), _abi.ptr(p.vb), _abi.ptr(p.gb), p.b.numel(), lr, b1, b2, eps, self.step_count,
        0, 0, p.cin, p.cout, st,
      )


# --------------------------------------------------------------------- weights on disk


def save(net: UNet, path, meta: dict) -> None:
  """Writes the FP32 master weights and `meta` (JSON; must hold "guide_range") as safetensors."""
  from safetensors.torch import save_file

  Path(path).parent.mkdir(parents=True, exist_ok=True)
  save_file(net.state_dict(), str(path), metadata={"meta": json.dumps(meta)})


def load(path, device) -> tuple[UNet, dict]:
  """(network, metadata) from a safetensors file written by `save`."""
  from safetensors import safe_open
  from safetensors.torch import load_file

  path = Path(path)
  if not path.exists():
    raise FileNotFoundError(f"learned denoiser weights {path} are missing (train them with python -m mujoco_warp_pbr_amd.train_denoiser)")
  with safe_open(str(path), framework="pt") as f:
    meta = json.loads((f.metadata() or {}).get("meta", "{}"))
  net = UNet(device)
  net.load_state_dict(load_file(str(path)))
  return net, meta


# --------------------------------------------------------------------- render-path integration


def _camera_adr(rc, cam: int) -> int:
  adr = int(rc.base.rgb_adr[cam])
  if adr < 0:
    raise ValueError(f"camera {cam} renders no RGB: nothing to denoise")
  return adr


def pack_input(rc, cam: int, nworld: int, out: torch.Tensor | None = None) -> torch.Tensor:
  """BF16 network input (nworld, h, w, 32) of rendering camera `cam` from the path tracer's noisy
  HDR and first-hit AOVs in the context."""
  base = rc.base
  width, height = int(base.cam_res[cam][0]), int(base.cam_res[cam][1])
  if out is None:
    out = _i16((nworld, height, width, IN_CHANNELS), rc.hdr_data.device)
  _abi.call(
    "mjpbr_unet_pack_input", _abi.ptr(rc.hdr_data), int(rc.hdr_data.shape[1]), _abi.ptr(rc.aov_albedo), _abi.ptr(rc.aov_normal),
    _abi.ptr(rc.aov_depth), int(rc.aov_depth.shape[1]), _camera_adr(rc, cam), _abi.ptr(rc.exposure), int(rc.exposure.shape[0]), nworld,
    height, width, IN_CHANNELS, _abi.ptr(out), _stream(out),
  )
  return out


def log_target(rc, nworld: int, cam: int = 0) -> torch.Tensor:
  """FP32 log-space training target (nworld * h * w, 3) = log1p(max(hdr * exposure, 0)) of camera
  `cam` from the context's HDR buffer (a high-sample render)."""
  base = rc.base
  width, height = int(base.cam_res[cam][0]), int(base.cam_res[cam][1])
  adr = _camera_adr(rc, cam)
  hdr = rc.hdr_data[:nworld, adr : adr + width * height]
  e = rc.exposure.reshape(-1)
  e = e.repeat((nworld + e.numel() - 1) // e.numel())[:nworld].reshape(nworld, 1, 1)
  return torch.log1p(torch.clamp(hdr * e, min=0.0)).reshape(-1, 3).contiguous()


class Denoiser:
  """The learned denoiser of one render context: network, guide range and per-camera buffers."""

  def __init__(self, rc, path=None):
    dev = rc.hdr_data.device
    self.net, self.meta = load(path or DEFAULT_WEIGHTS, dev)
    rng = self.meta.get("guide_range")
    if rng is None or len(rng) != 2 * NSTATS:
      raise ValueError("learned denoiser weights carry no guide_range (8 floats)")
    self.range = torch.tensor(rng, dtype=torch.float32, device=dev)
    self.inputs: dict[int, torch.Tensor] = {}
    n = rc.base.nworld
    self.stats = torch.zeros((n, NSTATS), dtype=torch.float32, device=dev)
    self.fallback = torch.zeros((n,), dtype=torch.int32, device=dev)

  def __call__(self, rc, nworld: int, stream: int) -> None:
    """Filters the noisy HDR in rc.hdr_data into the RGB (and HDR) outputs: the U-Net for worlds
    whose guide statistics are in the trained range, the a-trous filter for the others."""
    from mujoco_warp_pbr_amd import denoise as atrous

    base = rc.base
    cams = [i for i in range(base.nrender) if int(base.rgb_adr[i]) >= 0]
    for k, cam in enumerate(cams):
      x = pack_input(rc, cam, nworld, self.inputs.get(cam))
      self.inputs[cam] = x
      w, h = int(base.cam_res[cam][0]), int(base.cam_res[cam][1])
      _abi.call(
        "mjpbr_unet_guide_stats", _abi.ptr(x), nworld, h, w, IN_CHANNELS, _abi.ptr(self.range), _abi.ptr(self.stats),
        _abi.ptr(self.fallback), int(k > 0), stream,
      )
    atrous.denoise(rc, nworld, stream, world_select=self.fallback)
    for cam in cams:
      x = self.inputs[cam]
      w, h = int(base.cam_res[cam][0]), int(base.cam_res[cam][1])
      pred = self.net.forward(x)
      _abi.call(
        "mjpbr_unet_unpack_output", _abi.ptr(pred), _abi.ptr(x), IN_CHANNELS, nworld, h, w, _abi.ptr(rc.exposure),
        int(rc.exposure.shape[0]), int(rc.tone_mapper), _abi.ptr(rc.hdr_data) if rc.render_hdr else 0, int(rc.hdr_data.shape[1]),
        _abi.ptr(base.rgb_data) if bool(base.render_rgb[cam]) else 0, int(base.rgb_data.shape[1]), _camera_adr(rc, cam),
        _abi.ptr(self.fallback), stream,
      )


def denoiser_for(rc) -> Denoiser:
  d = rc._host_cache.get("learned_denoiser")
  if d is None:
    d = Denoiser(rc, rc.denoiser_weights)
    rc._host_cache["learned_denoiser"] = d
  return d


def denoise(rc, nworld: int, stream: int) -> None:
  denoiser_for(rc)(rc, nworld, stream)


def guide_range_from_stats(stats: np.ndarray, margin: float = 0.1) -> list[float]:
  """[lo(4), hi(4)] of per-image guide statistics widened by `margin` of their spread."""
  lo, hi = stats.min(axis=0), stats.max(axis=0)
  pad = margin * (hi - lo) + 1e-3
  return [float(v) for v in np.concatenate([lo - pad, hi + pad])]
