"""Learned denoiser of the path-traced tier: a 3-level U-Net (32 / 64 / 128 channels, 3x3
convolutions, ReLU, 2x2 average pooling, nearest upsampling with skip concatenation) whose
forward and backward passes run in the hand-written MFMA kernels of csrc/unet.hip (BF16
activations and weights, FP32 accumulation, FP32 master weights with Adam).

Input (10 channels, padded to 32): log1p of the exposed 1-spp radiance, first-hit albedo,
normal, depth / (1 + depth). The network predicts a residual in log space: denoised =
expm1(out + log1p(noisy)). Training pairs come from the renderer itself (low- and high-sample
renders of the same seeded scenes); the a-trous filter stays as the fallback.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch

from mujoco_warp_pbr_amd import _abi

IN_CHANNELS = 32  # 10 used
OUT_CHANNELS = 32  # 3 used
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
USED_IN = 10


def flops_per_pixel() -> float:
  """Forward multiply-adds x 2 per full-resolution pixel (padding channels included)."""
  return sum(2.0 * 9 * cin * cout / (4**level) for _, cin, cout, _, level in LAYERS)


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
    for name, cin, cout, relu, _ in LAYERS:
      fan_in = 9 * (USED_IN if name == "e0a" else cin)
      std = math.sqrt(2.0 / fan_in) if relu else 0.1 * math.sqrt(1.0 / fan_in)
      w = (torch.randn((cout, 9, cin), generator=gen) * std).float()
      if name == "e0a":
        w[:, :, USED_IN:] = 0.0
      dev = self.device
      p = Param(
        cin, cout, relu,
        w.to(dev), torch.zeros(cout, device=dev),
        _i16((cout, 9, cin), dev), _i16((cin, 9, cout), dev),
        torch.zeros_like(w, device=dev), torch.zeros_like(w, device=dev),
        torch.zeros(cout, device=dev), torch.zeros(cout, device=dev),
        torch.zeros_like(w, device=dev), torch.zeros(cout, device=dev),
      )
      self.params[name] = p
    self.step_count = 0
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
      out[f"{name}.w"] = p.w.detach().cpu()
      out[f"{name}.b"] = p.b.detach().cpu()
    return out

  def load_state_dict(self, state: dict[str, torch.Tensor]) -> None:
    for name, p in self.params.items():
      p.w.copy_(state[f"{name}.w"].to(p.w.device))
      p.b.copy_(state[f"{name}.b"].to(p.b.device))
    self.refresh_layouts()

  # ----------------------------------------------------------------- buffers
  def _alloc(self, n: int, h: int, w: int) -> None:
    if self._shape == (n, h, w):
      return
    if h % 64 or w % 64:
      raise ValueError(f"image size must be a multiple of 64 (three levels of 16x16 blocks), got {h}x{w}")
    dev = self.device
    self._shape = (n, h, w)
    sizes = {0: (n, h, w), 1: (n, h // 2, w // 2), 2: (n, h // 4, w // 4)}
    self.act = {}
    self.grad = {}
    for name, cin, cout, _, level in LAYERS:
      self.act[name] = _i16((*sizes[level], cout), dev)
      self.grad[name] = _i16((*sizes[level], cout), dev)
    self.act["p0"] = _i16((*sizes[1], 32), dev)
    self.act["p1"] = _i16((*sizes[2], 64), dev)
    self.act["u1"] = _i16((*sizes[1], 192), dev)
    self.act["u0"] = _i16((*sizes[0], 96), dev)
    for k in ("p0", "p1", "u1", "u0"):
      self.grad[k] = torch.zeros_like(self.act[k])
    self.pred = torch.zeros((n * h * w, 3), dtype=torch.float32, device=dev)
    self.loss_part = torch.zeros(((n * h * w + 255) // 256,), dtype=torch.float32, device=dev)
    self._ws = {}

  def _workspace(self, key, numel):
    t = self._ws.get(key)
    if t is None or t.numel() < numel:
      t = torch.zeros(max(int(numel), 1), dtype=torch.float32, device=self.device)
      self._ws[key] = t
    return t

  # ----------------------------------------------------------------- forward
  def _conv(self, name, x, level, mask=None, weights=None, out=None, relu=None, bias=True, y32=None, out_f32=0, cin=None, cout=None):
    p = self.params[name]
    n, h, w = x.shape[0], x.shape[1], x.shape[2]
    cin = p.cin if cin is None else cin
    cout = p.cout if cout is None else cout
    _abi.call(
      "mjpbr_unet_conv3x3",
      _abi.ptr(x),
      _abi.ptr(mask),
      _abi.ptr(p.wf if weights is None else weights),
      _abi.ptr(p.b) if bias else 0,
      _abi.ptr(out),
      _abi.ptr(y32),
      out_f32,
      n,
      h,
      w,
      cin,
      cout,
      int(p.relu if relu is None else relu),
      _stream(x),
    )

  def forward(self, x: torch.Tensor) -> torch.Tensor:
    """x: BF16 bits (n, h, w, 32). Returns the FP32 log-space residual (n * h * w, 3)."""
    n, h, w, c = x.shape
    if c != IN_CHANNELS or x.dtype != torch.int16:
      raise ValueError(f"input must be int16 (BF16 bits) with {IN_CHANNELS} channels, got {x.dtype} {tuple(x.shape)}")
    self._alloc(n, h, w)
    self.x = x
    A = self.act
    st = _stream(x)
    self._conv("e0a", x, 0, out=A["e0a"])
    self._conv("e0b", A["e0a"], 0, out=A["e0b"])
    _abi.call("mjpbr_unet_avgpool2", _abi.ptr(A["e0b"]), _abi.ptr(A["p0"]), n, h, w, 32, 0, st)
    self._conv("e1a", A["p0"], 1, out=A["e1a"])
    self._conv("e1b", A["e1a"], 1, out=A["e1b"])
    _abi.call("mjpbr_unet_avgpool2", _abi.ptr(A["e1b"]), _abi.ptr(A["p1"]), n, h // 2, w // 2, 64, 0, st)
    self._conv("b0", A["p1"], 2, out=A["b0"])
    self._conv("b1", A["b0"], 2, out=A["b1"])
    _abi.call("mjpbr_unet_upcat", _abi.ptr(A["b1"]), _abi.ptr(A["e1b"]), _abi.ptr(A["u1"]), n, h // 2, w // 2, 128, 64, st)
    self._conv("d1a", A["u1"], 1, out=A["d1a"])
    self._conv("d1b", A["d1a"], 1, out=A["d1b"])
    _abi.call("mjpbr_unet_upcat", _abi.ptr(A["d1b"]), _abi.ptr(A["e0b"]), _abi.ptr(A["u0"]), n, h, w, 64, 32, st)
    self._conv("d0a", A["u0"], 0, out=A["d0a"])
    self._conv("d0b", A["d0a"], 0, out=A["d0b"])
    self._conv("out", A["d0b"], 0, out=A["out"], y32=self.pred, out_f32=3)
    return self.pred

  # ----------------------------------------------------------------- backward
  def _layer_backward(self, name: str, x: torch.Tensor, g: torch.Tensor, gx: torch.Tensor | None) -> None:
    """Weight and bias gradient of layer `name` (input x, output gradient g) and, if gx is
    given, the input gradient (transposed convolution with the mirrored weights)."""
    p = self.params[name]
    n, h, w = x.shape[0], x.shape[1], x.shape[2]
    st = _stream(x)
    mask = self.act[name] if p.relu else None
    nblocks = n * (h // 8) * (w // 8)
    groups = (p.cin // 16) * ((p.cout + 63) // 64)
    nchunk = max(1, min(nblocks, 2048 // max(groups, 1)))
    part = self._workspace("wgrad", _abi.library().mjpbr_unet_wgrad_workspace(n, h, w, p.cin, p.cout, nchunk))
    bpart = self._workspace("bgrad", _abi.library().mjpbr_unet_bias_workspace(n, h, w, p.cout))
    _abi.call(
      "mjpbr_unet_conv3x3_wgrad",
      _abi.ptr(x),
      _abi.ptr(g),
      _abi.ptr(mask),
      _abi.ptr(part),
      nchunk,
      _abi.ptr(p.gw),
      _abi.ptr(bpart),
      _abi.ptr(p.gb),
      n,
      h,
      w,
      p.cin,
      p.cout,
      st,
    )
    if gx is not None:
      _abi.call(
        "mjpbr_unet_conv3x3",
        _abi.ptr(g), _abi.ptr(mask), _abi.ptr(p.wd), 0, _abi.ptr(gx), 0, 0, n, h, w, p.cout, p.cin, 0, st,
      )

  def backward(self, target: torch.Tensor) -> torch.Tensor:
    """L1 loss against FP32 log-space targets (n * h * w, 3); fills every parameter gradient.
    Returns the loss (a device scalar)."""
    n, h, w = self._shape
    A, G = self.act, self.grad
    st = _stream(target)
    npix = n * h * w
    _abi.call(
      "mjpbr_unet_l1", _abi.ptr(self.pred), _abi.ptr(self.x), IN_CHANNELS, _abi.ptr(target), npix, OUT_CHANNELS, _abi.ptr(G["out"]),
      _abi.ptr(self.loss_part), st,
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
    return self.loss_part.sum() / (npix * 3)

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


def bf16_bits(x: torch.Tensor) -> torch.Tensor:
  """float32 tensor -> BF16 bit patterns (int16), round to nearest even."""
  return x.to(torch.bfloat16).view(torch.int16)


def from_bf16_bits(x: torch.Tensor) -> torch.Tensor:
  return x.view(torch.bfloat16).float()
