"""NumPy restatement of the learned denoiser's U-Net — TEST INFRASTRUCTURE ONLY.

The reference has no learned denoiser (its only denoiser is the a-trous filter,
mujoco_warp_pbr/_src/denoise.py:63-201), so this oracle follows the network defined by
mujoco_warp_pbr_amd/learned.py (layer table LAYERS) and csrc/unet.hip: 3x3 convolutions with
zero padding (tap t = 3 dy + dx reads pixel (y + dy - 1, x + dx - 1)), ReLU, 2x2 average pooling,
nearest 2x upsampling with skip concatenation (upsampled channels first), an FP32 log-space
residual output over the first 3 input channels and an L1 loss.

Arithmetic is float64. With `bf16=True` every tensor the GPU stores in BF16 is rounded to BF16
(round to nearest even) at the same points, so the GPU and the oracle differ only by FP32
accumulation order. Pinned in tests/test_unet.py by brute-force loops, the adjoint identity
<conv(x), g> = <x, conv_T(g)>, and central finite differences of the whole network's loss.

Only tests/ may import this module.
"""

from __future__ import annotations

import numpy as np

# name, cin, cout, relu, level — must match mujoco_warp_pbr_amd.learned.LAYERS
LAYERS = (
  ("e0a", 32, 32, True, 0),
  ("e0b", 32, 32, True, 0),
  ("e1a", 32, 64, True, 1),
  ("e1b", 64, 64, True, 1),
  ("b0", 64, 128, True, 2),
  ("b1", 128, 128, True, 2),
  ("d1a", 192, 64, True, 1),
  ("d1b", 64, 64, True, 1),
  ("d0a", 96, 32, True, 0),
  ("d0b", 32, 32, True, 0),
  ("out", 32, 32, False, 0),
)


def bf16_round(x) -> np.ndarray:
  """Nearest-even rounding to BF16 (via float32), returned as float64."""
  f = np.asarray(x, dtype=np.float64).astype(np.float32)
  u = f.view(np.uint32).astype(np.uint64)
  nan = np.isnan(f)
  u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
  out = (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)
  out[nan] = np.nan
  return out


def to_bits(x) -> np.ndarray:
  """BF16 bit patterns (int16) of values already representable in BF16."""
  return (np.asarray(x, dtype=np.float64).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16).view(np.int16)


def from_bits(b) -> np.ndarray:
  return (np.asarray(b).view(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _r(x, bf16):
  return bf16_round(x) if bf16 else x


def conv3x3(x: np.ndarray, w: np.ndarray, bias=None) -> np.ndarray:
  """x (n, h, w, cin), w (cout, 9, cin) -> (n, h, w, cout), zero padding, no activation."""
  n, h, wd, _ = x.shape
  xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
  y = np.zeros((n, h, wd, w.shape[0]))
  for t in range(9):
    dy, dx = divmod(t, 3)
    y += np.einsum("nhwc,oc->nhwo", xp[:, dy : dy + h, dx : dx + wd], w[:, t, :], optimize=True)
  if bias is not None:
    y += bias
  return y


def conv3x3_input_grad(g: np.ndarray, w: np.ndarray) -> np.ndarray:
  """Adjoint of conv3x3 in x: g (n, h, w, cout) -> (n, h, w, cin)."""
  return conv3x3(g, mirrored_transpose(w))


def mirrored_transpose(w: np.ndarray) -> np.ndarray:
  """(cout, 9, cin) -> (cin, 9, cout) with taps reversed: the data-gradient weights."""
  return np.ascontiguousarray(w[:, ::-1, :].transpose(2, 1, 0))


def conv3x3_weight_grad(x: np.ndarray, g: np.ndarray) -> np.ndarray:
  """dL/dw (cout, 9, cin) for y = conv3x3(x, w) and dL/dy = g."""
  n, h, wd, cin = x.shape
  xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
  out = np.zeros((g.shape[-1], 9, cin))
  for t in range(9):
    dy, dx = divmod(t, 3)
    out[:, t, :] = np.einsum("nhwc,nhwo->oc", xp[:, dy : dy + h, dx : dx + wd], g, optimize=True)
  return out


def avgpool2(x):
  n, h, w, c = x.shape
  return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def upsample2(x):
  return x.repeat(2, axis=1).repeat(2, axis=2)


def upcat(low, skip):
  return np.concatenate([upsample2(low), skip], axis=-1)


def relu_mask(act):
  return act > 0


class Net:
  """Float64 U-Net with parameters {name: (w (cout, 9, cin), b (cout,))}."""

  def __init__(self, params: dict, bf16: bool = False):
    self.params = params
    self.bf16 = bf16

  def _w(self, name):
    w, b = self.params[name]
    return (bf16_round(w) if self.bf16 else w), b

  def _conv(self, name, x, relu):
    w, b = self._w(name)
    y = conv3x3(x, w, b)
    pre = y
    if relu:
      y = np.maximum(y, 0.0)
    return y, pre

  def forward(self, x: np.ndarray):
    """x (n, h, w, 32) -> (log-space prediction (n, h, w, 3) incl. the residual base, cache)."""
    bf = self.bf16
    A = {"x": x}
    relu = {name: r for name, _, _, r, _ in LAYERS}

    def conv(name, inp):
      y, pre = self._conv(name, inp, relu[name])
      A[name + ".pre"] = pre
      A[name] = _r(y, bf)
      return A[name]

    conv("e0a", x)
    conv("e0b", A["e0a"])
    A["p0"] = _r(avgpool2(A["e0b"]), bf)
    conv("e1a", A["p0"])
    conv("e1b", A["e1a"])
    A["p1"] = _r(avgpool2(A["e1b"]), bf)
    conv("b0", A["p1"])
    conv("b1", A["b0"])
    A["u1"] = upcat(A["b1"], A["e1b"])
    conv("d1a", A["u1"])
    conv("d1b", A["d1a"])
    A["u0"] = upcat(A["d1b"], A["e0b"])
    conv("d0a", A["u0"])
    conv("d0b", A["d0a"])
    conv("out", A["d0b"])
    pred = A["out.pre"][..., :3] + x[..., :3]  # FP32 output channels + residual base
    A["pred"] = pred
    return pred, A

  def loss(self, x, target):
    pred, _ = self.forward(x)
    return np.abs(pred - target).mean()

  def backward(self, A: dict, target: np.ndarray):
    """Gradients {name: (dw, db)} of the mean L1 loss, at the GPU's rounding points."""
    bf = self.bf16
    pred = A["pred"]
    count = pred.size
    d = pred - target
    g_out = np.zeros(A["out"].shape)
    g_out[..., :3] = np.sign(d) / count
    g_out = _r(g_out, bf)
    grads = {}
    G = {"out": g_out}
    relu = {name: r for name, _, _, r, _ in LAYERS}
    inputs = {
      "out": "d0b", "d0b": "d0a", "d0a": "u0", "d1b": "d1a", "d1a": "u1",
      "b1": "b0", "b0": "p1", "e1b": "e1a", "e1a": "p0", "e0b": "e0a", "e0a": "x",
    }

    def layer(name, want_input=True):
      g = G[name]
      if relu[name]:
        g = g * relu_mask(A[name])
      w, _ = self._w(name)
      grads[name] = (conv3x3_weight_grad(A[inputs[name]], g), g.sum(axis=(0, 1, 2)))
      if want_input:
        G[inputs[name]] = _r(conv3x3_input_grad(g, w), bf)

    layer("out")
    layer("d0b")
    layer("d0a")
    cl = 64
    u0 = G["u0"]
    n, h, wd, _ = u0.shape
    G["d1b"] = _r(u0[..., :cl].reshape(n, h // 2, 2, wd // 2, 2, cl).sum(axis=(2, 4)), bf)
    layer("d1b")
    layer("d1a")
    u1 = G["u1"]
    n1, h1, w1, _ = u1.shape
    G["b1"] = _r(u1[..., :128].reshape(n1, h1 // 2, 2, w1 // 2, 2, 128).sum(axis=(2, 4)), bf)
    layer("b1")
    layer("b0")
    G["e1b"] = _r(0.25 * upsample2(G["p1"]) + u1[..., 128:], bf)
    layer("e1b")
    layer("e1a")
    G["e0b"] = _r(0.25 * upsample2(G["p0"]) + u0[..., cl:], bf)
    layer("e0b")
    layer("e0a", want_input=False)
    return grads


def init_params(seed: int = 0, used_in: int = 10) -> dict:
  """He-normal weights (a small output layer) and zero biases, float64."""
  rng = np.random.default_rng(seed)
  params = {}
  for name, cin, cout, relu, _ in LAYERS:
    fan_in = 9 * (used_in if name == "e0a" else cin)
    std = np.sqrt(2.0 / fan_in) if relu else 0.1 * np.sqrt(1.0 / fan_in)
    w = rng.normal(0.0, std, (cout, 9, cin))
    if name == "e0a":
      w[:, :, used_in:] = 0.0
    params[name] = (w, rng.normal(0.0, 0.05, cout) if relu else np.zeros(cout))
  return params


def adam(w, m, v, g, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
  """One Adam step (Kingma & Ba 2015, bias-corrected); returns new (w, m, v)."""
  m = b1 * m + (1 - b1) * g
  v = b2 * v + (1 - b2) * g * g
  w = w - lr * (m / (1 - b1**step)) / (np.sqrt(v / (1 - b2**step)) + eps)
  return w, m, v
