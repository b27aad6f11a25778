"""Learned denoiser: the NumPy oracle of the U-Net pinned by brute force, adjoint identities and
finite differences; the GPU matrix-core kernels (forward, data and weight gradients, pooling,
concatenation, loss, Adam) against the oracle on identical BF16 inputs.

GPU bar: FP32 outputs within 1e-5 of the summed term magnitudes (the FP32 accumulation order is
the only difference), BF16 outputs within one BF16 ulp, elementwise layers bit-exact, whole-net
gradients within 2% (relative L2 per layer: BF16 activations of the two sides may differ by one
ulp and propagate).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd import learned
from oracle import unet as ou

# ------------------------------------------------------------------------------ oracle pins (CPU)


def test_bf16_rounding_matches_torch():
  rng = np.random.default_rng(0)
  x = np.concatenate([rng.normal(0, 1, 4096) * 10.0 ** rng.integers(-8, 8, 4096), [0.0, -0.0, 1.0, 1 + 2**-8, 1 + 3 * 2**-8]])
  ref = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).float().double().numpy()
  np.testing.assert_array_equal(ou.bf16_round(x), ref)
  np.testing.assert_array_equal(ou.from_bits(ou.to_bits(ref)), ref)


def test_conv_matches_brute_force():
  rng = np.random.default_rng(1)
  x = rng.normal(size=(2, 5, 6, 3))
  w = rng.normal(size=(4, 9, 3))
  b = rng.normal(size=4)
  y = ou.conv3x3(x, w, b)
  ref = np.zeros_like(y)
  for n in range(2):
    for i in range(5):
      for j in range(6):
        for o in range(4):
          s = b[o]
          for t in range(9):
            yy, xx = i + t // 3 - 1, j + t % 3 - 1
            if 0 <= yy < 5 and 0 <= xx < 6:
              s += float(x[n, yy, xx] @ w[o, t])
          ref[n, i, j, o] = s
  np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_conv_gradients_are_adjoints():
  """<conv(x), g> = <x, conv_T(g)> = <w, wgrad(x, g)> for the conv's two linear arguments."""
  rng = np.random.default_rng(2)
  x = rng.normal(size=(2, 8, 6, 5))
  w = rng.normal(size=(7, 9, 5))
  g = rng.normal(size=(2, 8, 6, 7))
  lhs = np.sum(ou.conv3x3(x, w) * g)
  np.testing.assert_allclose(np.sum(x * ou.conv3x3_input_grad(g, w)), lhs, rtol=1e-12)
  np.testing.assert_allclose(np.sum(w * ou.conv3x3_weight_grad(x, g)), lhs, rtol=1e-12)


def test_network_gradients_match_finite_differences():
  """Central differences of the float64 network's L1 loss against its backward pass."""
  rng = np.random.default_rng(3)
  params = ou.init_params(seed=4)
  net = ou.Net(params)
  x = np.zeros((1, 8, 8, 32))
  x[..., :10] = rng.normal(size=(1, 8, 8, 10))
  target = rng.normal(size=(1, 8, 8, 3))
  _, cache = net.forward(x)
  grads = net.backward(cache, target)
  h = 1e-6
  checked = 0
  for name, cin, cout, _, _ in ou.LAYERS:
    w, b = params[name]
    used = 10 if name == "e0a" else cin
    outs = 3 if name == "out" else cout
    for _ in range(3):
      idx = (rng.integers(outs), rng.integers(9), rng.integers(used))
      old = w[idx]
      w[idx] = old + h
      lp = net.loss(x, target)
      w[idx] = old - h
      lm = net.loss(x, target)
      w[idx] = old
      fd = (lp - lm) / (2 * h)
      assert abs(fd - grads[name][0][idx]) <= 1e-6 + 1e-4 * abs(fd), (name, idx, fd, grads[name][0][idx])
      checked += 1
    o = rng.integers(outs)
    old = b[o]
    b[o] = old + h
    lp = net.loss(x, target)
    b[o] = old - h
    lm = net.loss(x, target)
    b[o] = old
    fd = (lp - lm) / (2 * h)
    assert abs(fd - grads[name][1][o]) <= 1e-6 + 1e-4 * abs(fd), (name, "bias", o, fd, grads[name][1][o])
  assert checked == 3 * len(ou.LAYERS)


def test_adam_first_step_is_signed_learning_rate():
  """With bias correction the first Adam step is -lr * g / (|g| + eps)."""
  g = np.array([3.0, -0.5, 1e-3, 0.0])
  w, _, _ = ou.adam(np.zeros(4), np.zeros(4), np.zeros(4), g, step=1, lr=1e-3)
  np.testing.assert_allclose(w, -1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def test_layer_tables_agree():
  assert learned.LAYERS == ou.LAYERS
  # 2 * 9 * sum(cin * cout / 4^level): 0.26 MFLOP per full-resolution pixel as executed
  expect = sum(18.0 * cin * cout / 4**lv for _, cin, cout, _, lv in ou.LAYERS)
  assert learned.flops_per_pixel() == pytest.approx(expect) == 258048.0
  # algorithmic: 10 real input channels of e0a, 3 real outputs of the last layer
  assert learned.flops_per_pixel(padded=False) == pytest.approx(expect - 18.0 * (22 * 32 + 32 * 29))


# ------------------------------------------------------------------------------ GPU kernels


def _bits(x, device):
  return torch.from_numpy(ou.to_bits(x)).to(device)


def _vals(t):
  return ou.from_bits(t.cpu().numpy())


def _rand_bf16(rng, shape, scale=1.0):
  return ou.bf16_round(rng.normal(0.0, scale, shape))


def _stream():
  return torch.cuda.current_stream().cuda_stream


@pytest.mark.gpu
def test_gpu_mfma_lane_layout(cuda):
  rng = np.random.default_rng(10)
  a = _rand_bf16(rng, (16, 32))
  b = _rand_bf16(rng, (32, 16))
  d = torch.zeros((16, 16), dtype=torch.float32, device=cuda)
  ta, tb = _bits(a, cuda), _bits(b, cuda)  # held: a freed temporary's memory could be reused by the next
  _abi.call("mjpbr_unet_mfma_probe", _abi.ptr(ta), _abi.ptr(tb), _abi.ptr(d), _stream())
  ref = a @ b
  np.testing.assert_allclose(d.cpu().numpy(), ref, rtol=0, atol=1e-5 * (np.abs(a) @ np.abs(b)).max())


def _conv_case(cuda, n, h, w, cin, cout, relu, with_bias, with_mask, seed):
  rng = np.random.default_rng(seed)
  x = _rand_bf16(rng, (n, h, w, cin))
  wt = _rand_bf16(rng, (cout, 9, cin), 0.1)
  bias = rng.normal(0, 0.1, cout).astype(np.float32).astype(np.float64) if with_bias else None
  mask = _rand_bf16(rng, (n, h, w, cin)) if with_mask else None
  xin = x * (mask > 0) if with_mask else x
  ref = ou.conv3x3(xin, wt, bias)
  mag = ou.conv3x3(np.abs(xin), np.abs(wt), None if bias is None else np.abs(bias))
  if relu:
    ref = np.maximum(ref, 0.0)
  y = torch.zeros((n, h, w, cout), dtype=torch.int16, device=cuda)
  y32 = torch.zeros((n, h, w, cout), dtype=torch.float32, device=cuda)
  bias_t = None if bias is None else torch.from_numpy(bias.astype(np.float32)).to(cuda)
  tx, tw = _bits(x, cuda), _bits(wt, cuda)
  tm = None if mask is None else _bits(mask, cuda)
  _abi.call(
    "mjpbr_unet_conv3x3", _abi.ptr(tx), _abi.ptr(tm), _abi.ptr(tw), _abi.ptr(bias_t), _abi.ptr(y), _abi.ptr(y32), cout, n, h, w, cin,
    cout, int(relu), _stream(),
  )
  got32 = y32.cpu().numpy().astype(np.float64)
  assert np.all(np.abs(got32 - ref) <= 1e-5 * mag + 1e-7), np.max(np.abs(got32 - ref) - 1e-5 * mag)
  got = _vals(y)
  np.testing.assert_array_equal(got, ou.bf16_round(got32))
  # within one BF16 ulp of the rounded float64 result
  ulp = np.maximum(np.abs(ref), 1e-30) * 2.0**-7
  assert np.all(np.abs(got - ou.bf16_round(ref)) <= ulp + 1e-7)


@pytest.mark.gpu
@pytest.mark.parametrize(
  "n,h,w,cin,cout,relu,bias,mask",
  [
    (2, 32, 32, 64, 48, True, True, False),  # partial column tile (48 of 64)
    (1, 16, 48, 32, 16, False, False, False),  # one column tile, non-square
    (2, 32, 16, 96, 128, True, True, False),  # two column-tile groups
    (1, 32, 32, 64, 32, False, False, True),  # ReLU-backward mask fused into the load
    (2, 48, 32, 32, 32, True, True, False),  # single chunk, two tiles: B operands in registers
    (1, 32, 32, 32, 16, True, True, True),  # single chunk, one tile, masked
    (3, 16, 16, 32, 64, True, True, False),  # single chunk, four tiles: weights through LDS; 3 blocks
  ],
)
def test_gpu_conv3x3_matches_oracle(cuda, n, h, w, cin, cout, relu, bias, mask):
  _conv_case(cuda, n, h, w, cin, cout, relu, bias, mask, seed=11 + cin + cout)


@pytest.mark.gpu
@pytest.mark.parametrize("n,h,w,cin,cout,nchunk,mask", [(2, 16, 24, 32, 48, 3, True), (1, 32, 32, 64, 128, 7, False)])
def test_gpu_weight_and_bias_gradients_match_oracle(cuda, n, h, w, cin, cout, nchunk, mask):
  rng = np.random.default_rng(20 + cin)
  x = _rand_bf16(rng, (n, h, w, cin))
  g = _rand_bf16(rng, (n, h, w, cout))
  mk = _rand_bf16(rng, (n, h, w, cout)) if mask else None
  gm = g * (mk > 0) if mask else g
  ref = ou.conv3x3_weight_grad(x, gm)
  mag = ou.conv3x3_weight_grad(np.abs(x), np.abs(gm))
  lib = _abi.library()
  part = torch.zeros(lib.mjpbr_unet_wgrad_workspace(n, h, w, cin, cout, nchunk), dtype=torch.float32, device=cuda)
  bpart = torch.zeros(lib.mjpbr_unet_bias_workspace(n, h, w, cout), dtype=torch.float32, device=cuda)
  grad = torch.full((cout, 9, cin), float("nan"), dtype=torch.float32, device=cuda)
  bgrad = torch.full((cout,), float("nan"), dtype=torch.float32, device=cuda)
  tx, tg = _bits(x, cuda), _bits(g, cuda)
  tm = None if mk is None else _bits(mk, cuda)
  _abi.call(
    "mjpbr_unet_conv3x3_wgrad", _abi.ptr(tx), _abi.ptr(tg), _abi.ptr(tm), _abi.ptr(part), nchunk, _abi.ptr(grad), _abi.ptr(bpart), _abi.ptr(bgrad), n, h, w, cin, cout, _stream(),
  )
  got = grad.cpu().numpy().astype(np.float64)
  assert np.all(np.abs(got - ref) <= 1e-5 * mag + 1e-7)
  bref = gm.sum(axis=(0, 1, 2))
  np.testing.assert_allclose(bgrad.cpu().numpy(), bref, rtol=0, atol=1e-5 * np.abs(gm).sum(axis=(0, 1, 2)).max() + 1e-7)


@pytest.mark.gpu
def test_gpu_elementwise_layers_are_exact(cuda):
  rng = np.random.default_rng(30)
  n, h, w = 2, 16, 32
  x = _rand_bf16(rng, (n, h, w, 32))
  st = _stream()
  y = torch.zeros((n, h // 2, w // 2, 32), dtype=torch.int16, device=cuda)
  tx = _bits(x, cuda)
  _abi.call("mjpbr_unet_avgpool2", _abi.ptr(tx), _abi.ptr(y), n, h, w, 32, 0, st)
  np.testing.assert_array_equal(_vals(y), ou.bf16_round(ou.avgpool2(x)))
  low = _rand_bf16(rng, (n, h // 2, w // 2, 64))
  cat = torch.zeros((n, h, w, 96), dtype=torch.int16, device=cuda)
  tlow = _bits(low, cuda)
  _abi.call("mjpbr_unet_upcat", _abi.ptr(tlow), _abi.ptr(tx), _abi.ptr(cat), n, h, w, 64, 32, st)
  np.testing.assert_array_equal(_vals(cat), ou.upcat(low, x))
  gcat = _rand_bf16(rng, (n, h, w, 96))
  lg = torch.zeros((n, h // 2, w // 2, 64), dtype=torch.int16, device=cuda)
  tgcat = _bits(gcat, cuda)
  _abi.call("mjpbr_unet_upcat_low_grad", _abi.ptr(tgcat), _abi.ptr(lg), n, h, w, 64, 32, st)
  ref_lg = gcat[..., :64].reshape(n, h // 2, 2, w // 2, 2, 64).sum(axis=(2, 4))
  np.testing.assert_array_equal(_vals(lg), ou.bf16_round(ref_lg))
  pg = _rand_bf16(rng, (n, h // 2, w // 2, 32))
  sg = torch.zeros((n, h, w, 32), dtype=torch.int16, device=cuda)
  tpg = _bits(pg, cuda)
  _abi.call("mjpbr_unet_skip_grad", _abi.ptr(tpg), _abi.ptr(tgcat), _abi.ptr(sg), n, h, w, 32, 64, st)
  np.testing.assert_array_equal(_vals(sg), ou.bf16_round(0.25 * ou.upsample2(pg) + gcat[..., 64:]))


def _upload(net: learned.UNet, params):
  net.load_state_dict({f"{k}.{s}": torch.from_numpy(np.asarray(v[i], dtype=np.float32)) for k, v in params.items() for i, s in ((0, "w"), (1, "b"))})


def _inputs(rng, n, h, w):
  """Random inputs and a learnable target: a 3x3 box blur of channels 0-2 plus half of 3-5."""
  x = np.zeros((n, h, w, 32))
  x[..., :10] = ou.bf16_round(np.abs(rng.normal(0.5, 0.5, (n, h, w, 10))))
  xp = np.pad(x[..., :3], ((0, 0), (1, 1), (1, 1), (0, 0)), mode="edge")
  blur = sum(xp[:, dy : dy + h, dx : dx + w] for dy in range(3) for dx in range(3)) / 9.0
  target = blur + 0.5 * x[..., 3:6]
  return x, target.astype(np.float32).astype(np.float64)


@pytest.mark.gpu
def test_gpu_network_forward_and_backward_match_oracle(cuda):
  rng = np.random.default_rng(40)
  params = ou.init_params(seed=41)
  params = {k: (w.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64)) for k, (w, b) in params.items()}
  n, h, w = 2, 64, 64
  x, target = _inputs(rng, n, h, w)
  net = learned.UNet(cuda)
  _upload(net, params)
  pred = net.forward(_bits(x, cuda), training=True)
  loss = net.backward(torch.from_numpy(target.reshape(-1, 3).astype(np.float32)).to(cuda))
  torch.cuda.synchronize()
  ref_net = ou.Net(params, bf16=True)
  ref_pred, cache = ref_net.forward(x)
  ref_grads = ref_net.backward(cache, target)
  got_pred = pred.cpu().numpy().reshape(n, h, w, 3) + x[..., :3]
  assert np.abs(got_pred - ref_pred).max() <= 2e-2 * np.abs(ref_pred - x[..., :3]).max()
  assert float(loss) == pytest.approx(np.abs(ref_pred - target).mean(), rel=1e-3)
  for name, *_ in ou.LAYERS:
    p = net.params[name]
    gw = p.gw.cpu().numpy()
    gb = p.gb.cpu().numpy()
    rw, rb = ref_grads[name]
    assert np.linalg.norm(gw - rw) <= 2e-2 * np.linalg.norm(rw) + 1e-9, name
    assert np.linalg.norm(gb - rb) <= 2e-2 * np.linalg.norm(rb) + 1e-9, name


@pytest.mark.gpu
def test_gpu_adam_matches_oracle_and_training_reduces_loss(cuda):
  rng = np.random.default_rng(50)
  params = ou.init_params(seed=51)
  net = learned.UNet(cuda)
  _upload(net, params)
  x, target = _inputs(rng, 2, 64, 64)
  xb = _bits(x, cuda)
  tt = torch.from_numpy(target.reshape(-1, 3).astype(np.float32)).to(cuda)
  net.forward(xb, training=True)
  first = float(net.backward(tt))
  g = {k: (p.gw.cpu().numpy().astype(np.float64), p.gb.cpu().numpy().astype(np.float64)) for k, p in net.params.items()}
  w0 = {k: p.w.cpu().numpy().astype(np.float64) for k, p in net.params.items()}
  net.adam_step(lr=1e-3)
  torch.cuda.synchronize()
  for k, p in net.params.items():
    want, _, _ = ou.adam(w0[k], 0.0, 0.0, g[k][0], step=1)
    np.testing.assert_allclose(p.w.cpu().numpy(), want, rtol=1e-5, atol=1e-7)
    np.testing.assert_array_equal(p.wf.cpu().numpy(), ou.to_bits(ou.bf16_round(p.w.cpu().numpy())))
  for _ in range(100):
    net.forward(xb, training=True)
    last = float(net.backward(tt))
    net.adam_step(lr=1e-3)
  assert last < 0.7 * first, (first, last)


@pytest.mark.gpu
def test_gpu_fused_upcat_conv_matches_materialised_concat(cuda):
  """The decoder convolution reading concat(upsample2x(low), skip) from its two sources equals
  the oracle and, bit for bit, the convolution of the materialised concatenation."""
  rng = np.random.default_rng(60)
  n, h, w, cl, cs, cout = 2, 32, 32, 64, 32, 32
  low = _rand_bf16(rng, (n, h // 2, w // 2, cl))
  skip = _rand_bf16(rng, (n, h, w, cs))
  wt = _rand_bf16(rng, (cout, 9, cl + cs), 0.1)
  bias = rng.normal(0, 0.1, cout).astype(np.float32)
  tl, tsk, tw = _bits(low, cuda), _bits(skip, cuda), _bits(wt, cuda)
  tb = torch.from_numpy(bias).to(cuda)
  fused = torch.zeros((n, h, w, cout), dtype=torch.int16, device=cuda)
  _abi.call("mjpbr_unet_conv3x3_upcat", _abi.ptr(tl), _abi.ptr(tsk), cl, _abi.ptr(tw), _abi.ptr(tb), _abi.ptr(fused), n, h, w, cl + cs, cout, 1,
            _stream())
  cat = torch.zeros((n, h, w, cl + cs), dtype=torch.int16, device=cuda)
  _abi.call("mjpbr_unet_upcat", _abi.ptr(tl), _abi.ptr(tsk), _abi.ptr(cat), n, h, w, cl, cs, _stream())
  plain = torch.zeros_like(fused)
  _abi.call("mjpbr_unet_conv3x3", _abi.ptr(cat), 0, _abi.ptr(tw), _abi.ptr(tb), _abi.ptr(plain), 0, 0, n, h, w, cl + cs, cout, 1, _stream())
  torch.cuda.synchronize()
  np.testing.assert_array_equal(fused.cpu().numpy(), plain.cpu().numpy())
  ref = np.maximum(ou.conv3x3(ou.upcat(low, skip), wt, bias.astype(np.float64)), 0.0)
  got = _vals(fused)
  assert np.all(np.abs(got - ou.bf16_round(ref)) <= np.abs(ref) * 2.0**-7 + 1e-3)


@pytest.mark.gpu
def test_gpu_inference_forward_equals_training_forward(cuda):
  rng = np.random.default_rng(61)
  net = learned.UNet(cuda, seed=5)
  x, _ = _inputs(rng, 2, 64, 128)
  xb = _bits(x, cuda)
  a = net.forward(xb, training=True).clone()
  b = net.forward(xb).clone()
  torch.cuda.synchronize()
  assert torch.equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize(
  "n,h,w,out_f32,w2_rows",
  [(1, 16, 16, 3, 16), (2, 32, 48, 1, 32), (3, 48, 32, 4, 16), (9, 128, 128, 3, 32)],  # the last: 576 blocks, 2 a workgroup
)
def test_gpu_fused_layer_pair_equals_two_convolutions(cuda, n, h, w, out_f32, w2_rows):
  """The inference pair (32 -> 32 with ReLU, then the first 16 output channels of the next layer
  in FP32) equals, bit for bit, the two convolutions run one after the other; its second layer
  agrees with the float64 oracle convolution of the intermediate."""
  rng = np.random.default_rng(80 + n + out_f32)
  x = _rand_bf16(rng, (n, h, w, 32))
  w1 = _rand_bf16(rng, (32, 9, 32), 0.1)
  w2 = _rand_bf16(rng, (w2_rows, 9, 32), 0.1)
  b1 = rng.normal(0, 0.1, 32).astype(np.float32)
  b2 = rng.normal(0, 0.1, w2_rows).astype(np.float32)
  tx, tw1, tw2 = _bits(x, cuda), _bits(w1, cuda), _bits(w2, cuda)
  tb1, tb2 = torch.from_numpy(b1).to(cuda), torch.from_numpy(b2).to(cuda)
  fused = torch.full((n * h * w, out_f32), float("nan"), dtype=torch.float32, device=cuda)
  _abi.call(
    "mjpbr_unet_conv3x3_pair", _abi.ptr(tx), _abi.ptr(tw1), _abi.ptr(tb1), _abi.ptr(tw2), _abi.ptr(tb2), _abi.ptr(fused), out_f32, n, h,
    w, _stream(),
  )
  mid = torch.zeros((n, h, w, 32), dtype=torch.int16, device=cuda)
  _abi.call("mjpbr_unet_conv3x3", _abi.ptr(tx), 0, _abi.ptr(tw1), _abi.ptr(tb1), _abi.ptr(mid), 0, 0, n, h, w, 32, 32, 1, _stream())
  two = torch.full_like(fused, float("nan"))
  _abi.call(
    "mjpbr_unet_conv3x3", _abi.ptr(mid), 0, _abi.ptr(tw2), _abi.ptr(tb2), 0, _abi.ptr(two), out_f32, n, h, w, 32, 16, 0, _stream()
  )
  torch.cuda.synchronize()
  assert torch.equal(fused, two)
  m = _vals(mid)
  ref = ou.conv3x3(m, w2[:16], b2[:16].astype(np.float64))[..., :out_f32].reshape(-1, out_f32)
  mag = ou.conv3x3(np.abs(m), np.abs(w2[:16]), np.abs(b2[:16]).astype(np.float64))[..., :out_f32].reshape(-1, out_f32)
  got = fused.cpu().numpy().astype(np.float64)
  assert np.all(np.abs(got - ref) <= 1e-5 * mag + 1e-7)
  assert np.abs(m).sum() > 0 and (m == 0).any()  # the ReLU clipped some channels, kept others


@pytest.mark.gpu
def test_gpu_fused_layer_pair_rejects_bad_shapes(cuda):
  z = torch.zeros(16 * 16 * 32, dtype=torch.int16, device=cuda)
  f = torch.zeros(16 * 16 * 4, dtype=torch.float32, device=cuda)
  wt = torch.zeros(16 * 9 * 32, dtype=torch.int16, device=cuda)
  with pytest.raises(_abi.MjpbrError):
    _abi.call("mjpbr_unet_conv3x3_pair", _abi.ptr(z), _abi.ptr(wt), _abi.ptr(f), _abi.ptr(wt), _abi.ptr(f), _abi.ptr(f), 3, 1, 16, 24, _stream())
  with pytest.raises(_abi.MjpbrError):
    _abi.call("mjpbr_unet_conv3x3_pair", _abi.ptr(z), _abi.ptr(wt), _abi.ptr(f), _abi.ptr(wt), _abi.ptr(f), _abi.ptr(f), 5, 1, 16, 16, _stream())


@pytest.mark.gpu
@pytest.mark.parametrize("n,h,w,cin,cout", [(2, 32, 48, 32, 32), (1, 32, 32, 64, 64), (1, 16, 16, 32, 64)])
def test_gpu_pooled_convolution_equals_convolution_then_pool(cuda, n, h, w, cin, cout):
  """The encoder convolution with its 2x2 pooling fused into the epilogue: both outputs equal,
  bit for bit, the plain convolution followed by the pooling kernel."""
  rng = np.random.default_rng(70 + cin + cout)
  x = _rand_bf16(rng, (n, h, w, cin))
  wt = _rand_bf16(rng, (cout, 9, cin), 0.1)
  tb = torch.from_numpy(rng.normal(0, 0.1, cout).astype(np.float32)).to(cuda)
  tx, tw = _bits(x, cuda), _bits(wt, cuda)
  y = torch.zeros((n, h, w, cout), dtype=torch.int16, device=cuda)
  pooled = torch.zeros((n, h // 2, w // 2, cout), dtype=torch.int16, device=cuda)
  _abi.call("mjpbr_unet_conv3x3_pool", _abi.ptr(tx), _abi.ptr(tw), _abi.ptr(tb), _abi.ptr(y), _abi.ptr(pooled), n, h, w, cin, cout, 1,
            _stream())
  y2 = torch.zeros_like(y)
  _abi.call("mjpbr_unet_conv3x3", _abi.ptr(tx), 0, _abi.ptr(tw), _abi.ptr(tb), _abi.ptr(y2), 0, 0, n, h, w, cin, cout, 1, _stream())
  p2 = torch.zeros_like(pooled)
  _abi.call("mjpbr_unet_avgpool2", _abi.ptr(y2), _abi.ptr(p2), n, h, w, cout, 0, _stream())
  torch.cuda.synchronize()
  assert torch.equal(y, y2)
  assert torch.equal(pooled, p2)
  assert pooled.abs().sum() > 0
