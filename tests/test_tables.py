"""Context tables: mip chains, SH9 irradiance and importance-sampling tables.

The host restatement in oracle/tables.py is pinned by closed forms (block means of known images,
the irradiance of a hemispherical sky and of a cos^2 sky, the sampling density of a constant sky);
the HIP kernels that build the tables on the GPU are compared with it (mips bit-exact, the
double-precision reductions within float32 rounding).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest
import torch

from mujoco_warp_pbr_amd import _abi
from mujoco_warp_pbr_amd import envmap
from mujoco_warp_pbr_amd import scenes
from mujoco_warp_pbr_amd import texture
from oracle import tables


def _sh_basis(n):
  x, y, z = n
  return np.array([0.282095, 0.488603 * y, 0.488603 * z, 0.488603 * x, 1.092548 * x * y, 1.092548 * y * z,
                   0.315392 * (3 * z * z - 1), 1.092548 * x * z, 0.546274 * (x * x - y * y)])


def _equirect(fn, h=128, w=256):
  """An equirect image of radiance fn(direction), row 0 = +z (envmap.uv_to_direction)."""
  v, u = np.meshgrid((np.arange(h) + 0.5) / h, (np.arange(w) + 0.5) / w, indexing="ij")
  d = envmap.uv_to_direction(u, v)
  return np.repeat(fn(d)[..., None], 3, axis=-1).astype(np.float32)


def _normals(k=30, seed=0):
  n = np.random.default_rng(seed).normal(size=(k, 3))
  return n / np.linalg.norm(n, axis=1, keepdims=True)


# ----------------------------------------------------------------------------- host pins
def test_mip_chain_is_block_means():
  img = np.arange(8 * 8 * 4, dtype=np.float64).reshape(8, 8, 4)
  levels = tables.mip_chain(img, rgba16f=True)
  assert [lv.shape[:2] for lv in levels] == [(8, 8), (4, 4), (2, 2), (1, 1)]
  np.testing.assert_array_equal(levels[3][0, 0], img.reshape(-1, 4).mean(axis=0).astype(np.float16))
  np.testing.assert_array_equal(levels[2][1, 0], img[4:8, 0:4].reshape(-1, 4).mean(axis=0).astype(np.float16))
  # an axis of one texel stays; uint8 rounds half up
  u8 = tables.mip_chain(np.stack([np.full((8, 2), 200.0), np.full((8, 2), 201.0)], -1)[..., [0, 1, 0, 1]], rgba16f=False)
  assert [lv.shape[:2] for lv in u8] == [(8, 2), (4, 1), (2, 1), (1, 1)]
  tall = tables.mip_chain(np.tile(np.array([0.0, 1.0])[None, :, None], (4, 1, 4)), rgba16f=False)
  assert tall[1][0, 0, 0] == 1  # mean 0.5 rounds half up


def test_sh9_of_constant_sky_is_constant():
  sh = tables.sh9(np.full((64, 128, 3), 0.7, np.float32))
  for n in _normals():
    np.testing.assert_allclose(_sh_basis(n) @ sh, 0.7, rtol=5e-4)  # midpoint rule in theta over 64 rows


def test_sh9_of_hemispherical_sky_is_half_cosine():
  """Radiance 1 above the horizon: irradiance / pi at normal n is (1 + n_z) / 2 exactly; the
  clamped-cosine kernel has no odd bands above 1 and the step has no even bands above 0, so
  nine coefficients represent it without truncation."""
  sh = tables.sh9(_equirect(lambda d: (d[..., 2] > 0).astype(np.float64), 256, 512))
  for n in _normals():
    assert _sh_basis(n) @ sh[:, 0] == pytest.approx((1 + n[2]) / 2, abs=4e-3)


def test_sh9_of_cos2_sky_pins_the_second_band():
  """Radiance z^2 = 1/3 + (2/3) P2(z): irradiance / pi = 1/3 + (1/4)(2/3) P2(n_z)."""
  sh = tables.sh9(_equirect(lambda d: d[..., 2] ** 2, 256, 512))
  for n in _normals():
    want = 1.0 / 3.0 + (3 * n[2] ** 2 - 1) / 12.0
    assert _sh_basis(n) @ sh[:, 0] == pytest.approx(want, abs=2e-4)


def test_sampling_tables_are_distributions_proportional_to_sin_theta():
  prob, marginal, conditional = tables.sampling_tables(scenes.procedural_sky(0, 64, 128))
  assert prob.sum() == pytest.approx(1.0, rel=1e-5)
  assert marginal[0] == 0 and marginal[-1] == 1 and np.all(np.diff(marginal) >= 0)
  assert np.all(conditional[:, 0] == 0) and np.all(conditional[:, -1] == 1) and np.all(np.diff(conditional, axis=1) >= 0)
  # constant sky of luminance 2: rows weigh 2 sin(theta) + 1e-6 x 2, columns are uniform
  p, m, c = tables.sampling_tables(np.full((32, 64, 3), 2.0, np.float32))
  rows = np.sin(np.pi * (np.arange(32) + 0.5) / 32) + 1e-6
  np.testing.assert_allclose(p.sum(axis=1), rows / rows.sum(), rtol=1e-5)
  np.testing.assert_allclose(c[5], np.arange(65) / 64, atol=1e-6)
  # wider than 256: 2x2 block means first
  wide = tables.sampling_tables(np.random.default_rng(1).uniform(0, 1, (200, 600, 3)).astype(np.float32))
  assert wide[0].shape == (50, 150)


# ----------------------------------------------------------------------------- GPU vs oracle
def _mips_on_gpu(rgba8, fmt, cuda):
  h, w = rgba8.shape[:2]
  shapes = texture.level_shapes(w, h)
  texel = 8 if fmt == texture.FORMAT_RGBA16F else 4
  adr, off = [], 0
  for sw, sh in shapes:
    adr.append(off)
    off += -(-sw * sh * texel // 16) * 16
  pool = torch.zeros(off, dtype=torch.uint8, device=cuda)
  work = torch.empty(8 * w * h, dtype=torch.float64, device=cuda)
  table = torch.as_tensor(texture.SRGB_TABLE.view(np.int16)).to(cuda)
  src = torch.as_tensor(rgba8).to(cuda)
  arr = (ctypes.c_int64 * len(adr))(*adr)
  _abi.call("mjpbr_texture_mips", _abi.ptr(src), w, h, fmt, _abi.ptr(table), _abi.ptr(pool), ctypes.cast(arr, ctypes.c_void_p),
            len(adr), _abi.ptr(work), work.numel(), 0)
  torch.cuda.synchronize()
  data = pool.cpu().numpy()
  dt = np.float16 if fmt == texture.FORMAT_RGBA16F else np.uint8
  return [data[a : a + sw * sh * texel].view(dt).reshape(sh, sw, 4) for a, (sw, sh) in zip(adr, shapes)]


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(64, 64), (48, 20), (7, 33), (1, 16)])
@pytest.mark.parametrize("fmt", [texture.FORMAT_RGBA8, texture.FORMAT_RGBA16F])
def test_gpu_texture_mips_match_oracle_bit_for_bit(cuda, shape, fmt):
  rgba8 = np.random.default_rng(shape[0] * 7 + fmt).integers(0, 256, (*shape, 4)).astype(np.uint8)
  got = _mips_on_gpu(rgba8, fmt, cuda)
  f16 = fmt == texture.FORMAT_RGBA16F
  want = tables.mip_chain(tables.texture_level0(rgba8, f16, texture.SRGB_TABLE), f16)
  assert len(got) == len(want)
  for lv, (g, w) in enumerate(zip(got, want)):
    np.testing.assert_array_equal(g.view(np.uint16) if f16 else g, w.view(np.uint16) if f16 else w, err_msg=f"level {lv}")


@pytest.mark.gpu
def test_gpu_environment_tables_match_oracle(cuda):
  """Radiance mips (bit for bit), SH9 and sampling tables of a sky with a sharp sun, two maps of
  different sizes (the second's tables are built at the first's shape)."""
  skies = [scenes.procedural_sky(0, 256, 512), scenes.procedural_sky(3, 128, 256) * 2.0]
  env = envmap.build_environments(skies, cuda)
  torch.cuda.synchronize()
  rad = env.radiance.cpu().numpy().view(np.float16)
  adr = env.level_adr.cpu().numpy()
  for e, im in enumerate(env.images):
    for lv, want in enumerate(tables.mip_chain(tables.radiance_level0(im), True)):
      a = int(adr[e, lv])
      np.testing.assert_array_equal(rad[a : a + want.size].view(np.uint16), want.ravel().view(np.uint16), err_msg=f"map {e} level {lv}")
    np.testing.assert_allclose(env.sh[e].cpu().numpy(), tables.sh9(im), rtol=1e-5, atol=1e-7)
    p, m, c = tables.sampling_tables(env.table_images[e])
    np.testing.assert_allclose(env.prob[e].cpu().numpy(), p, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(env.marginal[e].cpu().numpy(), m, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(env.conditional[e].cpu().numpy(), c, rtol=1e-5, atol=1e-7)
  assert env.prob.shape[1:] == (128, 256)
