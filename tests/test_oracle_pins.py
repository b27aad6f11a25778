"""Pins of the CPU oracle's scalar building blocks against closed forms and published values.

Every renderer test compares the HIP kernels with the oracle, so the oracle itself must be fixed
by something other than its own code. Each check here derives the expected value a different way
from the oracle: ray-primitive hits from signed-distance sphere tracing or a world-frame
quadric in float64, triangles from hand-worked rays and a float64 linear solve, the display path
from the reference's own NumPy twin and Hill's published ACES fit, the spot cone and specular
occlusion from their published definitions, and the a-trous pass from inputs whose filtered value
follows in closed form. A plausible slip in the oracle (a flipped normal, a dropped 1/size^2, a
cap tested on the wrong half, a changed constant, variance propagated with w instead of w^2, a
depth footprint without its x step) fails one of them.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from mujoco_warp_pbr_amd import model as mdl

PLANE, SPHERE, CAPSULE, ELLIPSOID, CYLINDER, BOX = (
  mdl.GEOM_PLANE,
  mdl.GEOM_SPHERE,
  mdl.GEOM_CAPSULE,
  mdl.GEOM_ELLIPSOID,
  mdl.GEOM_CYLINDER,
  mdl.GEOM_BOX,
)


# ----------------------------------------------------------------------------- primitives
def _rotation(rng) -> np.ndarray:
  q, r = np.linalg.qr(rng.normal(size=(3, 3)))
  q = q * np.sign(np.diag(r))
  if np.linalg.det(q) < 0:
    q[:, 0] = -q[:, 0]
  return q.astype(np.float32).astype(np.float64)


def _sdf(kind: int, size: np.ndarray, p: np.ndarray) -> float:
  """Exact signed distance in the geom frame (capsule / cylinder axis z)."""
  if kind == SPHERE:
    return float(np.linalg.norm(p) - size[0])
  if kind == CAPSULE:
    r, h = size[0], size[1]
    return float(np.linalg.norm(p - np.array([0.0, 0.0, np.clip(p[2], -h, h)])) - r)
  if kind == CYLINDER:
    r, h = size[0], size[1]
    d = np.array([math.hypot(p[0], p[1]) - r, abs(p[2]) - h])
    return float(min(max(d[0], d[1]), 0.0) + np.linalg.norm(np.maximum(d, 0.0)))
  if kind == BOX:
    q = np.abs(p) - size
    return float(np.linalg.norm(np.maximum(q, 0.0)) + min(q.max(), 0.0))
  raise ValueError(kind)


def _normal(kind: int, size: np.ndarray, p: np.ndarray) -> np.ndarray | None:
  """Outward normal at surface point p (geom frame), None within 1e-4 of an edge or rim."""
  if kind == SPHERE:
    return p / np.linalg.norm(p)
  if kind == CAPSULE:
    c = np.array([0.0, 0.0, np.clip(p[2], -size[1], size[1])])
    return (p - c) / np.linalg.norm(p - c)
  if kind == CYLINDER:
    r, h = size[0], size[1]
    on_cap = abs(abs(p[2]) - h) < 1e-6
    on_side = abs(math.hypot(p[0], p[1]) - r) < 1e-6
    if on_cap and on_side or (on_cap and math.hypot(p[0], p[1]) > r - 1e-4) or (on_side and abs(p[2]) > h - 1e-4):
      return None
    if on_cap:
      return np.array([0.0, 0.0, math.copysign(1.0, p[2])])
    return np.array([p[0], p[1], 0.0]) / math.hypot(p[0], p[1])
  if kind == BOX:
    rel = np.abs(p) / size
    order = np.argsort(rel)
    if rel[order[1]] > 1.0 - 1e-4:
      return None
    n = np.zeros(3)
    n[order[2]] = math.copysign(1.0, p[order[2]])
    return n
  raise ValueError(kind)


def _sphere_trace(kind, size, lo, ld):
  """First hit of a ray (geom frame, float64) by sphere tracing the exact SDF: (t, point) or
  None for a miss; 'graze' when the ray passes within 1e-6 of the surface without a clear hit."""
  t, closest = 0.0, np.inf
  for _ in range(20000):
    p = lo + t * ld
    d = _sdf(kind, size, p)
    closest = min(closest, d)
    if d < 1e-13:
      return t, p
    t += d
    if t > 50.0:
      return "graze" if closest < 1e-6 else None
  return "graze"


def _rays(rng, n: int, extent: float, axial: bool):
  """Origins on a sphere of radius 4 around the geom, aimed at points inside its bounds; with
  axial=True also rays along +-z (caps)."""
  rays = []
  for _ in range(n):
    o = rng.normal(size=3)
    o = 4.0 * o / np.linalg.norm(o)
    target = rng.uniform(-1.2, 1.2, size=3) * extent
    rays.append((o, target - o))
  if axial:
    for _ in range(n // 4):
      sign = rng.choice([-1.0, 1.0])
      off = rng.uniform(-0.7, 0.7, size=2) * extent[:2]
      o = np.array([off[0], off[1], 4.0 * sign])
      tgt = np.array([off[0] * 0.5, off[1] * 0.5, 0.0])
      rays.append((o, tgt - o))
  return rays


PRIMS = [
  (SPHERE, np.array([0.35, 0.0, 0.0])),
  (CAPSULE, np.array([0.2, 0.45, 0.0])),
  (CYLINDER, np.array([0.3, 0.25, 0.0])),
  (BOX, np.array([0.4, 0.15, 0.25])),
]


@pytest.mark.parametrize("kind,size", PRIMS, ids=["sphere", "capsule", "cylinder", "box"])
def test_primitive_hits_match_sphere_traced_sdf(kind, size):
  """traverse.py:131-192 (MuJoCo ray_sphere / capsule / cylinder / box): distance and outward
  normal of every hit against exact-SDF sphere tracing and the analytic surface normal."""
  rng = np.random.default_rng(7 + kind)
  size = size.astype(np.float32).astype(np.float64)
  extent = np.array([size[0], size[0], size[1] + size[0]]) if kind in (CAPSULE, CYLINDER) else np.array(
    [size[0]] * 3 if kind == SPHERE else size
  )
  if kind == CYLINDER:
    extent = np.array([size[0], size[0], size[1]])
  hits = caps = 0
  for trial in range(3):
    R = _rotation(rng) if trial else np.eye(3)
    pos = rng.uniform(-0.5, 0.5, size=3).astype(np.float32).astype(np.float64)
    for lo_dir in _rays(rng, 120, extent, kind in (CAPSULE, CYLINDER)):
      lo, ld = lo_dir
      ld = ld / np.linalg.norm(ld)
      o32 = (pos + R @ lo).astype(np.float32)
      d32 = (R @ ld).astype(np.float32)
      # the ray exactly as the oracle sees it, back in the geom frame (float64)
      lo64 = R.T @ (o32.astype(np.float64) - pos)
      ld64 = R.T @ d32.astype(np.float64)
      dist, n = oracle.ray_primitive(kind, pos, R.reshape(9), size, o32, d32)
      ref = _sphere_trace(kind, size, lo64, ld64 / np.linalg.norm(ld64))
      if isinstance(ref, str):
        continue
      if ref is None:
        assert dist < 0.0, (lo64, ld64, dist)
        continue
      t_ref, p_ref = ref
      t_ref = t_ref / np.linalg.norm(ld64)  # distance in units of the (unit float32) direction
      n_ref = _normal(kind, size, p_ref)
      if n_ref is None:
        continue
      if abs(np.dot(n_ref, ld64 / np.linalg.norm(ld64))) < 0.05:
        continue  # grazing: float32 quadratic loses digits
      assert dist >= 0.0, ("missed", kind, lo64, ld64, t_ref)
      assert dist == pytest.approx(t_ref, rel=2e-5, abs=2e-6)
      n_world = R @ n_ref
      assert float(np.dot(n, n_world)) > 1.0 - 1e-5, (kind, p_ref, n, n_world)
      hits += 1
      if kind in (CAPSULE, CYLINDER) and abs(p_ref[2]) > size[1] - 1e-9:
        caps += 1  # hemispherical end (capsule) or flat cap (cylinder)
  assert hits > 150
  if kind in (CAPSULE, CYLINDER):
    assert caps > 20, caps


def test_ellipsoid_hits_match_world_frame_quadric():
  """MuJoCo ray_ellipsoid: the hit solves (p - c)^T M (p - c) = 1 with M = R diag(1/s^2) R^T,
  and the normal is the quadric's gradient M (p - c)."""
  rng = np.random.default_rng(3)
  size = np.array([0.45, 0.2, 0.3], np.float32).astype(np.float64)
  n_checked = 0
  for _ in range(200):
    R = _rotation(rng)
    c = rng.uniform(-0.5, 0.5, 3).astype(np.float32).astype(np.float64)
    M = R @ np.diag(1.0 / size**2) @ R.T
    o = rng.normal(size=3)
    o = c + 3.0 * o / np.linalg.norm(o)
    d = c + rng.uniform(-0.3, 0.3, 3) - o
    o32, d32 = o.astype(np.float32), (d / np.linalg.norm(d)).astype(np.float32)
    o64, d64 = o32.astype(np.float64), d32.astype(np.float64)
    a = d64 @ M @ d64
    b = 2.0 * d64 @ M @ (o64 - c)
    cc = (o64 - c) @ M @ (o64 - c) - 1.0
    disc = b * b - 4 * a * cc
    dist, n = oracle.ray_primitive(ELLIPSOID, c, R.reshape(9), size, o32, d32)
    if disc < 0:
      assert dist < 0
      continue
    t = (-b - math.sqrt(disc)) / (2 * a)
    p = o64 + t * d64
    g = M @ (p - c)
    g /= np.linalg.norm(g)
    if abs(g @ d64) < 0.05:
      continue
    assert dist == pytest.approx(t, rel=2e-5, abs=2e-6)
    assert float(np.dot(n, g)) > 1.0 - 1e-5, (n, g)
    n_checked += 1
  assert n_checked > 60


def test_plane_hits_front_side_within_extent():
  """MuJoCo ray_plane: hits only from the normal's side, inside |x| <= size.x, |y| <= size.y
  (size <= 0: unbounded), at t = -(n . (o - c)) / (n . d), normal = the frame's z axis."""
  rng = np.random.default_rng(5)
  for _ in range(100):
    R = _rotation(rng)
    c = rng.uniform(-1, 1, 3).astype(np.float32).astype(np.float64)
    size = np.array([0.5, 0.8, 0.0]) if rng.random() < 0.5 else np.zeros(3)
    nz = R[:, 2]
    o = c + nz * rng.uniform(0.5, 2.0) + R[:, 0] * rng.uniform(-1, 1) + R[:, 1] * rng.uniform(-1, 1)
    d = -nz + 0.5 * rng.normal(size=3)
    d /= np.linalg.norm(d)
    o32, d32 = o.astype(np.float32), d.astype(np.float32)
    dist, n = oracle.ray_primitive(PLANE, c, R.reshape(9), size, o32, d32)
    o64, d64 = o32.astype(np.float64), d32.astype(np.float64)
    den = nz @ d64
    if den >= 0:
      assert dist < 0
      continue
    t = -(nz @ (o64 - c)) / den
    local = R.T @ (o64 + t * d64 - c)
    inside = size[0] <= 0 or (abs(local[0]) <= size[0] and abs(local[1]) <= size[1])
    if size[0] > 0 and min(abs(abs(local[0]) - size[0]), abs(abs(local[1]) - size[1])) < 1e-5:
      continue
    if not inside:
      assert dist < 0
      continue
    assert dist == pytest.approx(t, rel=1e-5, abs=1e-6)
    np.testing.assert_allclose(n, nz, atol=1e-6)
    # from behind: never a hit
    back, _ = oracle.ray_primitive(PLANE, c, R.reshape(9), size, (c - (o64 - c)).astype(np.float32), (-d64).astype(np.float32))
    assert back < 0


# ----------------------------------------------------------------------------- triangles
TRI = ([0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])


@pytest.mark.parametrize(
  "origin,direction,tmax,expect",
  [
    ((0.25, 0.25, 1.0), (0, 0, -1), 10.0, (1.0, 0.25, 0.25)),  # interior
    ((0.5, 0.0, 1.0), (0, 0, -1), 10.0, (1.0, 0.5, 0.0)),  # on edge v0-v1
    ((0.5, 0.5, 2.0), (0, 0, -1), 10.0, (2.0, 0.5, 0.5)),  # on edge v1-v2
    ((0.0, 0.0, 1.0), (0, 0, -1), 10.0, (1.0, 0.0, 0.0)),  # on vertex v0
    ((0.0, 1.0, 0.5), (0, 0, -1), 10.0, (0.5, 0.0, 1.0)),  # on vertex v2
    ((0.25, 0.25, -1.0), (0, 0, 1), 10.0, (1.0, 0.25, 0.25)),  # back face: double-sided
    ((0.6, 0.6, 1.0), (0, 0, -1), 10.0, None),  # outside u + v > 1
    ((-0.01, 0.5, 1.0), (0, 0, -1), 10.0, None),  # outside u < 0
    ((0.25, 0.25, 1.0), (0, 0, 1), 10.0, None),  # triangle behind the origin
    ((0.25, 0.25, 1.0), (1, 0, 0), 10.0, None),  # parallel to the plane
    ((0.25, 0.25, 1.0), (0, 0, -1), 1.0, None),  # t == tmax is excluded
    ((0.25, 0.25, 1.0), (0, 0, -1), 1.0001, (1.0, 0.25, 0.25)),
  ],
)
def test_triangle_known_answers(origin, direction, tmax, expect):
  hit, t, b1, b2 = oracle.ray_triangle(origin, direction, *TRI, tmax)
  if expect is None:
    assert not hit
  else:
    assert hit
    assert (t, b1, b2) == pytest.approx(expect, abs=1e-7)


def test_triangle_matches_linear_solve():
  """o + t d = v0 + b1 (v1 - v0) + b2 (v2 - v0) solved in float64; hit iff t >= 0, b1, b2 >= 0,
  b1 + b2 <= 1 (rays within 1e-5 of an edge are skipped)."""
  rng = np.random.default_rng(11)
  checked = 0
  for _ in range(500):
    v = rng.uniform(-1, 1, (3, 3)).astype(np.float32).astype(np.float64)
    o = rng.uniform(-2, 2, 3).astype(np.float32).astype(np.float64)
    tgt = v.mean(0) + rng.normal(scale=0.6, size=3)
    d = (tgt - o).astype(np.float32).astype(np.float64)
    A = np.stack([-d, v[1] - v[0], v[2] - v[0]], axis=1)
    if abs(np.linalg.det(A)) < 1e-3:
      continue
    t, b1, b2 = np.linalg.solve(A, o - v[0])
    margin = min(b1, b2, 1 - b1 - b2)
    if abs(margin) < 1e-5 or abs(t) < 1e-5:
      continue
    hit, tt, bb1, bb2 = oracle.ray_triangle(o, d, v[0], v[1], v[2], 1e30)
    assert hit == (margin > 0 and t > 0)
    if hit:
      assert (tt, bb1, bb2) == pytest.approx((t, b1, b2), rel=1e-4, abs=1e-5)
    checked += 1
  assert checked > 400


# ----------------------------------------------------------------------------- display path
def _reference_tone_map(hdr: np.ndarray) -> np.ndarray:
  """The reference's NumPy twin of its display path, before the sRGB curve
  (benchmarks/scenes.py:313-329: exposure 1, PBR Neutral, clip)."""
  c = np.maximum(hdr, 0.0)
  x = c.min(axis=-1, keepdims=True)
  offset = np.where(x < 0.08, x - 6.25 * x * x, 0.04)
  c = c - offset
  peak = c.max(axis=-1, keepdims=True)
  start, desat = 0.76, 0.15
  d = 1.0 - start
  new_peak = 1.0 - d * d / (peak + d - start)
  compressed = c * (new_peak / np.maximum(peak, 1e-9))
  g = 1.0 - 1.0 / (desat * (peak - new_peak) + 1.0)
  compressed = compressed * (1 - g) + new_peak * g
  return np.clip(np.where(peak < start, c, compressed), 0.0, 1.0)


def _hdr_grid() -> np.ndarray:
  rng = np.random.default_rng(2)
  levels = np.array([0.0, 0.01, 0.05, 0.079, 0.2, 0.5, 0.75, 0.77, 0.9, 1.0, 1.5, 3.0, 10.0, 100.0])
  grid = np.stack(np.meshgrid(levels, levels, levels, indexing="ij"), -1).reshape(-1, 3)
  return np.concatenate([grid, rng.lognormal(0.0, 1.5, (500, 3)), -rng.uniform(0, 1, (20, 3))])


def test_pbr_neutral_matches_reference_numpy_twin():
  """color.py:40-60 (Khronos PBR Neutral) against benchmarks/scenes.py:313-329."""
  hdr = _hdr_grid()
  got = np.stack([oracle.display(c, 0) for c in hdr])
  np.testing.assert_allclose(got, _reference_tone_map(hdr), atol=2e-6, rtol=0)


def _hill_aces(c: np.ndarray) -> np.ndarray:
  """Stephen Hill's ACES RRT + ODT fit (BakingLab ACES.hlsl, 2016), float64."""
  inp = np.array([[0.59719, 0.35458, 0.04823], [0.07600, 0.90834, 0.01566], [0.02840, 0.13383, 0.83777]])
  out = np.array([[1.60475, -0.53108, -0.07367], [-0.10208, 1.10813, -0.00605], [-0.00327, -0.07276, 1.07602]])
  v = c @ inp.T
  v = (v * (v + 0.0245786) - 0.000090537) / (v * (0.983729 * v + 0.4329510) + 0.238081)
  return v @ out.T


def test_aces_matches_hill_fit_and_keeps_grey_neutral():
  """color.py:63-73: display = clip(ACESFitted(max(c, 0))). Grey stays grey (both of Hill's
  matrices have unit row sums to 1e-5), so on the grey axis the curve alone decides the value."""
  hdr = _hdr_grid()
  got = np.stack([oracle.display(c, 1) for c in hdr])
  np.testing.assert_allclose(got, np.clip(_hill_aces(np.maximum(hdr, 0.0)), 0, 1), atol=3e-6, rtol=0)
  for x in (0.01, 0.18, 0.5, 1.0, 2.0, 8.0):
    g = oracle.display((x, x, x), 1)
    assert np.ptp(g) < 1e-5, (x, g)
    v = x
    curve = (v * (v + 0.0245786) - 0.000090537) / (v * (0.983729 * v + 0.4329510) + 0.238081)
    assert g[0] == pytest.approx(min(curve, 1.0), abs=1e-5)


def test_clamp_tone_mapper_is_a_clip():
  hdr = _hdr_grid()
  got = np.stack([oracle.display(c, 2) for c in hdr])
  np.testing.assert_allclose(got, np.clip(hdr, 0, 1), atol=0)


def _srgb8(x: float) -> int:
  """IEC 61966-2-1 sRGB transfer function, 8-bit with +0.5 rounding (color.py:76-119)."""
  s = 12.92 * x if x <= 0.0031308 else 1.055 * x ** (1 / 2.4) - 0.055
  return int(s * 255.0 + 0.5)


def test_encode_is_srgb8_in_mujoco_warp_byte_order():
  """encode_display packs with MuJoCo Warp's pack_rgba_to_uint32: 0xAARRGGBB (the reference's
  Newton adapter and viewer read R from bits 16-23: newton-pbr/newton_pbr/sensor.py:72,
  viewer.py:44-51)."""
  assert oracle.encode((1.0, 0.0, 0.0)) == 0xFFFF0000
  assert oracle.encode((0.0, 1.0, 0.0)) == 0xFF00FF00
  assert oracle.encode((0.0, 0.0, 1.0)) == 0xFF0000FF
  assert oracle.encode((0.5, 0.5, 0.5)) == 0xFF000000 | (188 << 16) | (188 << 8) | 188
  for x in np.linspace(0.0, 1.0, 1001):
    word = oracle.encode((x, 0.0, 0.0))
    assert abs(((word >> 16) & 0xFF) - _srgb8(float(np.float32(x)))) <= 1
  # exact below the linear segment's end, where no power is involved
  for x in np.linspace(0.0, 0.0031308, 50):
    assert (oracle.encode((0.0, 0.0, x)) & 0xFF) == _srgb8(float(np.float32(x)))


# ----------------------------------------------------------------------------- lights
def _light_at_angle(theta: float, model: int, cutoff: float = 40.0, exponent: float = 0.0, spot: bool = True):
  """Spot light at the origin pointing down -z; shade point at distance 2 and angle theta."""
  point = (2.0 * math.sin(theta), 0.0, -2.0 * math.cos(theta))
  lt = mdl.LIGHT_SPOT if spot else mdl.LIGHT_POINT
  return oracle.eval_light(lt, model, True, (0, 0, 0), (0, 0, -1), cutoff, exponent, (1, 0, 0), 1.0, 1e6, point)


def test_physical_spot_cone():
  """lights.py:151-198 with SPOT_INNER_FRACTION 0.75 (lights.py:37): full intensity inside
  0.75 x cutoff, none outside the cutoff, a smooth monotone falloff in between."""
  cutoff = 40.0
  outer = math.radians(cutoff)
  full = _light_at_angle(0.0, 0, spot=False)["diffuse"][0]
  assert full == pytest.approx(0.25, rel=1e-5)  # 1 / d^2 at d = 2, window ~ 1 at range 1e6
  for f in (0.0, 0.3, 0.6, 0.74):
    assert _light_at_angle(f * outer, 0)["diffuse"][0] == pytest.approx(full, rel=1e-5), f
  for f in (1.0, 1.1, 1.5):
    assert _light_at_angle(f * outer, 0)["diffuse"][0] == 0.0, f
  inner = 0.75 * outer
  prev = full
  for f in np.linspace(0.76, 0.99, 12):
    v = _light_at_angle(f * outer, 0)["diffuse"][0]
    assert 0.0 < v < prev
    prev = v
  # smooth step squared in cos theta between the cones
  theta = 0.5 * (inner + outer)
  s = (math.cos(theta) - math.cos(outer)) / (math.cos(inner) - math.cos(outer))
  assert _light_at_angle(theta, 0)["diffuse"][0] == pytest.approx(full * s * s, rel=1e-4)


def test_mujoco_spot_cone_and_exponent():
  """MuJoCo (OpenGL) spot: zero outside the cutoff, cos(theta)^exponent inside, pi x the
  attenuated diffuse colour (lights.py:180-198)."""
  for theta, expo in ((0.2, 0.0), (0.3, 2.0), (0.6, 5.0)):
    v = _light_at_angle(theta, 1, cutoff=45.0, exponent=expo)["diffuse"][0]
    assert v == pytest.approx(math.pi * math.cos(theta) ** expo, rel=1e-5)
  assert _light_at_angle(math.radians(46.0), 1, cutoff=45.0)["diffuse"][0] == 0.0


def test_physical_point_light_window():
  """Physical point light: I / max(d^2, 1e-4) x (clamp(1 - (d/range)^4, 0, 1))^2."""
  for d, rng_ in ((1.0, 2.0), (1.5, 2.0), (0.5, 10.0), (3.0, 2.0)):
    out = oracle.eval_light(mdl.LIGHT_POINT, 0, False, (0, 0, d), (0, 0, -1), 0, 0, (1, 0, 0), 3.0, rng_, (0, 0, 0))
    window = min(max(1 - (d / rng_) ** 4, 0.0), 1.0) ** 2
    assert out["diffuse"][0] == pytest.approx(3.0 * window / d**2, rel=1e-5, abs=1e-9)
    assert out["dist"] == pytest.approx(d, rel=1e-6)
    np.testing.assert_allclose(out["l"], (0, 0, 1), atol=1e-6)


# ----------------------------------------------------------------------------- specular occlusion
def test_specular_occlusion_matches_lagarde():
  """render.py:217-219: Lagarde & de Rousiers, "Moving Frostbite to PBR" (2014), listing 26:
  saturate(pow(NoV + AO, exp2(-16 alpha - 1)) - 1 + AO)."""
  for nov in (1e-4, 0.1, 0.3, 0.7, 1.0):
    for ao in (0.0, 0.2, 0.5, 0.9, 1.0):
      for alpha in (0.002025, 0.05, 0.2, 0.5, 1.0):
        want = min(max((nov + ao) ** (2.0 ** (-16.0 * alpha - 1.0)) - 1.0 + ao, 0.0), 1.0)
        assert oracle.spec_ao(nov, ao, alpha) == pytest.approx(want, abs=3e-6), (nov, ao, alpha)
  # no occlusion keeps the full specular term; full occlusion removes it at grazing view
  assert oracle.spec_ao(0.3, 1.0, 0.3) == 1.0
  assert oracle.spec_ao(1e-4, 0.0, 0.3) == 0.0


# ----------------------------------------------------------------------------- a-trous pass
def _features(h, w):
  normal = np.zeros((h, w, 3))
  normal[..., 2] = 1.0
  return normal, np.ones((h, w)), np.full((h, w, 3), 0.5)


def test_atrous_variance_propagates_with_squared_weights():
  """denoise.py:160-178: on a flat image every edge weight is 1, so the centre of a 5 x 5 image
  at step 1 takes the B3 x B3 weights: colour unchanged, variance x sum(w^2) / (sum w)^2 =
  (1/16^2 + 1/4^2 + 3/8^2 + 1/4^2 + 1/16^2)^2."""
  normal, depth, albedo = _features(5, 5)
  src = np.zeros((5, 5, 4))
  src[..., :3] = (0.3, 0.6, 0.2)
  src[..., 3] = 0.8
  out = oracle.atrous_pass(src, normal, depth, albedo, 1)
  b2 = (2 * 0.0625**2 + 2 * 0.25**2 + 0.375**2) ** 2
  np.testing.assert_allclose(out[2, 2, :3], (0.3, 0.6, 0.2), rtol=1e-6)
  assert out[2, 2, 3] == pytest.approx(0.8 * b2, rel=1e-6)
  # a corner keeps only the in-image taps: (3/8, 1/4, 1/16)^2 normalised by their sum
  w = np.array([0.375, 0.25, 0.0625])
  corner = (np.sum(w**2) ** 2) / (np.sum(w) ** 4)
  assert out[0, 0, 3] == pytest.approx(0.8 * corner, rel=1e-6)


def test_atrous_depth_weight_scales_with_step():
  """denoise.py:146-147: the depth term is |dz| / (sigma_z z0 step). Centre column at depth 1,
  the other columns at 1 + 2 sigma_z: at step 2 each off-column tap weighs exp(-1) (exp(-2)
  without the step). Colours of equal luminance keep the luminance term at zero."""
  normal, depth, albedo = _features(5, 5)
  sigma_z = 0.02
  depth[:, [0, 1, 3, 4]] = 1.0 + 2.0 * sigma_z
  lum = 0.5
  red = np.array([lum / 0.2126, 0.0, 0.0])
  blue = np.array([0.0, 0.0, lum / 0.0722])
  src = np.zeros((5, 5, 4))
  src[..., :3] = blue
  src[:, 2, :3] = red
  src[..., 3] = 0.01
  out = oracle.atrous_pass(src, normal, depth, albedo, 2, sigma=(4.0, 128.0, sigma_z, 0.05))
  # step 2 at (2, 2): taps at x, y in {0, 2, 4} with B3 weights 1/4, 3/8, 1/4
  by = np.array([0.25, 0.375, 0.25])
  col_w = np.array([0.25 * math.exp(-1.0), 0.375, 0.25 * math.exp(-1.0)])
  w = by.sum() * col_w
  expect = (w[1] * red + (w[0] + w[2]) * blue) / w.sum()
  np.testing.assert_allclose(out[2, 2, :3], expect, rtol=2e-5)
  # variance: sum(w^2) var / (sum w)^2 with the same weights
  w2 = np.outer(by, col_w)
  assert out[2, 2, 3] == pytest.approx(0.01 * np.sum(w2**2) / np.sum(w2) ** 2, rel=2e-5)


def test_atrous_background_passes_through_and_gets_no_weight():
  """denoise.py:118-121 and :152: a zero normal passes the pixel through, and a background
  neighbour gets weight zero in its neighbours' sums."""
  normal, depth, albedo = _features(5, 5)
  normal[2, 3] = 0.0
  src = np.zeros((5, 5, 4))
  src[..., :3] = 0.2
  src[2, 3, :3] = 5.0
  src[..., 3] = 0.5
  out = oracle.atrous_pass(src, normal, depth, albedo, 1)
  np.testing.assert_array_equal(out[2, 3], src[2, 3].astype(np.float32))
  np.testing.assert_allclose(out[2, 2, :3], 0.2, rtol=1e-6)


# ----------------------------------------------------------------------------- BRDF
def test_ggx_distribution_is_normalised_and_matches_walter():
  """brdf.py:32-37: D_GGX = alpha^2 / (pi cos^4 (alpha^2 + tan^2)^2) (Walter et al. 2007), and
  its projected integral over the hemisphere, int D(h) (n.h) dw, is one."""
  for alpha in (0.05, 0.2, 0.5, 0.9):
    # quadrature in theta (the lobe is rotationally symmetric)
    theta = np.linspace(0.0, math.pi / 2, 200001)
    c = np.cos(theta)
    tan2 = np.tan(theta) ** 2
    d = alpha**2 / (math.pi * c**4 * (alpha**2 + tan2) ** 2)
    integral = np.trapezoid(d * c * np.sin(theta) * 2 * math.pi, theta)
    assert integral == pytest.approx(1.0, abs=2e-3)
    for th in (0.0, 0.3, 0.8, 1.2):
      # lobe with F = 1 (f0 = 1), V divided out through nov = nol = 1 (V = 1/4 there)
      got = oracle.spec_lobe((1, 1, 1), alpha, math.cos(th), 1.0, 1.0, 1.0)[0]
      want = alpha**2 / (math.pi * math.cos(th) ** 4 * (alpha**2 + math.tan(th) ** 2) ** 2) * 0.25
      assert got == pytest.approx(want, rel=2e-5)


def test_height_correlated_smith_matches_heitz():
  """brdf.py:40-46: V = G2 / (4 NoV NoL) with the height-correlated Smith G2 = 1 / (1 + L(v) +
  L(l)), L = (-1 + sqrt(1 + alpha^2 tan^2 theta)) / 2 (Heitz 2014); checked with D = 1/pi
  (n.h = 1) and F = 1."""
  for alpha in (0.05, 0.3, 0.8):
    for nov in (0.1, 0.5, 1.0):
      for nol in (0.2, 0.7, 1.0):
        lam = [(-1 + math.sqrt(1 + alpha**2 * (1 - c * c) / (c * c))) / 2 for c in (nov, nol)]
        g2 = 1.0 / (1.0 + lam[0] + lam[1])
        want = (1.0 / (math.pi * alpha**2)) * g2 / (4 * nov * nol)
        got = oracle.spec_lobe((1, 1, 1), alpha, 1.0, nov, nol, 1.0)[0]
        assert got == pytest.approx(want, rel=3e-5), (alpha, nov, nol)


def test_schlick_fresnel_end_points():
  """brdf.py:49-55: F(l.h = 1) = f0 and F(l.h = 0) = 1, F = f0 + (1 - f0)(1 - l.h)^5."""
  f0 = (0.04, 0.5, 0.9)
  ref = oracle.spec_lobe((1, 1, 1), 0.3, 1.0, 1.0, 1.0, 1.0)[0]
  np.testing.assert_allclose(oracle.spec_lobe(f0, 0.3, 1.0, 1.0, 1.0, 1.0) / ref, f0, rtol=1e-6)
  np.testing.assert_allclose(oracle.spec_lobe(f0, 0.3, 1.0, 1.0, 1.0, 0.0) / ref, 1.0, rtol=1e-6)
  for loh in (0.2, 0.6):
    want = np.array(f0) + (1 - np.array(f0)) * (1 - loh) ** 5
    np.testing.assert_allclose(oracle.spec_lobe(f0, 0.3, 1.0, 1.0, 1.0, loh) / ref, want, rtol=1e-5)


def test_dfg_lookup_is_bilinear_at_texel_centres():
  """ibl.py:117-119: a clamped bilinear texture read at (u = NoV, v = perceptual roughness) of
  a table indexed [roughness][NoV]; bilinear filtering reproduces a linear ramp exactly."""
  n = 8
  j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
  lut = np.stack([i, j], -1).astype(np.float32)
  for nov, pr in ((0.5, 0.5), (0.1, 0.9), (0.03, 0.97), (0.33, 0.61), (1.0, 0.0)):
    want = (np.clip(nov * n - 0.5, 0, n - 1), np.clip(pr * n - 0.5, 0, n - 1))
    np.testing.assert_allclose(oracle.dfg_lookup(lut, nov, pr), want, atol=1e-5)


def test_shading_lobes_follow_filament():
  """material.py:506-539 / brdf.py:58-110 on a constant DFG table (x, y) = (0.3, 0.8): roughness
  clamp 0.045, alpha = pr^2, dielectric F0 = 0.16 r^2 (4 % at r = 0.5), metals tint F0 by the
  base colour, E = (1 - F0) x + F0 y, compensation 1 + F0 (1 / y - 1), diffuse base (1 - m)(1 - E)."""
  lut = np.tile(np.array([0.3, 0.8], np.float32), (4, 4, 1))
  lb = oracle.lobes((0.2, 0.4, 0.6), 0.0, 0.5, 0.01, 0.5, lut)
  assert lb["pr"] == pytest.approx(0.045) and lb["alpha"] == pytest.approx(0.045**2)
  np.testing.assert_allclose(lb["f0"], 0.04, rtol=1e-6)
  e = (1 - 0.04) * 0.3 + 0.04 * 0.8
  np.testing.assert_allclose(lb["spec_albedo"], e, rtol=1e-6)
  np.testing.assert_allclose(lb["energy_comp"], 1 + 0.04 * (1 / 0.8 - 1), rtol=1e-6)
  np.testing.assert_allclose(lb["diffuse"], np.array([0.2, 0.4, 0.6]) * (1 - e), rtol=1e-6)
  metal = oracle.lobes((0.9, 0.6, 0.3), 1.0, 0.5, 0.7, 0.5, lut)
  np.testing.assert_allclose(metal["f0"], (0.9, 0.6, 0.3), rtol=1e-6)
  np.testing.assert_allclose(metal["diffuse"], 0.0, atol=0)
  half = oracle.lobes((1.0, 1.0, 1.0), 0.5, 1.0, 0.7, 0.5, lut)
  np.testing.assert_allclose(half["f0"], 0.5 + 0.16 * 0.5, rtol=1e-6)
