"""Host-side material and lighting resolution (reference material.py:93-222, io.py:223-300,
lights.py:134-148, color.py:35-87)."""

from __future__ import annotations

import dataclasses
import math

import numpy as np

from mujoco_warp_pbr_amd import model as mdl
from mujoco_warp_pbr_amd.types import LightModel
from mujoco_warp_pbr_amd.types import UnsetMaterial

LUMINANCE_WEIGHTS = (0.2126, 0.7152, 0.0722)
_DEFAULT_GEOM_RGBA = (0.5, 0.5, 0.5, 1.0)
_DEFAULT_SHININESS = 0.5
_DEFAULT_SPECULAR = 0.5


def exposure_from_ev100(ev100: float) -> float:
  """Photometric exposure for a camera EV100, 1 / (1.2 * 2^EV100) (color.py:35-37)."""
  return 1.0 / (1.2 * math.pow(2.0, ev100))


def srgb_to_linear(x) -> np.ndarray:
  """sRGB display values in [0, 1] -> linear (color.py:102-106)."""
  x = np.asarray(x, dtype=np.float64)
  return np.where(x <= 0.04045, x / 12.92, ((x + 0.055) / 1.055) ** 2.4)


def effective_geom_rgba(mjm: mdl.HostModel) -> np.ndarray:
  """(ngeom, 4) rgba each geom renders with (MuJoCo Warp's _geom_rgba on the host): a geom that
  has a material and still carries MuJoCo's default rgba (0.5, 0.5, 0.5, 1) shows the
  material's colour; any other geom its own."""
  own = np.asarray(mjm.geom_rgba, np.float64).reshape(-1, 4)
  if not mjm.nmat:
    return own.copy()
  default = np.asarray(_DEFAULT_GEOM_RGBA)
  untouched = np.all(np.abs(own - default) <= 1e-8 + 1e-5 * default, axis=1)
  matid = np.asarray(mjm.geom_matid)
  from_material = untouched & (matid >= 0)
  mat_rgba = np.asarray(mjm.mat_rgba, np.float64).reshape(-1, 4)
  return np.where(from_material[:, None], mat_rgba[np.maximum(matid, 0)], own)


# The colour heuristics of the HEURISTIC unset-material policy (reference material.py:106-129),
# as data: the first matching class gives (metallic, perceptual roughness, reflectance).
# Geoms without a material, by the mean of r, g, b:
_PLAIN_CLASSES = (
  ("dark", lambda mean, r, g: mean < 0.25, (0.4, 0.45, 0.35)),
  ("bright or red", lambda mean, r, g: (mean > 0.8) | ((r > 0.7) & (g < 0.3)), (0.0, 0.40, 0.25)),
  ("other", lambda mean, r, g: np.ones_like(mean, bool), (0.0, 0.50, 0.20)),
)
# Geoms with a material, by luminance and chroma (max - min of r, g, b):
_MATERIAL_CLASSES = (
  ("white", lambda lum, chroma: (lum > 0.75) & (chroma < 0.10), (0.0, 0.38, 0.65)),
  ("black", lambda lum, chroma: (lum < 0.28) & (chroma < 0.12), (0.0, 0.28, 0.60)),
  ("grey metal", lambda lum, chroma: (chroma < 0.08) & (lum >= 0.35) & (lum <= 0.75), (0.6, 0.40, 0.70)),
  ("coloured", lambda lum, chroma: np.ones_like(lum, bool), (0.0, 0.40, 0.60)),
)
# A material shininess other than MuJoCo's default scales that roughness by (1.3 - shininess),
# clamped to [0.06, 1].
_ROUGHNESS_CLAMP = (0.06, 1.0)


def _first_match(classes, *keys) -> np.ndarray:
  """(n, 3) values of the first class whose predicate holds for each entry."""
  tests = [pred(*keys) for _, pred, _ in classes]
  table = np.array([vals for _, _, vals in classes], np.float64)
  return table[np.argmax(np.stack(tests), axis=0)]


def heuristic_materials(rgba: np.ndarray, shininess: np.ndarray, has_material: np.ndarray) -> np.ndarray:
  """(n, 3) (metallic, roughness, reflectance) guessed from flat colours (HEURISTIC policy)."""
  rgb = np.asarray(rgba, np.float64)[:, :3]
  r, g = rgb[:, 0], rgb[:, 1]
  plain = _first_match(_PLAIN_CLASSES, rgb.mean(axis=1), r, g)
  lum = rgb @ np.asarray(LUMINANCE_WEIGHTS)
  chroma = rgb.max(axis=1) - rgb.min(axis=1)
  with_mat = _first_match(_MATERIAL_CLASSES, lum, chroma)
  glossy = (shininess > 0.0) & (shininess != _DEFAULT_SHININESS)
  with_mat[:, 1] = np.where(glossy, np.clip(with_mat[:, 1] * (1.3 - shininess), *_ROUGHNESS_CLAMP), with_mat[:, 1])
  return np.where(np.asarray(has_material)[:, None], with_mat, plain)


def resolve_geom_materials(mjm: mdl.HostModel, unset_material: int = UnsetMaterial.GLTF, default_material: dict | None = None):
  """(metallic, perceptual roughness, reflectance) per geom, float32 (ngeom,) each (reference
  material.py:132-182). Authored material values win. Unset ones (MuJoCo's -1) come from
  `default_material` when it names them, else from the policy: GLTF gives metallic 0, roughness
  1 - shininess and reflectance = the material's specular; HEURISTIC (for geoms whose metallic
  and roughness are both unset) the colour classes above, and matte ground planes."""
  default_material = default_material or {}
  matid = np.asarray(mjm.geom_matid)
  has = matid >= 0
  m_idx = np.maximum(matid, 0)

  def per_material(field, fallback):
    values = np.asarray(getattr(mjm, field), np.float64) if mjm.nmat else np.zeros(1)
    return np.where(has, values[m_idx] if mjm.nmat else fallback, fallback)

  shininess = per_material("mat_shininess", _DEFAULT_SHININESS)
  specular = per_material("mat_specular", _DEFAULT_SPECULAR)
  authored = np.stack([per_material("mat_metallic", -1.0), per_material("mat_roughness", -1.0)], axis=1)
  unset = authored < 0.0
  # the policy's guess for every geom: metallic, roughness, reflectance
  guess = np.stack([np.zeros_like(shininess), np.clip(1.0 - shininess, 0.0, 1.0), specular], axis=1)
  if unset_material == UnsetMaterial.HEURISTIC:
    heuristic = heuristic_materials(effective_geom_rgba(mjm), shininess, has)
    heuristic[np.asarray(mjm.geom_type) == mdl.GEOM_PLANE] = (0.0, 1.0, 0.0)
    use = unset.all(axis=1)
    guess = np.where(use[:, None], heuristic, guess)
  for k, key in enumerate(("metallic", "roughness")):
    if key in default_material:
      guess[:, k] = np.where(unset[:, k], float(default_material[key]), guess[:, k])
  resolved = np.where(unset, guess[:, :2], authored)
  return resolved[:, 0].astype(np.float32), resolved[:, 1].astype(np.float32), guess[:, 2].astype(np.float32)


def _texels_below_opaque(mjm: mdl.HostModel, tex_ids, channel: int) -> bool:
  """True when some texture of `tex_ids` has a texel below 255 in `channel` (when it has it)."""
  for t in set(int(x) for x in np.asarray(tex_ids).ravel() if x >= 0):
    nc = int(mjm.tex_nchannel[t])
    if channel >= nc:
      continue
    n = int(mjm.tex_width[t]) * int(mjm.tex_height[t]) * nc
    texels = np.asarray(mjm.tex_data[int(mjm.tex_adr[t]) : int(mjm.tex_adr[t]) + n])
    if texels[channel::nc].min(initial=255) < 255:
      return True
  return False


class SurfaceFeatures:
  """Which optional material inputs the rendered materials `mats` use (reference
  material.py:185-221): texture roles (when textures are on), whether any alpha texture is not
  fully opaque (translucent layers), and specular anti-aliasing."""

  def __init__(self, mjm: mdl.HostModel, use_textures: bool, specular_aa: bool, mats: np.ndarray):
    texid = np.asarray(mjm.mat_texid)[mats] if mjm.nmat else np.full((0, mdl.NTEXROLE), -1)
    used = (texid >= 0).any(axis=0) if use_textures and len(texid) else np.zeros(mdl.NTEXROLE, bool)
    self.base_color_map = bool(used[mdl.ROLE_RGB] or used[mdl.ROLE_RGBA])
    self.normal_map = bool(used[mdl.ROLE_NORMAL])
    self.orm_map = bool(used[[mdl.ROLE_ORM, mdl.ROLE_OCCLUSION, mdl.ROLE_ROUGHNESS, mdl.ROLE_METALLIC]].any())
    self.emissive_map = bool(used[mdl.ROLE_EMISSIVE])
    self.opacity_map = bool(used[mdl.ROLE_OPACITY])
    self.textures = self.base_color_map or self.normal_map or self.orm_map or self.emissive_map or self.opacity_map
    # RGBA textures carry alpha in channel 3, opacity textures in channel 0
    self.alpha_map = bool(use_textures and len(texid)) and (
      _texels_below_opaque(mjm, texid[:, mdl.ROLE_RGBA], 3) or _texels_below_opaque(mjm, texid[:, mdl.ROLE_OPACITY], 0)
    )
    self.specular_aa = specular_aa


def fallback_intensity(mjm: mdl.HostModel, light_id: int, exposure: float) -> float:
  """Physical intensity of a light with MuJoCo intensity 0 (lights.py:134-148)."""
  illuminance = math.pi / exposure
  if mjm.light_type[light_id] == mdl.LIGHT_DIRECTIONAL:
    return illuminance
  distance = float(np.linalg.norm(mjm.light_xpos[light_id].astype(np.float64) - mjm.stat_center))
  distance = max(distance, 0.1 * float(mjm.stat_extent), 1e-3)
  return illuminance * distance * distance


@dataclasses.dataclass
class Lighting:
  exposure: float
  light_intensity: np.ndarray
  headlight_intensity: float
  ambient: np.ndarray


def resolve_lighting(mjm: mdl.HostModel, light_model: int, ev100: float) -> Lighting:
  """Exposure, light intensities and ambient radiance (io.py:223-248)."""
  physical = light_model == LightModel.PHYSICAL
  exposure = exposure_from_ev100(ev100) if physical else 1.0
  light_intensity = np.array(mjm.light_intensity, dtype=np.float32)
  for i in range(mjm.nlight):
    if light_intensity[i] > 0.0:
      light_intensity[i] /= max(float(np.max(mjm.light_diffuse[i])), 1e-6)
    else:
      light_intensity[i] = fallback_intensity(mjm, i, exposure)
  ambient = np.zeros(3, dtype=np.float32)
  if mjm.headlight_active:
    ambient += mjm.headlight_ambient
  if physical:
    ambient /= exposure
  elif mjm.nlight:
    ambient += np.sum(mjm.light_ambient[mjm.light_active.astype(bool)], axis=0)
  return Lighting(exposure, light_intensity, math.pi, ambient)
