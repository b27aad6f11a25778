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
  """(ngeom, 4) rgba each geom renders with: material colour when the geom keeps the default."""
  rgba = np.array(mjm.geom_rgba, dtype=np.float64).reshape(-1, 4)
  if not mjm.nmat:
    return rgba
  inherit = (mjm.geom_matid >= 0) & np.all(np.isclose(rgba, _DEFAULT_GEOM_RGBA), axis=1)
  rgba[inherit] = mjm.mat_rgba[mjm.geom_matid[inherit]]
  return rgba


def _heuristic(rgba, shininess: float, has_material: bool) -> tuple[float, float, float]:
  """Colour-based (metallic, roughness, reflectance) guess (material.py:106-129)."""
  r, g, b = rgba[:3]
  if not has_material:
    brightness = (r + g + b) / 3.0
    if brightness < 0.25:
      return 0.4, 0.45, 0.35
    if brightness > 0.8 or (r > 0.7 and g < 0.3):
      return 0.0, 0.4, 0.25
    return 0.0, 0.5, 0.2
  lum = float(np.dot(rgba[:3], LUMINANCE_WEIGHTS))
  chroma = max(r, g, b) - min(r, g, b)
  metallic, roughness, reflectance = 0.0, 0.42, 0.6
  if lum > 0.75 and chroma < 0.10:
    roughness, reflectance = 0.38, 0.65
  elif lum < 0.28 and chroma < 0.12:
    roughness, reflectance = 0.28, 0.6
  elif chroma < 0.08 and 0.35 <= lum <= 0.75:
    metallic, roughness, reflectance = 0.6, 0.40, 0.7
  else:
    roughness = 0.40
  if shininess > 0.0 and shininess != _DEFAULT_SHININESS:
    roughness = float(np.clip(roughness * (1.3 - shininess), 0.06, 1.0))
  return metallic, roughness, reflectance


def resolve_geom_materials(mjm: mdl.HostModel, unset_material: int = UnsetMaterial.GLTF, default_material: dict | None = None):
  """(metallic, perceptual roughness, reflectance) per geom, float32 (ngeom,) each
  (material.py:132-182). Authored values win; unset ones (-1) follow `default_material`, else
  the `unset_material` policy."""
  default_material = default_material or {}
  n = mjm.ngeom
  metallic = np.zeros(n, np.float32)
  roughness = np.zeros(n, np.float32)
  reflectance = np.zeros(n, np.float32)
  rgba_all = effective_geom_rgba(mjm)
  for g in range(n):
    mat = int(mjm.geom_matid[g])
    has_material = mat >= 0
    shininess = float(mjm.mat_shininess[mat]) if has_material else _DEFAULT_SHININESS
    specular = float(mjm.mat_specular[mat]) if has_material else _DEFAULT_SPECULAR
    authored_m = float(mjm.mat_metallic[mat]) if has_material else -1.0
    authored_r = float(mjm.mat_roughness[mat]) if has_material else -1.0
    if unset_material == UnsetMaterial.HEURISTIC and authored_m < 0.0 and authored_r < 0.0:
      m, r, f = _heuristic(rgba_all[g], shininess, has_material)
      if mjm.geom_type[g] == mdl.GEOM_PLANE:
        m, r, f = 0.0, 1.0, 0.0
    else:
      m = 0.0
      r = float(np.clip(1.0 - shininess, 0.0, 1.0))
      f = specular
    if authored_m < 0.0 and "metallic" in default_material:
      m = default_material["metallic"]
    if authored_r < 0.0 and "roughness" in default_material:
      r = default_material["roughness"]
    metallic[g] = authored_m if authored_m >= 0.0 else m
    roughness[g] = authored_r if authored_r >= 0.0 else r
    reflectance[g] = f
  return metallic, roughness, reflectance


class SurfaceFeatures:
  """Which optional material inputs the rendered geoms use (material.py:185-221)."""

  def __init__(self, mjm: mdl.HostModel, use_textures: bool, specular_aa: bool, mats: np.ndarray):
    texid = mjm.mat_texid[mats] if mjm.nmat else np.full((0, mdl.NTEXROLE), -1)

    def has(*roles):
      return use_textures and bool(np.any(texid[:, list(roles)] >= 0))

    self.base_color_map = has(mdl.ROLE_RGB, mdl.ROLE_RGBA)
    self.normal_map = has(mdl.ROLE_NORMAL)
    self.orm_map = has(mdl.ROLE_ORM, mdl.ROLE_OCCLUSION, mdl.ROLE_ROUGHNESS, mdl.ROLE_METALLIC)
    self.emissive_map = has(mdl.ROLE_EMISSIVE)
    self.opacity_map = has(mdl.ROLE_OPACITY)
    self.textures = self.base_color_map or self.normal_map or self.orm_map or self.emissive_map or self.opacity_map
    self.alpha_map = False
    if use_textures:
      for role, channel in ((mdl.ROLE_RGBA, 3), (mdl.ROLE_OPACITY, 0)):
        for t in np.unique(texid[:, role]):
          if t < 0:
            continue
          w, h, nc = int(mjm.tex_width[t]), int(mjm.tex_height[t]), int(mjm.tex_nchannel[t])
          data = mjm.tex_data[mjm.tex_adr[t] : mjm.tex_adr[t] + w * h * nc].reshape(-1, nc)
          if channel < nc and np.any(data[:, channel] < 255):
            self.alpha_map = True
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
