"""Per-pixel camera rays (ray bundles) replacing the cameras' pinhole projection
(reference mujoco_warp_pbr/_src/rays.py:84-143; kernels: csrc/camera.hpp bundle_ray / bundle_cone,
reference camera.py:194-347)."""

from __future__ import annotations

from collections.abc import Sequence
from typing import Any, Union

import numpy as np
import torch

from mujoco_warp_pbr_amd.types import PbrRenderContext

# Rays of every rendering camera: (nrender, height, width, 2, 3) when the cameras share a
# resolution, else one (height, width, 2, 3) array per camera; [..., 0, :] is the camera-frame
# origin, [..., 1, :] the direction. NumPy arrays or torch tensors.
CameraRays = Union[np.ndarray, torch.Tensor, Sequence]


def _images(camera_rays) -> list:
  if isinstance(camera_rays, (np.ndarray, torch.Tensor)):
    if camera_rays.ndim != 5 or tuple(camera_rays.shape[-2:]) != (2, 3):
      raise ValueError(f"camera_rays as one array must have shape (ncam, height, width, 2, 3), got {tuple(camera_rays.shape)}")
    return [camera_rays[i] for i in range(camera_rays.shape[0])]
  return list(camera_rays)


def _as_tensor(image, device) -> torch.Tensor:
  if isinstance(image, np.ndarray):
    if not np.issubdtype(image.dtype, np.floating):
      raise ValueError(f"camera rays must be floating point, got {image.dtype}")
    image = torch.from_numpy(np.ascontiguousarray(image, dtype=np.float32))
  elif isinstance(image, torch.Tensor):
    if not image.is_floating_point():
      raise ValueError(f"camera rays must be floating point, got {image.dtype}")
  else:
    raise TypeError(f"camera rays must be NumPy arrays or torch tensors, got {type(image).__name__}")
  if image.ndim != 4 or tuple(image.shape[-2:]) != (2, 3):
    raise ValueError(f"a camera's rays must have shape (height, width, 2, 3), got {tuple(image.shape)}")
  return image.to(device=device, dtype=torch.float32)


def set_camera_rays(rc: PbrRenderContext, camera_rays: CameraRays | None) -> None:
  """Renders the context's cameras from per-pixel rays instead of their pinhole projection
  (reference rays.py:84-143, same conventions).

  Each rendering camera gets one ray per pixel in its camera frame (-z forward, +y up, +x
  right), placed in the world by the camera's pose (per world). A zero direction marks a pixel
  without a ray (background, depth 0). Depth is the distance along each ray and zfar bounds it.
  Sub-pixel samples interpolate between neighbouring rays; texture filtering takes each pixel's
  footprint from them. The rays are copied: call again after changing them; None renders the
  pinhole cameras again.

  Raises:
    ValueError: for a camera count, resolution or dtype that does not match the context.
    TypeError: for rays that are neither NumPy arrays nor torch tensors.
  """
  if camera_rays is None:
    rc.ray_origin = rc.ray_direction = None
    return
  base = rc.base
  images = _images(camera_rays)
  if len(images) != base.nrender:
    raise ValueError(f"camera_rays must have rays for each of the {base.nrender} rendering cameras, got {len(images)}")
  device = rc.hdr_data.device
  tensors = []
  for i, image in enumerate(images):
    t = _as_tensor(image, device)
    width, height = int(base.cam_res[i][0]), int(base.cam_res[i][1])
    if tuple(t.shape[:2]) != (height, width):
      raise ValueError(f"the rays of rendering camera {i} must have shape ({height}, {width}, 2, 3) (its resolution), got {tuple(t.shape)}")
    tensors.append(t.reshape(height * width, 2, 3))
  rays = torch.cat(tensors, dim=0)
  origin = rays[:, 0, :].contiguous()
  direction = rays[:, 1, :].contiguous()
  if rc.ray_direction is not None and tuple(rc.ray_direction.shape) == tuple(direction.shape):
    rc.ray_origin.copy_(origin)
    rc.ray_direction.copy_(direction)
  else:
    rc.ray_origin, rc.ray_direction = origin, direction
