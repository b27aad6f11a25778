"""MI355X-native batched PBR renderer with the API of mujoco_warp_pbr
(reference mujoco_warp_pbr/__init__.py:18-54).

    mjm = scenes.full_scene(seed=0)                 # host model (MjModel fields)
    m, d = put_model(mjm), put_data(mjm, nworld)    # device model / data
    rc = create_render_context(mjm, nworld=nworld, cam_res=(128, 128), ...)
    refit_bvh(m, d, rc)                             # after the poses change
    render(m, d, rc, seed=0)
    get_rgb(rc, 0, out)

All rendering runs in hand-written HIP kernels (libmjpbr_hip.so, gfx950) called through a C
ABI with ctypes; torch tensors hold the buffers.
"""

from mujoco_warp_pbr_amd.envmap import load_hdr
from mujoco_warp_pbr_amd.io import create_render_context
from mujoco_warp_pbr_amd.model import HostModel
from mujoco_warp_pbr_amd.model import put_data
from mujoco_warp_pbr_amd.model import put_model
from mujoco_warp_pbr_amd.output import get_albedo
from mujoco_warp_pbr_amd.output import get_depth
from mujoco_warp_pbr_amd.output import get_hdr
from mujoco_warp_pbr_amd.output import get_normal
from mujoco_warp_pbr_amd.output import get_rgb
from mujoco_warp_pbr_amd.output import get_segmentation
from mujoco_warp_pbr_amd.output import albedo_view
from mujoco_warp_pbr_amd.output import bgra_view
from mujoco_warp_pbr_amd.output import depth_view
from mujoco_warp_pbr_amd.output import hdr_view
from mujoco_warp_pbr_amd.output import normal_view
from mujoco_warp_pbr_amd.output import packed_rgb_view
from mujoco_warp_pbr_amd.output import segmentation_view
from mujoco_warp_pbr_amd.rays import CameraRays
from mujoco_warp_pbr_amd.rays import set_camera_rays
from mujoco_warp_pbr_amd.render import count_path_rays
from mujoco_warp_pbr_amd.render import render
from mujoco_warp_pbr_amd.bvh import refit_bvh
from mujoco_warp_pbr_amd.types import LightModel
from mujoco_warp_pbr_amd.types import PbrRenderContext
from mujoco_warp_pbr_amd.types import RenderOrder
from mujoco_warp_pbr_amd.types import Tier
from mujoco_warp_pbr_amd.types import ToneMapper
from mujoco_warp_pbr_amd.types import UnsetMaterial
from mujoco_warp_pbr_amd.wavefront import WavefrontBatch
from mujoco_warp_pbr_amd.wavefront import WavefrontConfig
from mujoco_warp_pbr_amd.wavefront import render_batch

__all__ = [
  "CameraRays",
  "HostModel",
  "LightModel",
  "PbrRenderContext",
  "RenderOrder",
  "Tier",
  "ToneMapper",
  "UnsetMaterial",
  "WavefrontBatch",
  "WavefrontConfig",
  "create_render_context",
  "get_albedo",
  "get_depth",
  "get_hdr",
  "get_normal",
  "get_rgb",
  "get_segmentation",
  "albedo_view",
  "bgra_view",
  "depth_view",
  "hdr_view",
  "normal_view",
  "packed_rgb_view",
  "segmentation_view",
  "load_hdr",
  "put_data",
  "put_model",
  "refit_bvh",
  "count_path_rays",
  "render",
  "render_batch",
  "set_camera_rays",
]
