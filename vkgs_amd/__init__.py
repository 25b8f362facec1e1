"""MI355X-native (gfx950) implementation of vkgs's splat pipeline.

Layers:
  include/vkgs_hip.h   C ABI (one entry point per reference stage + a renderer)
  vkgs_amd/csrc        HIP kernels and C++ host code behind it
  vkgs_amd._lib        ctypes binding; raises LibraryNotBuilt when the library is absent
  vkgs_amd.pipeline    stage-by-stage Python mirror on torch device memory
  vkgs_amd.scene       synthetic 3DGS scenes, PLY bytes, orbit cameras
  vkgs_amd.pygs        the reference's pygs surface (show/load/close) plus render
"""
from ._lib import Camera, LibraryNotBuilt, RasterConfig, VkgsError, lib  # noqa: F401

__all__ = ["Camera", "LibraryNotBuilt", "RasterConfig", "VkgsError", "lib"]
