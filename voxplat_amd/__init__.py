"""voxplat_amd: MI355X (gfx950) voxel pipeline -- remesh (cull + splat list +
LOD), RLE codec and splat resolve as HIP kernels behind the C ABI of
include/vp_gpu.h.

    from voxplat_amd import GPU, abi, workload
"""
from . import abi, workload
from ._lib import (GPU, FrameGraph, RemeshGraph, VPError, frustum_planes, host_shadow_spheres, load_library,
                   octree_cull, world_read_header)

__all__ = ["GPU", "FrameGraph", "RemeshGraph", "VPError", "abi", "workload", "load_library", "host_shadow_spheres",
           "frustum_planes", "octree_cull", "world_read_header"]
