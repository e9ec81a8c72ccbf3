"""ctypes mirror of include/vp_gpu.h (structs, constants, prototypes)."""
from __future__ import annotations

import ctypes as C

VP_OK = 0
VP_ERR_ARG = -1
VP_ERR_CAPACITY = -2
VP_ERR_HIP = -3
VP_ERR_UNSUPPORTED = -4
VP_ERR_DATA = -5
VP_ERR_NOMEM = -6

VP_STATUS_OVERFLOW = 1
VP_STATUS_BAD_RLE = 2
VP_STATUS_BAD_ID = 4

VP_CM_LOD = 1
VP_CM_RECORD_V2 = 2
VP_CM_COUNT_ONLY = 4
VP_NOALLOC_MAX_CHUNKS = 524288

VP_RLE_ENCODE = 0
VP_RLE_DECODE = 1


class WorldDesc(C.Structure):
    """vp_world_desc: ChunkSet::root_bitw / max_bitw (src/chunkset.h:36-40)."""
    _fields_ = [("root_bitw", C.c_uint8), ("max_bitw", C.c_uint8 * 3)]

    @classmethod
    def make(cls, root_bitw: int, max_bitw) -> "WorldDesc":
        w = cls()
        w.root_bitw = root_bitw
        for i in range(3):
            w.max_bitw[i] = max_bitw[i]
        return w

    @property
    def R(self) -> int:
        return 1 << self.root_bitw

    @property
    def chunks(self) -> int:
        return 1 << (self.max_bitw[0] + self.max_bitw[1] + self.max_bitw[2])

    @property
    def size(self):
        return tuple((1 << self.max_bitw[i]) << self.root_bitw for i in range(3))

    def shadow_dims(self):
        X, Y, Z = self.size
        return X + Y, Z


def shadow_padded_len(size_x: int, size_z: int) -> int:
    return size_x * (size_z + 17) + 64


class RemeshShape(C.Structure):
    _fields_ = [("block", C.c_int), ("wg_per_cu", C.c_int), ("records_in_flight", C.c_int)]


class ShadowDesc(C.Structure):
    _fields_ = [("d_map", C.c_void_p), ("size_x", C.c_uint32), ("size_z", C.c_uint32)]


class SplatOut(C.Structure):
    _fields_ = [("d_splats", C.c_void_p), ("capacity_records", C.c_uint64),
                ("d_chunk_offsets", C.c_void_p), ("d_level_counts", C.c_void_p)]


class MeshOut(C.Structure):
    _fields_ = [("d_geometry", C.c_void_p), ("d_index", C.c_void_p),
                ("capacity_faces", C.c_uint64), ("d_face_offsets", C.c_void_p)]


class NodeOut(C.Structure):
    _fields_ = [("d_node_splats", C.c_void_p), ("capacity_records", C.c_uint64),
                ("d_node_offsets", C.c_void_p)]


class Draw(C.Structure):
    """vp_draw: one node draw of a frame."""
    _fields_ = [("level", C.c_uint32), ("node", C.c_uint32), ("first", C.c_uint64)]


def node_bits(w: "WorldDesc", level: int, axis: int) -> int:
    b = w.max_bitw[axis]
    return b - min(level, b)


def node_count(w: "WorldDesc", level: int) -> int:
    return 1 << (node_bits(w, level, 0) + node_bits(w, level, 1) + node_bits(w, level, 2))


class Camera(C.Structure):
    """vp_camera: column-major matrices as uploaded by gfx_svl_draw (vsplat.c:126-158)."""
    _fields_ = [("view_translation", C.c_float * 16), ("proj_view_rot", C.c_float * 16),
                ("inv_proj_view_rot", C.c_float * 16), ("width", C.c_uint32),
                ("height", C.c_uint32), ("far_plane", C.c_float)]


class Frame(C.Structure):
    _fields_ = [("d_depth", C.c_void_p), ("d_id", C.c_void_p), ("d_rgb", C.c_void_p),
                ("d_weight", C.c_void_p)]


# name -> (restype, argtypes) of every extern "C" symbol in include/vp_gpu.h
PROTOTYPES = {
    "vp_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "vp_ctx_destroy": (C.c_int, [C.c_void_p]),
    "vp_strerror": (C.c_char_p, [C.c_int]),
    "vp_ctx_status": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)]),
    "vp_ctx_num_cus": (C.c_int, [C.c_void_p]),
    "vp_ctx_set_splat_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_uint64]),
    "vp_gpu_cull_mesh": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p, C.c_void_p,
                                   C.POINTER(ShadowDesc), C.c_void_p, C.c_uint32, C.c_void_p,
                                   C.c_uint32, C.POINTER(SplatOut), C.c_void_p]),
    "vp_gpu_cull_mesh_noalloc": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p,
                                           C.c_void_p, C.POINTER(ShadowDesc), C.c_void_p,
                                           C.c_uint32, C.c_void_p, C.c_uint32,
                                           C.POINTER(SplatOut), C.c_void_p]),
    "vp_gpu_mesh": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p, C.c_void_p,
                              C.POINTER(ShadowDesc), C.c_void_p, C.c_uint32,
                              C.POINTER(MeshOut), C.c_void_p]),
    "vp_gpu_make_mask": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "vp_gpu_lod_mip": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p,
                                 C.c_uint32, C.c_void_p]),
    "vp_gpu_rle": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(WorldDesc), C.c_void_p, C.c_void_p,
                             C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                             C.c_void_p]),
    "vp_gpu_splat": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64,
                               C.POINTER(Camera), C.c_void_p, C.POINTER(Frame), C.c_void_p]),
    "vp_gpu_splat_backward": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(Camera), C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vp_gpu_splat_bounds": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64,
                                      C.c_void_p, C.c_void_p]),
    "vp_gpu_splat_blocks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64,
                                      C.c_void_p, C.POINTER(Camera), C.c_void_p, C.POINTER(Frame),
                                      C.c_void_p]),
    "vp_gpu_svl_nodes": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_uint32, C.c_void_p,
                                   C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                   C.POINTER(NodeOut), C.c_void_p]),
    "vp_gpu_svl_nodes_all": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p, C.c_void_p,
                                       C.c_uint32, C.c_void_p, C.c_void_p, NodeOut * 5,
                                       C.c_void_p]),
    "vp_gpu_draw_gather": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p * 5,
                                     C.c_void_p * 5, C.c_void_p, C.c_uint32, C.c_void_p,
                                     C.c_void_p, C.c_uint64, C.c_void_p]),
}
VP_SPLAT_BLOCK = 256


PROTOTYPES.update({
    "vp_gpu_octree_cull": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.POINTER(C.c_float),
                                     C.POINTER(C.c_float), C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "vp_gpu_draw_gather_counted": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "vp_gpu_splat_counted": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                       C.c_uint64, C.c_void_p, C.c_void_p, C.POINTER(Camera),
                                       C.c_void_p, C.POINTER(Frame), C.c_void_p]),
})


# ---- include/vp_workload.h ---------------------------------------------------
class GenParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("period", C.c_uint32), ("hmin", C.c_int32),
                ("hmax", C.c_int32)]


class Sphere(C.Structure):
    _fields_ = [("cx", C.c_int32), ("cy", C.c_int32), ("cz", C.c_int32),
                ("radius", C.c_uint32), ("voxel", C.c_uint32)]


PROTOTYPES.update({
    "vp_gpu_generate": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.POINTER(GenParams),
                                  C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "vp_gpu_classify": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p, C.c_uint32,
                                  C.c_uint32, C.c_void_p, C.c_void_p]),
    "vp_gpu_shadow_build": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "vp_gpu_edit_spheres": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p, C.c_void_p,
                                      C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vp_host_shadow_spheres": (C.c_int, [C.POINTER(WorldDesc), C.c_void_p, C.c_void_p,
                                         C.c_uint32]),
    "vp_gpu_shadow_spheres": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p, C.c_void_p,
                                        C.c_uint32, C.c_void_p]),
})


# ---- include/vp_host.h ---------------------------------------------------------
class FrameOut(C.Structure):
    _fields_ = [("d_dirty_ids", C.c_void_p), ("d_idle_ids", C.c_void_p), ("d_counts", C.c_void_p),
                ("remesh", SplatOut), ("d_words", C.c_void_p), ("capacity_words", C.c_uint64),
                ("d_word_offsets", C.c_void_p), ("d_idle_uniform", C.c_void_p)]


PROTOTYPES.update({
    "vp_host_frame_graph_create": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p,
                                             C.c_void_p, C.POINTER(ShadowDesc), C.c_void_p,
                                             C.c_uint32, C.POINTER(FrameOut),
                                             C.POINTER(C.c_void_p)]),
    "vp_host_frame_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vp_host_frame_graph_destroy": (C.c_int, [C.c_void_p]),
    "vp_ctx_reserve": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64]),
    "vp_ctx_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "vp_ctx_get_launch": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vp_ctx_set_remesh_shape": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc),
                                          C.POINTER(RemeshShape)]),
    "vp_ctx_get_remesh_shape": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc),
                                          C.POINTER(RemeshShape)]),
    "vp_ctx_remesh_resident": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_uint32,
                                         C.POINTER(RemeshShape), C.POINTER(C.c_int)]),
    "vp_host_remesh_per_chunk": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p,
                                           C.c_void_p, C.POINTER(ShadowDesc), C.c_void_p,
                                           C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vp_host_remesh_graph_create": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p,
                                              C.c_void_p, C.POINTER(ShadowDesc), C.c_void_p,
                                              C.c_uint32, C.c_void_p, C.c_uint32,
                                              C.POINTER(SplatOut), C.POINTER(C.c_void_p)]),
    "vp_host_remesh_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vp_host_world_save": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(WorldDesc), C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]),
    "vp_host_world_read_header": (C.c_int, [C.c_char_p, C.POINTER(WorldDesc)]),
    "vp_host_world_load": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(WorldDesc), C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p]),
    "vp_host_remesh_graph_destroy": (C.c_int, [C.c_void_p]),
    "vp_host_autotune_remesh": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_void_p,
                                          C.c_void_p, C.POINTER(ShadowDesc), C.c_void_p,
                                          C.c_uint32, C.c_uint32, C.POINTER(SplatOut), C.c_int,
                                          C.c_void_p, C.POINTER(RemeshShape),
                                          C.POINTER(C.c_float), C.POINTER(C.c_int)]),
    "vp_host_frustum_planes": (None, [C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "vp_host_octree_cull": (C.c_int, [C.POINTER(WorldDesc), C.POINTER(C.c_float),
                                      C.POINTER(C.c_float), C.c_void_p * 5, C.c_void_p,
                                      C.POINTER(Draw), C.c_uint32, C.POINTER(C.c_uint32),
                                      C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint64)]),
})
