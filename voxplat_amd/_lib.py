"""Loader and thin torch-facing wrapper of libvp_gpu.so (the C ABI).

The HIP library is the only implementation: if it is missing this module
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

from . import abi

_PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = pathlib.Path(os.environ.get("VP_GPU_LIB", _PKG / "libvp_gpu.so"))

_lib = None


def load_library() -> C.CDLL:
    """Load libvp_gpu.so and bind every prototype of include/vp_gpu.h."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m voxplat_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in abi.PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class VPError(RuntimeError):
    def __init__(self, code: int, what: str):
        self.code = code
        msg = load_library().vp_strerror(code).decode()
        super().__init__(f"{what}: {msg} ({code})")


def _check(code: int, what: str) -> None:
    if code != abi.VP_OK:
        raise VPError(code, what)


def _ptr(t) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


def _need(t, nbytes: int, what: str) -> None:
    """A device tensor handed to the library holds at least nbytes, contiguous.
    The C ABI takes bare pointers and cannot check their sizes, so the Python
    side checks before a kernel would read or write past a buffer."""
    if t is None:
        return
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{what}: needs a contiguous device tensor")
    have = t.numel() * t.element_size()
    if have < nbytes:
        raise ValueError(f"{what}: {have} bytes, needs at least {nbytes}")


def _need_world(world, voxels, uniform=None, shadow_map=None) -> None:
    _need(voxels, world.chunks * world.R ** 3, "voxels")
    _need(uniform, world.chunks * 2, "uniform")
    if shadow_map is not None:
        sx, sz = world.shadow_dims()
        _need(shadow_map, abi.shadow_padded_len(sx, sz) * 2, "shadow_map")


class GPU:
    """One vp_ctx on a device, with helpers that take torch tensors."""

    def __init__(self, device: int = 0):
        import torch

        self.torch = torch
        self.lib = load_library()
        self.device = device
        torch.cuda.set_device(device)
        h = C.c_void_p()
        _check(self.lib.vp_ctx_create(device, C.byref(h)), "vp_ctx_create")
        self.ctx = h

    def close(self) -> None:
        if self.ctx:
            self.lib.vp_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self, stream=None) -> int:
        s = stream if stream is not None else self.torch.cuda.current_stream(self.device)
        return s.cuda_stream

    def status(self, stream=None) -> int:
        v = C.c_uint32()
        _check(self.lib.vp_ctx_status(self.ctx, C.c_void_p(self._stream(stream)), C.byref(v)),
               "vp_ctx_status")
        return v.value

    @property
    def num_cus(self) -> int:
        return self.lib.vp_ctx_num_cus(self.ctx)

    def set_splat_tiles(self, enable: bool = True, max_entries: int = 0) -> None:
        """Screen-tile resolve of block-culled frames on or off (off by
        default), and its entry capacity (0: the default); the frames are
        identical either way."""
        _check(self.lib.vp_ctx_set_splat_tiles(self.ctx, 1 if enable else 0, max_entries),
               "vp_ctx_set_splat_tiles")

    # ---- remesh --------------------------------------------------------------
    def cull_mesh(self, world: abi.WorldDesc, voxels, ids, shadow_map=None, uniform=None,
                  lod: bool = True, v2: bool = False, capacity: int | None = None,
                  count_only: bool = False, d_n=None, out=None, records=None, stream=None,
                  noalloc: bool = False):
        """Remesh chunks `ids` (int32 tensor). Returns (records int16[N,4] or None,
        offsets int64[n+1] in int16 items, level_counts int32[n,5]).
        noalloc: call vp_gpu_cull_mesh_noalloc (static scratch, n limited to
        VP_NOALLOC_MAX_CHUNKS) instead of vp_gpu_cull_mesh."""
        torch = self.torch
        n = int(ids.numel())
        if noalloc and n > abi.VP_NOALLOC_MAX_CHUNKS:
            raise ValueError(f"vp_gpu_cull_mesh_noalloc takes at most "
                             f"{abi.VP_NOALLOC_MAX_CHUNKS} chunks, got {n}")
        entry = self.lib.vp_gpu_cull_mesh_noalloc if noalloc else self.lib.vp_gpu_cull_mesh
        name = "vp_gpu_cull_mesh_noalloc" if noalloc else "vp_gpu_cull_mesh"
        dev = voxels.device
        _need_world(world, voxels, uniform, shadow_map)
        _need(ids, n * 4, "ids")
        if out is not None:
            _need(out[0], (n + 1) * 8, "offsets")
            _need(out[1], n * 5 * 4, "level_counts")
        flags = (abi.VP_CM_LOD if lod else 0) | (abi.VP_CM_RECORD_V2 if v2 else 0)
        sh = abi.ShadowDesc()
        if shadow_map is not None:
            sx, sz = world.shadow_dims()
            sh.d_map, sh.size_x, sh.size_z = shadow_map.data_ptr(), sx, sz
        st = C.c_void_p(self._stream(stream))
        if out is None:
            offsets = torch.empty(n + 1, dtype=torch.int64, device=dev)
            counts = torch.empty((n, 5), dtype=torch.int32, device=dev)
        else:
            offsets, counts = out
        so = abi.SplatOut()
        so.d_chunk_offsets, so.d_level_counts = offsets.data_ptr(), counts.data_ptr()
        if count_only or capacity is None:
            so.d_splats, so.capacity_records = None, 0
            _check(entry(
                self.ctx, C.byref(world), C.c_void_p(voxels.data_ptr()),
                C.c_void_p(_ptr(uniform)), C.byref(sh), C.c_void_p(ids.data_ptr()), n,
                C.c_void_p(_ptr(d_n)), flags | abi.VP_CM_COUNT_ONLY, C.byref(so), st),
                name + "(count)")
            if count_only:
                return None, offsets, counts
            capacity = int(offsets[n].item()) // 4
        recs = records if records is not None else \
            torch.empty((max(capacity, 1), 4), dtype=torch.int16, device=dev)
        _need(recs, capacity * 8, "records")
        so.d_splats, so.capacity_records = recs.data_ptr(), capacity
        _check(entry(
            self.ctx, C.byref(world), C.c_void_p(voxels.data_ptr()), C.c_void_p(_ptr(uniform)),
            C.byref(sh), C.c_void_p(ids.data_ptr()), n, C.c_void_p(_ptr(d_n)), flags,
            C.byref(so), st), name)
        return recs, offsets, counts

    def mesh(self, world: abi.WorldDesc, voxels, ids, shadow_map, uniform=None,
             capacity: int | None = None, count_only: bool = False, out=None, stream=None):
        """Near-field quad mesh of chunks `ids` (int32 tensor).  Returns
        (geometry int16[F,4,4] or None, index int32[F,6] or None, face offsets
        int64[n+1]).  capacity=None sizes the output with a counting pass;
        out=(geometry, index, offsets) reuses buffers."""
        torch = self.torch
        n = int(ids.numel())
        dev = voxels.device
        _need_world(world, voxels, uniform, shadow_map)
        _need(ids, n * 4, "ids")
        if out is not None:
            _need(out[2], (n + 1) * 8, "face offsets")
        sh = abi.ShadowDesc()
        sx, sz = world.shadow_dims()
        sh.d_map, sh.size_x, sh.size_z = shadow_map.data_ptr(), sx, sz
        st = C.c_void_p(self._stream(stream))
        if out is None:
            geom = index = None
            offsets = torch.empty(n + 1, dtype=torch.int64, device=dev)
        else:
            geom, index, offsets = out
        mo = abi.MeshOut()
        mo.d_face_offsets = offsets.data_ptr()

        def run(what):
            _check(self.lib.vp_gpu_mesh(
                self.ctx, C.byref(world), C.c_void_p(voxels.data_ptr()),
                C.c_void_p(_ptr(uniform)), C.byref(sh), C.c_void_p(ids.data_ptr()), n,
                C.byref(mo), st), what)
        if count_only or (capacity is None and geom is None):
            mo.d_geometry, mo.d_index, mo.capacity_faces = None, None, 0
            run("vp_gpu_mesh(count)")
            if count_only:
                return None, None, offsets
            capacity = int(offsets[n].item())
        if geom is None:
            geom = torch.empty((max(capacity, 1), 4, 4), dtype=torch.int16, device=dev)
            index = torch.empty((max(capacity, 1), 6), dtype=torch.int32, device=dev)
        else:
            capacity = min(int(geom.shape[0]), int(index.shape[0])) if capacity is None else capacity
        _need(geom, capacity * 32, "geometry")
        _need(index, capacity * 24, "index")
        mo.d_geometry, mo.d_index, mo.capacity_faces = geom.data_ptr(), index.data_ptr(), capacity
        run("vp_gpu_mesh")
        return geom, index, offsets

    def make_mask(self, world: abi.WorldDesc, voxels, ids, uniform=None, stream=None):
        torch = self.torch
        n = int(ids.numel())
        N0 = world.R + 1
        masks = torch.empty((n, N0, N0, N0), dtype=torch.uint8, device=voxels.device)
        _check(self.lib.vp_gpu_make_mask(
            self.ctx, C.byref(world), C.c_void_p(voxels.data_ptr()), C.c_void_p(_ptr(uniform)),
            C.c_void_p(ids.data_ptr()), n, C.c_void_p(masks.data_ptr()),
            C.c_void_p(self._stream(stream))), "vp_gpu_make_mask")
        return masks

    def lod_mip(self, root_bitw: int, level: int, src, stream=None):
        dst = self.torch.empty_like(src)
        n = src.shape[0]
        _check(self.lib.vp_gpu_lod_mip(self.ctx, root_bitw, level, C.c_void_p(src.data_ptr()),
                                       C.c_void_p(dst.data_ptr()), n,
                                       C.c_void_p(self._stream(stream))), "vp_gpu_lod_mip")
        return dst

    # ---- RLE -----------------------------------------------------------------
    def rle_encode(self, world: abi.WorldDesc, voxels, ids, capacity: int | None = None,
                   uniform_out=True, count_only: bool = False, out=None, stream=None):
        """Encode chunks `ids`.  Returns (words int32 or None, offsets int64[n+1]
        in words, uniform int16[n] or None).  capacity=None sizes the streams
        with a counting pass first; count_only returns after that pass;
        out=(words, offsets, uniform) reuses buffers (words may be None)."""
        torch = self.torch
        n = int(ids.numel())
        dev = voxels.device
        _need_world(world, voxels)
        _need(ids, n * 4, "ids")
        if out is None:
            offsets = torch.empty(n + 1, dtype=torch.int64, device=dev)
            uni = torch.empty(n, dtype=torch.int16, device=dev) if uniform_out else None
            words = None
        else:
            words, offsets, uni = out
            _need(offsets, (n + 1) * 8, "word offsets")
            _need(uni, n * 2, "uniform out")
        st = C.c_void_p(self._stream(stream))

        def run(wptr, cap, what):
            _check(self.lib.vp_gpu_rle(self.ctx, abi.VP_RLE_ENCODE, C.byref(world),
                                       C.c_void_p(voxels.data_ptr()), C.c_void_p(ids.data_ptr()),
                                       n, C.c_void_p(wptr), cap, C.c_void_p(offsets.data_ptr()),
                                       C.c_void_p(_ptr(uni)), st), what)
        if count_only or (capacity is None and words is None):
            run(None, 0, "vp_gpu_rle(encode, count)")
            if count_only:
                return None, offsets, uni
            capacity = int(offsets[n].item()) if n else 0
        if words is None:
            words = torch.empty(max(capacity, 1), dtype=torch.int32, device=dev)
        elif capacity is None:
            capacity = int(words.numel())
        _need(words, capacity * 4, "words")
        run(words.data_ptr(), capacity, "vp_gpu_rle(encode)")
        return words, offsets, uni

    def rle_decode(self, world: abi.WorldDesc, words, offsets, ids, voxels, stream=None):
        n = int(ids.numel())
        _need_world(world, voxels)
        _need(ids, n * 4, "ids")
        _need(offsets, (n + 1) * 8, "word offsets")
        _need(words, 4, "words")
        _check(self.lib.vp_gpu_rle(self.ctx, abi.VP_RLE_DECODE, C.byref(world),
                                   C.c_void_p(voxels.data_ptr()), C.c_void_p(ids.data_ptr()), n,
                                   C.c_void_p(words.data_ptr()), int(words.numel()),
                                   C.c_void_p(offsets.data_ptr()), None,
                                   C.c_void_p(self._stream(stream))), "vp_gpu_rle(decode)")
        return voxels

    # ---- splat -----------------------------------------------------------------
    def splat_bounds(self, splats, lod=None, lod_log2: int = 0, out=None, stream=None):
        """Per-block bounds (float32 [ceil(n/256), 8]) for splat(..., bounds=...);
        out: a buffer of at least that many rows to write them to."""
        torch = self.torch
        n = int(splats.shape[0])
        nb = (n + abi.VP_SPLAT_BLOCK - 1) // abi.VP_SPLAT_BLOCK
        _need(splats, n * 8, "splats")
        _need(lod, n, "lod")
        if out is not None:
            if out.dtype != torch.float32:
                raise ValueError("bounds: needs float32")
            _need(out, nb * 32, "bounds")
            b = out[:max(nb, 1)]
        else:
            b = torch.empty((max(nb, 1), 8), dtype=torch.float32, device=splats.device)
        _check(self.lib.vp_gpu_splat_bounds(self.ctx, C.c_void_p(splats.data_ptr()),
                                            C.c_void_p(_ptr(lod)), lod_log2, n,
                                            C.c_void_p(b.data_ptr()),
                                            C.c_void_p(self._stream(stream))),
               "vp_gpu_splat_bounds")
        return b

    def splat(self, splats, cam: abi.Camera, lod=None, lod_log2: int = 0, voxel_rgb=None,
              rgb: bool = False, weight: bool = False, out=None, stream=None, bounds=None):
        torch = self.torch
        W, H = cam.width, cam.height
        dev = splats.device
        n = int(splats.shape[0])
        if out is None:
            depth = torch.empty((H, W), dtype=torch.float32, device=dev)
            ids = torch.empty((H, W), dtype=torch.int32, device=dev)
            col = torch.empty((H, W, 3), dtype=torch.float32, device=dev) if rgb else None
            wt = torch.empty((H, W), dtype=torch.float32, device=dev) if weight else None
        else:
            depth, ids, col, wt = out
        _need(splats, n * 8, "splats")
        _need(lod, n, "lod")
        _need(voxel_rgb, n * 12, "voxel_rgb")
        _need(bounds, ((n + abi.VP_SPLAT_BLOCK - 1) // abi.VP_SPLAT_BLOCK) * 32, "bounds")
        for t, per, what in ((depth, 4, "depth"), (ids, 4, "ids"), (col, 12, "rgb"), (wt, 4, "weight")):
            _need(t, W * H * per, what)
        fr = abi.Frame()
        fr.d_depth, fr.d_id, fr.d_rgb, fr.d_weight = (depth.data_ptr(), ids.data_ptr(),
                                                      _ptr(col), _ptr(wt))
        if bounds is None:
            _check(self.lib.vp_gpu_splat(self.ctx, C.c_void_p(splats.data_ptr()),
                                         C.c_void_p(_ptr(lod)), lod_log2, n, C.byref(cam),
                                         C.c_void_p(_ptr(voxel_rgb)), C.byref(fr),
                                         C.c_void_p(self._stream(stream))), "vp_gpu_splat")
        else:
            _check(self.lib.vp_gpu_splat_blocks(self.ctx, C.c_void_p(splats.data_ptr()),
                                                C.c_void_p(_ptr(lod)), lod_log2, n,
                                                C.c_void_p(bounds.data_ptr()), C.byref(cam),
                                                C.c_void_p(_ptr(voxel_rgb)), C.byref(fr),
                                                C.c_void_p(self._stream(stream))),
                   "vp_gpu_splat_blocks")
        return depth, ids, col, wt

    def splat_backward(self, n: int, cam: abi.Camera, ids, weight, dL_drgb, grad, stream=None):
        np_ = cam.width * cam.height
        for t, nb, what in ((ids, np_ * 4, "ids"), (weight, np_ * 4, "weight"),
                            (dL_drgb, np_ * 12, "dL_drgb"), (grad, n * 12, "grad")):
            _need(t, nb, what)
        _check(self.lib.vp_gpu_splat_backward(self.ctx, n, C.byref(cam),
                                              C.c_void_p(ids.data_ptr()),
                                              C.c_void_p(weight.data_ptr()),
                                              C.c_void_p(dL_drgb.data_ptr()),
                                              C.c_void_p(grad.data_ptr()),
                                              C.c_void_p(self._stream(stream))),
               "vp_gpu_splat_backward")
        return grad

    # ---- workload builders (include/vp_workload.h) -------------------------------
    # ---- draw lists --------------------------------------------------------------
    def svl_nodes(self, world: abi.WorldDesc, level: int, splats, ids, offsets, counts,
                  capacity: int | None = None, out=None, stream=None):
        """Node lists of one level from a remesh output.  Returns (records
        int16[N,4] or None, node offsets int64[nodes+1] in records)."""
        torch = self.torch
        n = int(ids.numel())
        dev = ids.device
        nn = abi.node_count(world, level)
        _need(ids, n * 4, "ids")
        _need(offsets, (n + 1) * 8, "chunk offsets")
        _need(counts, n * 5 * 4, "level counts")
        if out is None:
            recs, noffs = None, torch.empty(nn + 1, dtype=torch.int64, device=dev)
        else:
            recs, noffs = out
            _need(noffs, (nn + 1) * 8, "node offsets")
        st = C.c_void_p(self._stream(stream))
        no = abi.NodeOut()
        no.d_node_offsets = noffs.data_ptr()

        def run(what):
            _check(self.lib.vp_gpu_svl_nodes(
                self.ctx, C.byref(world), level, C.c_void_p(_ptr(splats)), C.c_void_p(_ptr(ids)), n,
                C.c_void_p(_ptr(offsets)), C.c_void_p(_ptr(counts)), C.byref(no), st), what)
        if recs is None and capacity is None:
            no.d_node_splats, no.capacity_records = None, 0
            run("vp_gpu_svl_nodes(count)")
            capacity = int(noffs[-1].item())
        if recs is None:
            recs = torch.empty((max(capacity, 1), 4), dtype=torch.int16, device=dev)
        elif capacity is None:
            capacity = int(recs.shape[0])
        _need(recs, capacity * 8, "node records")
        no.d_node_splats, no.capacity_records = recs.data_ptr(), capacity
        run("vp_gpu_svl_nodes")
        return recs, noffs

    def svl_nodes_all(self, world: abi.WorldDesc, splats, ids, offsets, counts, out=None,
                      stream=None):
        """Node lists of all five levels in one launch.  Returns a list of five
        (records int16[N,4], node offsets int64[nodes+1]); out: such a list
        to reuse (its record buffers' lengths are the capacities)."""
        torch = self.torch
        n = int(ids.numel())
        dev = ids.device
        st = C.c_void_p(self._stream(stream))
        outs = (abi.NodeOut * 5)()
        _need(ids, n * 4, "ids")
        _need(offsets, (n + 1) * 8, "chunk offsets")
        _need(counts, n * 5 * 4, "level counts")

        def run(what):
            _check(self.lib.vp_gpu_svl_nodes_all(
                self.ctx, C.byref(world), C.c_void_p(_ptr(splats)), C.c_void_p(_ptr(ids)), n,
                C.c_void_p(_ptr(offsets)), C.c_void_p(_ptr(counts)), outs, st), what)
        if out is None:
            noffs = [torch.empty(abi.node_count(world, l) + 1, dtype=torch.int64, device=dev)
                     for l in range(5)]
            for l in range(5):
                outs[l].d_node_splats, outs[l].capacity_records = None, 0
                outs[l].d_node_offsets = noffs[l].data_ptr()
            run("vp_gpu_svl_nodes_all(count)")
            caps = [int(o[-1].item()) for o in noffs]
            out = [(torch.empty((max(c, 1), 4), dtype=torch.int16, device=dev), o)
                   for c, o in zip(caps, noffs)]
        for l in range(5):
            recs, noff = out[l]
            _need(recs, int(recs.shape[0]) * 8, "node records")
            _need(noff, (abi.node_count(world, l) + 1) * 8, "node offsets")
            outs[l].d_node_splats = recs.data_ptr()
            outs[l].capacity_records = int(recs.shape[0])
            outs[l].d_node_offsets = noff.data_ptr()
        run("vp_gpu_svl_nodes_all")
        return out

    def draw_gather(self, world: abi.WorldDesc, nodes, node_offsets, draws, total: int,
                    out=None, stream=None):
        """Assemble a frame's list from its draws (structured array from
        octree_cull).  Returns (records int16[total,4], lod uint8[total])."""
        import numpy as np
        torch = self.torch
        dev = node_offsets[0].device
        if out is None:
            recs = torch.empty((max(total, 1), 4), dtype=torch.int16, device=dev)
            lod = torch.empty(max(total, 1), dtype=torch.uint8, device=dev)
        else:
            recs, lod = out
        _need(lod, int(recs.shape[0]), "lod")
        for l in range(5):
            _need(node_offsets[l], (abi.node_count(world, l) + 1) * 8, "node offsets")
        d = torch.from_numpy(np.ascontiguousarray(draws).view(np.uint8).copy()).to(dev)
        nptr = (C.c_void_p * 5)(*[t.data_ptr() for t in nodes])
        optr = (C.c_void_p * 5)(*[t.data_ptr() for t in node_offsets])
        _check(self.lib.vp_gpu_draw_gather(self.ctx, C.byref(world), nptr, optr,
                                           C.c_void_p(d.data_ptr()), len(draws),
                                           C.c_void_p(recs.data_ptr()), C.c_void_p(lod.data_ptr()),
                                           int(recs.shape[0]), C.c_void_p(self._stream(stream))),
               "vp_gpu_draw_gather")
        return recs, lod

    def octree_cull(self, world: abi.WorldDesc, planes, origin, node_offsets, chunk_mesh=None,
                    draw_cap: int = 65536, mesh_cap: int = 65536, out=None, stream=None):
        """vp_gpu_octree_cull: the frame's draw queue selected on the device.
        node_offsets: the five levels' device node offsets.  Returns (draws
        int64[draw_cap, 2] (vp_draw {u32 level, u32 node, u64 first}), mesh chunk
        ids int32[mesh_cap], counts int64[3] = draws, mesh draws, records)."""
        import numpy as np
        torch = self.torch
        dev = node_offsets[0].device
        if out is None:
            draws = torch.empty((max(draw_cap, 1), 2), dtype=torch.int64, device=dev)
            mesh = torch.empty(max(mesh_cap, 1), dtype=torch.int32, device=dev)
            counts = torch.empty(3, dtype=torch.int64, device=dev)
        else:
            draws, mesh, counts = out
        for l in range(5):
            _need(node_offsets[l], (abi.node_count(world, l) + 1) * 8, "node offsets")
        _need(chunk_mesh, world.chunks, "chunk_mesh")
        pl = np.ascontiguousarray(planes, np.float32)
        org = np.ascontiguousarray(origin, np.float32)
        optr = (C.c_void_p * 5)(*[o.data_ptr() for o in node_offsets])
        _check(self.lib.vp_gpu_octree_cull(
            self.ctx, C.byref(world), pl.ctypes.data_as(C.POINTER(C.c_float)),
            org.ctypes.data_as(C.POINTER(C.c_float)), optr, C.c_void_p(_ptr(chunk_mesh)),
            C.c_void_p(draws.data_ptr()), draw_cap, C.c_void_p(mesh.data_ptr()), mesh_cap,
            C.c_void_p(counts.data_ptr()), C.c_void_p(self._stream(stream))), "vp_gpu_octree_cull")
        return draws, mesh, counts

    def draw_frame(self, world: abi.WorldDesc, node_splats, node_offsets, planes, origin,
                   cam: abi.Camera, draw_cap: int, record_capacity: int, chunk_mesh=None,
                   stream=None):
        """The LOD-selected frame with no host read between its steps: the draw
        queue on the device (vp_gpu_octree_cull), the assembled list
        (vp_gpu_draw_gather_counted) and its block-culled resolve
        (vp_gpu_splat_counted).  Returns (depth, ids, records drawn); the record
        count is read only after the frame is queued."""
        torch = self.torch
        dev = node_offsets[0].device
        draws, _, counts = self.octree_cull(world, planes, origin, node_offsets,
                                            chunk_mesh=chunk_mesh, draw_cap=draw_cap,
                                            stream=stream)
        lst = torch.empty((max(record_capacity, 1), 4), dtype=torch.int16, device=dev)
        lod = torch.empty(max(record_capacity, 1), dtype=torch.uint8, device=dev)
        st = C.c_void_p(self._stream(stream))
        nptr = (C.c_void_p * 5)(*[n.data_ptr() for n in node_splats])
        optr = (C.c_void_p * 5)(*[o.data_ptr() for o in node_offsets])
        _check(self.lib.vp_gpu_draw_gather_counted(
            self.ctx, C.byref(world), nptr, optr, C.c_void_p(draws.data_ptr()), draw_cap,
            C.c_void_p(counts.data_ptr()), C.c_void_p(lst.data_ptr()), C.c_void_p(lod.data_ptr()),
            record_capacity, st), "vp_gpu_draw_gather_counted")
        nb = (record_capacity + abi.VP_SPLAT_BLOCK - 1) // abi.VP_SPLAT_BLOCK
        bounds = torch.empty((max(nb, 1), 8), dtype=torch.float32, device=dev)
        depth = torch.empty((cam.height, cam.width), dtype=torch.float32, device=dev)
        ids = torch.empty((cam.height, cam.width), dtype=torch.int32, device=dev)
        fr = abi.Frame()
        fr.d_depth, fr.d_id, fr.d_rgb, fr.d_weight = depth.data_ptr(), ids.data_ptr(), None, None
        # the records drawn are the queue's running total: counts[2]
        total_ptr = counts.data_ptr() + 16
        _check(self.lib.vp_gpu_splat_counted(
            self.ctx, C.c_void_p(lst.data_ptr()), C.c_void_p(lod.data_ptr()), 0, record_capacity,
            C.c_void_p(total_ptr), C.c_void_p(bounds.data_ptr()), C.byref(cam), None,
            C.byref(fr), st), "vp_gpu_splat_counted")
        torch.cuda.synchronize()
        return depth, ids, int(counts[2].item())

    def generate(self, world: abi.WorldDesc, params: abi.GenParams, voxels, uniform=None,
                 chunk0: int = 0, count: int | None = None, stream=None):
        count = world.chunks - chunk0 if count is None else count
        _check(self.lib.vp_gpu_generate(self.ctx, C.byref(world), C.byref(params),
                                        C.c_void_p(voxels.data_ptr()), chunk0, count,
                                        C.c_void_p(_ptr(uniform)),
                                        C.c_void_p(self._stream(stream))), "vp_gpu_generate")

    def classify(self, world: abi.WorldDesc, voxels, uniform, chunk0: int = 0,
                 count: int | None = None, stream=None):
        count = world.chunks - chunk0 if count is None else count
        _check(self.lib.vp_gpu_classify(self.ctx, C.byref(world), C.c_void_p(voxels.data_ptr()),
                                        chunk0, count, C.c_void_p(uniform.data_ptr()),
                                        C.c_void_p(self._stream(stream))), "vp_gpu_classify")

    def shadow_build(self, world: abi.WorldDesc, voxels, smap, stream=None):
        _check(self.lib.vp_gpu_shadow_build(self.ctx, C.byref(world),
                                            C.c_void_p(voxels.data_ptr()),
                                            C.c_void_p(smap.data_ptr()),
                                            C.c_void_p(self._stream(stream))),
               "vp_gpu_shadow_build")

    def edit_spheres(self, world: abi.WorldDesc, voxels, spheres, dirty=None, uniform=None,
                     stream=None):
        arr = (abi.Sphere * len(spheres))(*spheres)
        _check(self.lib.vp_gpu_edit_spheres(self.ctx, C.byref(world),
                                            C.c_void_p(voxels.data_ptr()), arr, len(spheres),
                                            C.c_void_p(_ptr(dirty)), C.c_void_p(_ptr(uniform)),
                                            C.c_void_p(self._stream(stream))),
               "vp_gpu_edit_spheres")

    def shadow_spheres(self, world: abi.WorldDesc, smap, spheres, stream=None):
        """The add-spheres' shadow_place_update calls, in order, on the device map."""
        sx, sz = world.shadow_dims()
        _need(smap, abi.shadow_padded_len(sx, sz) * 2, "shadow_map")
        arr = (abi.Sphere * len(spheres))(*spheres)
        _check(self.lib.vp_gpu_shadow_spheres(self.ctx, C.byref(world), C.c_void_p(smap.data_ptr()),
                                              arr, len(spheres), C.c_void_p(self._stream(stream))),
               "vp_gpu_shadow_spheres")


def _draw_dtype():
    import numpy as np
    return np.dtype([("level", "<u4"), ("node", "<u4"), ("first", "<u8")])


def frustum_planes(view_projection):
    """vp_host_frustum_planes: 24 floats (6 normalised planes) of a
    column-major view-projection matrix."""
    import numpy as np
    m = np.ascontiguousarray(view_projection, np.float32)
    out = np.empty(24, np.float32)
    load_library().vp_host_frustum_planes(m.ctypes.data_as(C.POINTER(C.c_float)),
                                          out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def octree_cull(world: abi.WorldDesc, planes, origin, node_offsets, chunk_mesh=None):
    """vp_host_octree_cull.  node_offsets: five host uint64 arrays (the
    levels' node offsets).  Returns (draws: structured array of
    (level, node, first), mesh chunk ids uint32, total records)."""
    import numpy as np
    lib = load_library()
    planes = np.ascontiguousarray(planes, np.float32)
    origin = np.ascontiguousarray(origin, np.float32)
    offs = [np.ascontiguousarray(o, np.uint64) for o in node_offsets]
    optr = (C.c_void_p * 5)(*[o.ctypes.data for o in offs])
    cm = None if chunk_mesh is None else np.ascontiguousarray(chunk_mesh, np.uint8)
    nd, nm, tot = C.c_uint32(), C.c_uint32(), C.c_uint64()
    cap_d, cap_m = 256, 256
    while True:
        draws = np.zeros(cap_d, _draw_dtype())
        mesh = np.zeros(cap_m, np.uint32)
        rc = lib.vp_host_octree_cull(C.byref(world), planes.ctypes.data_as(C.POINTER(C.c_float)),
                                     origin.ctypes.data_as(C.POINTER(C.c_float)), optr,
                                     None if cm is None else cm.ctypes.data,
                                     draws.ctypes.data_as(C.POINTER(abi.Draw)), cap_d, C.byref(nd),
                                     mesh.ctypes.data, cap_m, C.byref(nm), C.byref(tot))
        if rc == abi.VP_ERR_CAPACITY:
            cap_d, cap_m = max(cap_d, nd.value), max(cap_m, nm.value)
            continue
        _check(rc, "vp_host_octree_cull")
        return draws[:nd.value].copy(), mesh[:nm.value].copy(), tot.value


def host_shadow_spheres(world: abi.WorldDesc, smap_host, spheres) -> None:
    """Sequential shadow_place_update of add-sphere voxels on a host map (numpy u16)."""
    lib = load_library()
    arr = (abi.Sphere * len(spheres))(*spheres)
    _check(lib.vp_host_shadow_spheres(C.byref(world), C.c_void_p(smap_host.ctypes.data), arr,
                                      len(spheres)), "vp_host_shadow_spheres")


class RemeshGraph:
    """The batched remesh of one frame captured in a hipGraph (vp_host.h)."""

    def __init__(self, gpu: GPU, world: abi.WorldDesc, voxels, ids, shadow_map, offsets, counts,
                 splats, capacity: int, uniform=None, d_n=None, lod: bool = True,
                 v2: bool = False):
        self.gpu = gpu
        self._keep = (world, voxels, ids, shadow_map, offsets, counts, splats, uniform, d_n)
        sh = abi.ShadowDesc()
        if shadow_map is not None:
            sx, sz = world.shadow_dims()
            sh.d_map, sh.size_x, sh.size_z = shadow_map.data_ptr(), sx, sz
        so = abi.SplatOut()
        so.d_splats, so.capacity_records = splats.data_ptr(), capacity
        so.d_chunk_offsets, so.d_level_counts = offsets.data_ptr(), counts.data_ptr()
        self._sh, self._so = sh, so
        flags = (abi.VP_CM_LOD if lod else 0) | (abi.VP_CM_RECORD_V2 if v2 else 0)
        h = C.c_void_p()
        _check(gpu.lib.vp_host_remesh_graph_create(
            gpu.ctx, C.byref(world), C.c_void_p(voxels.data_ptr()), C.c_void_p(_ptr(uniform)),
            C.byref(sh), C.c_void_p(ids.data_ptr()), int(ids.numel()), C.c_void_p(_ptr(d_n)),
            flags, C.byref(so), C.byref(h)), "vp_host_remesh_graph_create")
        self.h = h

    def launch(self, stream=None):
        _check(self.gpu.lib.vp_host_remesh_graph_launch(self.h, C.c_void_p(self.gpu._stream(stream))),
               "vp_host_remesh_graph_launch")

    def close(self):
        if self.h:
            self.gpu.lib.vp_host_remesh_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FrameGraph:
    """One frame of chunk management over a resident world, captured once
    (vp_host_frame_graph_*, vp_host.h): the dirty bytes become a dirty and an
    idle list on the device, the dirty list is remeshed and the idle list
    RLE-encoded.  Output buffers are allocated here for `record_capacity`
    records and `word_capacity` RLE words; the lists and counts are
    re-made by every launch."""

    def __init__(self, gpu: GPU, world: abi.WorldDesc, voxels, uniform, shadow_map, dirty,
                 record_capacity: int, word_capacity: int, lod: bool = True):
        torch = gpu.torch
        self.gpu, self.world = gpu, world
        nch = world.chunks
        dev = voxels.device
        _need_world(world, voxels, uniform, shadow_map)
        _need(dirty, nch, "dirty")
        self.dirty_ids = torch.empty(nch, dtype=torch.int32, device=dev)
        self.idle_ids = torch.empty(nch, dtype=torch.int32, device=dev)
        self.counts = torch.zeros(2, dtype=torch.int32, device=dev)
        self.offsets = torch.zeros(nch + 1, dtype=torch.int64, device=dev)
        self.level_counts = torch.zeros((nch, 5), dtype=torch.int32, device=dev)
        self.splats = torch.empty((max(record_capacity, 1), 4), dtype=torch.int16, device=dev)
        self.words = torch.empty(max(word_capacity, 1), dtype=torch.int32, device=dev)
        self.word_offsets = torch.zeros(nch + 1, dtype=torch.int64, device=dev)
        self.idle_uniform = torch.empty(nch, dtype=torch.int16, device=dev)
        self.record_capacity, self.word_capacity = record_capacity, word_capacity
        sh = abi.ShadowDesc()
        if shadow_map is not None:
            sx, sz = world.shadow_dims()
            sh.d_map, sh.size_x, sh.size_z = shadow_map.data_ptr(), sx, sz
        fo = abi.FrameOut()
        fo.d_dirty_ids, fo.d_idle_ids = self.dirty_ids.data_ptr(), self.idle_ids.data_ptr()
        fo.d_counts = self.counts.data_ptr()
        fo.remesh.d_splats, fo.remesh.capacity_records = self.splats.data_ptr(), record_capacity
        fo.remesh.d_chunk_offsets = self.offsets.data_ptr()
        fo.remesh.d_level_counts = self.level_counts.data_ptr()
        fo.d_words, fo.capacity_words = self.words.data_ptr(), word_capacity
        fo.d_word_offsets, fo.d_idle_uniform = self.word_offsets.data_ptr(), self.idle_uniform.data_ptr()
        self._keep = (world, voxels, uniform, shadow_map, dirty, sh, fo)
        h = C.c_void_p()
        _check(gpu.lib.vp_host_frame_graph_create(
            gpu.ctx, C.byref(world), C.c_void_p(voxels.data_ptr()), C.c_void_p(_ptr(uniform)),
            C.byref(sh), C.c_void_p(dirty.data_ptr()), abi.VP_CM_LOD if lod else 0, C.byref(fo),
            C.byref(h)), "vp_host_frame_graph_create")
        self.h = h

    def launch(self, stream=None):
        _check(self.gpu.lib.vp_host_frame_graph_launch(self.h, C.c_void_p(self.gpu._stream(stream))),
               "vp_host_frame_graph_launch")

    def lists(self):
        """(dirty ids, idle ids) of the last replay, as device views."""
        nd, ni = (int(v) for v in self.counts.cpu().tolist())
        return self.dirty_ids[:nd], self.idle_ids[:ni]

    def close(self):
        if self.h:
            self.gpu.lib.vp_host_frame_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _gpu_reserve(self: GPU, max_chunks: int, max_pixels: int = 0) -> None:
    _check(self.lib.vp_ctx_reserve(self.ctx, max_chunks, max_pixels), "vp_ctx_reserve")


def _gpu_set_launch(self: GPU, block: int, wg_per_cu: int) -> None:
    _check(self.lib.vp_ctx_set_launch(self.ctx, block, wg_per_cu), "vp_ctx_set_launch")


def _gpu_get_launch(self: GPU):
    b, w = C.c_int(), C.c_int()
    _check(self.lib.vp_ctx_get_launch(self.ctx, C.byref(b), C.byref(w)), "vp_ctx_get_launch")
    return b.value, w.value


def _shape_dict(s: abi.RemeshShape) -> dict:
    return {"block": s.block, "wg_per_cu": s.wg_per_cu, "records_in_flight": s.records_in_flight}


def _gpu_set_remesh_shape(self: GPU, world: abi.WorldDesc | None, block: int, wg_per_cu: int,
                          records_in_flight: int = 4) -> None:
    """The remesh launch shape for `world` (None: the default for worlds
    without their own)."""
    s = abi.RemeshShape(block, wg_per_cu, records_in_flight)
    _check(self.lib.vp_ctx_set_remesh_shape(self.ctx, None if world is None else C.byref(world),
                                            C.byref(s)), "vp_ctx_set_remesh_shape")


def _gpu_get_remesh_shape(self: GPU, world: abi.WorldDesc | None = None) -> dict:
    s = abi.RemeshShape()
    _check(self.lib.vp_ctx_get_remesh_shape(self.ctx, None if world is None else C.byref(world),
                                            C.byref(s)), "vp_ctx_get_remesh_shape")
    return _shape_dict(s)


def _gpu_remesh_resident(self: GPU, world: abi.WorldDesc, block: int,
                         records_in_flight: int = 4, v2: bool = False) -> int:
    """Workgroups of that remesh kernel that fit on one CU at once."""
    s = abi.RemeshShape(block, 1, records_in_flight)
    n = C.c_int()
    _check(self.lib.vp_ctx_remesh_resident(self.ctx, C.byref(world),
                                           abi.VP_CM_RECORD_V2 if v2 else 0, C.byref(s),
                                           C.byref(n)), "vp_ctx_remesh_resident")
    return n.value


def _gpu_remesh_per_chunk(self: GPU, world: abi.WorldDesc, voxels, ids, shadow_map, splats,
                          rec_offsets, rec_caps, offsets2, counts, uniform=None, lod=True,
                          stream=None):
    """The plain path: one launch per chunk, driven by the C host loop."""
    import numpy as np

    sh = abi.ShadowDesc()
    if shadow_map is not None:
        sx, sz = world.shadow_dims()
        sh.d_map, sh.size_x, sh.size_z = shadow_map.data_ptr(), sx, sz
    ro = np.ascontiguousarray(rec_offsets, np.uint64)
    rc = np.ascontiguousarray(rec_caps, np.uint64)
    flags = abi.VP_CM_LOD if lod else 0
    _check(self.lib.vp_host_remesh_per_chunk(
        self.ctx, C.byref(world), C.c_void_p(voxels.data_ptr()), C.c_void_p(_ptr(uniform)),
        C.byref(sh), C.c_void_p(ids.data_ptr()), int(ids.numel()), flags,
        C.c_void_p(splats.data_ptr()), C.c_void_p(ro.ctypes.data), C.c_void_p(rc.ctypes.data),
        C.c_void_p(offsets2.data_ptr()), C.c_void_p(counts.data_ptr()),
        C.c_void_p(self._stream(stream))), "vp_host_remesh_per_chunk")


def _gpu_autotune(self: GPU, world: abi.WorldDesc, voxels, ids, shadow_map, splats, capacity,
                  offsets, counts, uniform=None, reps: int = 5, lod=True, stream=None):
    sh = abi.ShadowDesc()
    if shadow_map is not None:
        sx, sz = world.shadow_dims()
        sh.d_map, sh.size_x, sh.size_z = shadow_map.data_ptr(), sx, sz
    so = abi.SplatOut()
    so.d_splats, so.capacity_records = splats.data_ptr(), capacity
    so.d_chunk_offsets, so.d_level_counts = offsets.data_ptr(), counts.data_ptr()
    best, ms, timed = abi.RemeshShape(), C.c_float(), C.c_int()
    _check(self.lib.vp_host_autotune_remesh(
        self.ctx, C.byref(world), C.c_void_p(voxels.data_ptr()), C.c_void_p(_ptr(uniform)),
        C.byref(sh), C.c_void_p(ids.data_ptr()), int(ids.numel()),
        abi.VP_CM_LOD if lod else 0, C.byref(so), reps, C.c_void_p(self._stream(stream)),
        C.byref(best), C.byref(ms), C.byref(timed)), "vp_host_autotune_remesh")
    out = _shape_dict(best)
    out.update(ms=ms.value, shapes_timed=timed.value)
    return out


def _gpu_world_save(self: GPU, path, world: abi.WorldDesc, voxels, shadow_map=None,
                    stream=None) -> int:
    """Write the on-disk RLE world (vp_host_world_save); returns the bytes written."""
    _need_world(world, voxels, None, shadow_map)
    n = C.c_uint64()
    _check(self.lib.vp_host_world_save(
        self.ctx, str(path).encode(), C.byref(world), C.c_void_p(voxels.data_ptr()),
        C.c_void_p(_ptr(shadow_map)), C.c_void_p(self._stream(stream)), C.byref(n)),
        "vp_host_world_save")
    return n.value


def world_read_header(path) -> abi.WorldDesc:
    """The world geometry in an on-disk RLE world's header."""
    w = abi.WorldDesc()
    _check(load_library().vp_host_world_read_header(str(path).encode(), C.byref(w)),
           "vp_host_world_read_header")
    return w


def _gpu_world_load(self: GPU, path, world: abi.WorldDesc, voxels, uniform=None, shadow_map=None,
                    stream=None) -> None:
    """Decode an on-disk RLE world into `voxels` (and `uniform`, `shadow_map`)."""
    _need_world(world, voxels, uniform, shadow_map)
    _check(self.lib.vp_host_world_load(
        self.ctx, str(path).encode(), C.byref(world), C.c_void_p(voxels.data_ptr()),
        C.c_void_p(_ptr(uniform)), C.c_void_p(_ptr(shadow_map)),
        C.c_void_p(self._stream(stream))), "vp_host_world_load")


GPU.world_save = _gpu_world_save
GPU.world_load = _gpu_world_load
GPU.reserve = _gpu_reserve
GPU.set_launch = _gpu_set_launch
GPU.set_remesh_shape = _gpu_set_remesh_shape
GPU.get_remesh_shape = _gpu_get_remesh_shape
GPU.remesh_resident = _gpu_remesh_resident
GPU.get_launch = _gpu_get_launch
GPU.remesh_per_chunk = _gpu_remesh_per_chunk
GPU.autotune_remesh = _gpu_autotune
