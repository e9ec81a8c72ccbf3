"""ctypes wrapper of the CPU oracle (libvp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the CPU baseline.  The
product (voxplat_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

from voxplat_amd.abi import Camera, GenParams, WorldDesc, shadow_padded_len

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libvp_oracle.so"


_P = C.c_void_p
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        from . import build_oracle
        build_oracle.build()
    L = C.CDLL(str(LIB_PATH))
    W = C.POINTER(WorldDesc)
    G = C.POINTER(GenParams)
    sig = {
        "vpo_splitmix64": (C.c_uint64, [C.c_uint64]),
        "vpo_gen_height": (C.c_int32, [G, C.c_uint32, C.c_uint32]),
        "vpo_gen_chunk": (None, [W, G, C.c_uint32, _P]),
        "vpo_gen_world": (None, [W, G, _P, C.c_int]),
        "vpo_shadow_dims": (None, [W, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "vpo_shadow_sample": (C.c_int, [_P, C.c_uint32, _P]),
        "vpo_shadow_place": (None, [_P, C.c_uint32, _P]),
        "vpo_shadow_build": (None, [W, _P, _P, C.c_int, C.c_int]),
        "vpo_make_mask": (None, [W, _P, C.c_uint32, _P]),
        "vpo_mask_downsample": (None, [C.c_uint32, C.c_int, _P, _P]),
        "vpo_make_splatlist": (C.c_uint32, [W, C.c_uint32, C.c_int, _P, _P, C.c_uint32, _P]),
        "vpo_chunk_remesh": (C.c_uint32, [W, _P, C.c_uint32, _P, C.c_uint32, C.c_int, _P, _P,
                                          _P, _P]),
        "vpo_remesh_list": (C.c_int64, [W, _P, _P, C.c_uint32, _P, C.c_uint32, C.c_int, C.c_int,
                                        _P, C.c_uint64, _P, _P]),
        "vpo_remesh_time": (C.c_double, [W, _P, _P, C.c_uint32, _P, C.c_uint32, C.c_int, _P]),
        "vpo_rle_compress": (C.c_uint32, [_P, C.c_uint32, _P, C.c_uint32]),
        "vpo_rle_decompress": (C.c_int64, [_P, C.c_uint32, _P, C.c_uint32]),
        "vpo_edit_sphere": (None, [W, _P, _P, C.c_uint32, _P, C.c_uint32, C.c_uint8, _P]),
        "vpo_splat_render": (None, [_P, _P, C.c_uint32, C.c_uint64, C.POINTER(Camera), _P, _P,
                                    _P, _P, _P, C.c_int]),
        "vpo_splat_backward": (None, [C.c_uint64, C.POINTER(Camera), _P, _P, _P, _P]),
        "vpo_chunk_make_mesh": (C.c_uint32, [W, _P, C.c_uint32, _P, C.c_uint32, _P, _P]),
        "vpo_mesh_list": (C.c_int64, [W, _P, _P, C.c_uint32, _P, C.c_uint32, C.c_int, _P, _P,
                                      C.c_uint64, _P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags.c_contiguous
    return a.ctypes.data


def gen_params(seed: int, period: int, hmin: int, hmax: int) -> GenParams:
    return GenParams(seed, period, hmin, hmax)


def gen_height(p: GenParams, x: int, z: int) -> int:
    return lib().vpo_gen_height(C.byref(p), x, z)


def gen_chunk(w: WorldDesc, p: GenParams, chunk: int) -> np.ndarray:
    out = np.empty(w.R ** 3, np.uint8)
    lib().vpo_gen_chunk(C.byref(w), C.byref(p), chunk, _p(out))
    return out


def gen_world(w: WorldDesc, p: GenParams, threads: int = 8) -> np.ndarray:
    out = np.empty(w.chunks * w.R ** 3, np.uint8)
    lib().vpo_gen_world(C.byref(w), C.byref(p), _p(out), threads)
    return out


def shadow_build(w: WorldDesc, voxels: np.ndarray, serial: bool = False,
                 threads: int = 8) -> np.ndarray:
    sx, sz = w.shadow_dims()
    m = np.empty(shadow_padded_len(sx, sz), np.uint16)
    lib().vpo_shadow_build(C.byref(w), _p(voxels), _p(m), int(serial), threads)
    return m


def shadow_sample(m: np.ndarray, sx: int, ws) -> int:
    a = np.asarray(ws, np.uint32)
    return lib().vpo_shadow_sample(_p(m), sx, _p(a))


def shadow_place(m: np.ndarray, sx: int, ws) -> None:
    a = np.asarray(ws, np.uint32)
    lib().vpo_shadow_place(_p(m), sx, _p(a))


def make_mask(w: WorldDesc, voxels: np.ndarray, chunk: int) -> np.ndarray:
    n = w.R + 1
    m = np.zeros(n ** 3, np.uint8)
    lib().vpo_make_mask(C.byref(w), _p(voxels), chunk, _p(m))
    return m


def mask_downsample(R: int, level: int, src: np.ndarray) -> np.ndarray:
    dst = np.zeros_like(src)
    lib().vpo_mask_downsample(R, level, _p(src), _p(dst))
    return dst


def make_splatlist(w: WorldDesc, chunk: int, level: int, mask: np.ndarray,
                   smap: np.ndarray | None) -> np.ndarray:
    sx, _ = w.shadow_dims()
    n = (w.R >> level) + 1
    g = np.empty(4 * n ** 3, np.int16)
    if smap is None:
        smap = np.zeros(shadow_padded_len(*w.shadow_dims()), np.uint16)
    k = lib().vpo_make_splatlist(C.byref(w), chunk, level, _p(mask), _p(smap), sx, _p(g))
    return g[:k].copy()


def chunk_remesh(w: WorldDesc, voxels: np.ndarray, chunk: int, smap: np.ndarray,
                 lod: bool = True):
    sx, _ = w.shadow_dims()
    n = (w.R + 1) ** 3
    geom = np.empty(4 * sum(((w.R >> l) + 1) ** 3 for l in range(5)), np.int16)
    items = np.zeros(5, np.uint32)
    mask = np.empty(n, np.uint8)
    work = np.empty(n, np.uint8)
    t = lib().vpo_chunk_remesh(C.byref(w), _p(voxels), chunk, _p(smap), sx, int(lod), _p(geom),
                               _p(items), _p(mask), _p(work))
    return geom[:t].copy(), items


def remesh_list(w: WorldDesc, voxels: np.ndarray, ids: np.ndarray, smap: np.ndarray,
                lod: bool = True, threads: int = 8):
    """Returns (records int16[N,4], offsets u64[n+1] in int16 items, counts u32[n,5])."""
    ids = np.ascontiguousarray(ids, np.uint32)
    n = len(ids)
    sx, _ = w.shadow_dims()
    offsets = np.zeros(n + 1, np.uint64)
    counts = np.zeros((n, 5), np.uint32)
    # sizing pass
    lib().vpo_remesh_list(C.byref(w), _p(voxels), _p(ids), n, _p(smap), sx, int(lod), threads,
                          None, 0, _p(offsets), _p(counts))
    total = int(offsets[n])
    out = np.empty(max(total, 4), np.int16)
    r = lib().vpo_remesh_list(C.byref(w), _p(voxels), _p(ids), n, _p(smap), sx, int(lod),
                              threads, _p(out), total, _p(offsets), _p(counts))
    assert r == total
    return out[:total].reshape(-1, 4), offsets, counts


def chunk_make_mesh(w: WorldDesc, voxels: np.ndarray, chunk: int, smap: np.ndarray):
    """(geometry int16[f, 4, 4], index u32[f, 6]) of one chunk."""
    sx, _ = w.shadow_dims()
    fmax = 3 * w.R ** 3
    g = np.empty(16 * fmax, np.int16)
    ix = np.empty(6 * fmax, np.uint32)
    f = lib().vpo_chunk_make_mesh(C.byref(w), _p(voxels), chunk, _p(smap), sx, _p(g), _p(ix))
    return g[:16 * f].reshape(f, 4, 4).copy(), ix[:6 * f].reshape(f, 6).copy()


def mesh_list(w: WorldDesc, voxels: np.ndarray, ids: np.ndarray, smap: np.ndarray,
              threads: int = 8):
    """(geometry int16[F, 4, 4], index u32[F, 6], face offsets u64[n+1])."""
    ids = np.ascontiguousarray(ids, np.uint32)
    n = len(ids)
    sx, _ = w.shadow_dims()
    offs = np.zeros(n + 1, np.uint64)
    lib().vpo_mesh_list(C.byref(w), _p(voxels), _p(ids), n, _p(smap), sx, threads, None, None,
                        0, _p(offs))
    total = int(offs[n])
    g = np.empty(max(16 * total, 16), np.int16)
    ix = np.empty(max(6 * total, 6), np.uint32)
    r = lib().vpo_mesh_list(C.byref(w), _p(voxels), _p(ids), n, _p(smap), sx, threads, _p(g),
                            _p(ix), total, _p(offs))
    assert r == total
    return g[:16 * total].reshape(total, 4, 4), ix[:6 * total].reshape(total, 6), offs


def remesh_time(w: WorldDesc, voxels: np.ndarray, ids: np.ndarray, smap: np.ndarray,
                threads: int):
    ids = np.ascontiguousarray(ids, np.uint32)
    sx, _ = w.shadow_dims()
    tot = C.c_uint64()
    t = lib().vpo_remesh_time(C.byref(w), _p(voxels), _p(ids), len(ids), _p(smap), sx, threads,
                              C.byref(tot))
    return t, tot.value


def rle_compress(data: np.ndarray) -> np.ndarray:
    data = np.ascontiguousarray(data, np.uint8)
    cap = len(data) + 2
    w = np.empty(cap, np.uint32)
    k = lib().vpo_rle_compress(_p(data), len(data), _p(w), cap)
    assert k > 0
    return w[:k].copy()


def rle_decompress(words: np.ndarray, cap: int):
    words = np.ascontiguousarray(words, np.uint32)
    out = np.zeros(cap, np.uint8)
    n = lib().vpo_rle_decompress(_p(words), len(words), _p(out), cap)
    return out, n


WORLD_MAGIC = bytes([0x89]) + b"VOXPLAT"


def world_file(w: WorldDesc, voxels: np.ndarray, smap: np.ndarray | None = None) -> bytes:
    """The on-disk RLE world as command_export lays it out (src/deadcode.c:320-346):
    0x89 "VOXPLAT", root_bitw, max_bitw[3], then every chunk's RLE stream in
    chunk-index order, then the shadow map.  Each stream is written whole
    (its u32 words and the 0 terminator, little endian) and the map as all of
    its size_x * size_z u16 entries, where the dead export's two-byte scan and
    byte count would cut both short."""
    R3 = w.R ** 3
    vox = np.ascontiguousarray(voxels, np.uint8).reshape(-1)
    parts = [WORLD_MAGIC + bytes([w.root_bitw, *w.max_bitw])]
    for c in range(w.chunks):
        parts.append(rle_compress(vox[c * R3:(c + 1) * R3]).astype("<u4").tobytes())
    if smap is not None:
        sx, sz = w.shadow_dims()
        parts.append(np.ascontiguousarray(smap, np.uint16).reshape(-1)[:sx * sz]
                     .astype("<u2").tobytes())
    return b"".join(parts)


def world_file_parse(data: bytes):
    """Read a world file back as chunkset_clear_import walks it
    (src/deadcode.c:159-229): the header, then for each chunk in index order
    the words up to and including its 0 terminator, then what follows (the
    shadow map, or nothing).  Returns (world, voxels, map or None)."""
    if len(data) < 12 or data[:8] != WORLD_MAGIC:
        raise ValueError("not a voxplat world file")
    w = WorldDesc.make(data[8], tuple(data[9:12]))
    R3 = w.R ** 3
    words = np.frombuffer(data[12:12 + (len(data) - 12) // 4 * 4], dtype="<u4")
    ends = np.flatnonzero(words == 0)
    if len(ends) < w.chunks:
        raise ValueError("truncated stream section")
    vox = np.zeros(w.chunks * R3, np.uint8)
    start = 0
    for c in range(w.chunks):
        stream = words[start:ends[c] + 1]
        out, n = rle_decompress(stream, R3)
        if n != R3:
            raise ValueError(f"chunk {c}: stream expands to {n} voxels")
        vox[c * R3:(c + 1) * R3] = out
        start = ends[c] + 1
    rest = data[12 + 4 * start:]
    sx, sz = w.shadow_dims()
    if len(rest) == 0:
        return w, vox, None
    if len(rest) != 2 * sx * sz:
        raise ValueError("trailer is neither empty nor a whole shadow map")
    return w, vox, np.frombuffer(rest, dtype="<u2").astype(np.uint16)


def edit_sphere(w: WorldDesc, voxels: np.ndarray, smap: np.ndarray | None, centre, radius: int,
                voxel: int, dirty: np.ndarray | None = None) -> None:
    sx, _ = w.shadow_dims()
    c = np.asarray(centre, np.int32)
    lib().vpo_edit_sphere(C.byref(w), _p(voxels), _p(smap), sx, _p(c), radius, voxel, _p(dirty))


def splat_render(splats: np.ndarray, cam: Camera, lod: np.ndarray | None = None,
                 lod_log2: int = 0, voxel_rgb: np.ndarray | None = None, rgb: bool = False,
                 weight: bool = False, threads: int = 1):
    """threads > 1 splits the splats over OpenMP threads; the result is the
    same for every thread count."""
    splats = np.ascontiguousarray(splats, np.int16).reshape(-1, 4)
    n = len(splats)
    H, W = cam.height, cam.width
    depth = np.empty((H, W), np.float32)
    ids = np.empty((H, W), np.uint32)
    col = np.empty((H, W, 3), np.float32) if rgb else None
    wt = np.empty((H, W), np.float32) if weight else None
    if voxel_rgb is not None:
        voxel_rgb = np.ascontiguousarray(voxel_rgb, np.float32)
    if lod is not None:
        lod = np.ascontiguousarray(lod, np.uint8)
    lib().vpo_splat_render(_p(splats), _p(lod), lod_log2, n, C.byref(cam), _p(voxel_rgb),
                           _p(depth), _p(ids), _p(col), _p(wt), threads)
    return depth, ids, col, wt


def splat_backward(n: int, cam: Camera, ids: np.ndarray, weight: np.ndarray,
                   dL: np.ndarray) -> np.ndarray:
    g = np.zeros((n, 3), np.float64)
    lib().vpo_splat_backward(n, C.byref(cam), _p(np.ascontiguousarray(ids, np.uint32)),
                             _p(np.ascontiguousarray(weight, np.float32)),
                             _p(np.ascontiguousarray(dL, np.float32)), _p(g))
    return g


# ---- draw lists (pure Python: small host-side walks) ------------------------
def node_bits(w: WorldDesc, level: int):
    """Node index bits per axis at a LOD level (vsplat.c:213-216)."""
    return [w.max_bitw[a] - min(level, w.max_bitw[a]) for a in range(3)]


def svl_nodes(w: WorldDesc, level: int, recs: np.ndarray, ids, offsets, counts):
    """gfx_update_svl's node buffers for one level (vsplat.c:209-305).

    recs / offsets (n+1, int16 items) / counts (n x 5, int16 items) are a
    remesh output for the list `ids`; the last entry naming a chunk is its
    svl.  Node (nx, ny, nz) concatenates, over chunks x in [nx<<l, (nx+1)<<l)
    (outer), then y, then z (inner), clipped to the world, each chunk's
    level-l records.  Returns (records int16[N, 4], node offsets u64[nodes+1])."""
    recs = np.asarray(recs).reshape(-1, 4)
    offsets = np.asarray(offsets, np.int64)
    counts = np.asarray(counts, np.int64).reshape(-1, 5)
    entry = {}
    for k, c in enumerate(np.asarray(ids, np.int64)):
        entry[int(c)] = k
    nb = node_bits(w, level)
    mx = [1 << b for b in w.max_bitw]
    parts, offs = [], [0]
    for nz in range(1 << nb[2]):
        for ny in range(1 << nb[1]):
            for nx in range(1 << nb[0]):
                lo = [nx << level, ny << level, nz << level]
                hi = [min((v + 1) << level, m) for v, m in zip((nx, ny, nz), mx)]
                total = 0
                for x in range(lo[0], hi[0]):
                    for y in range(lo[1], hi[1]):
                        for z in range(lo[2], hi[2]):
                            c = ((z << w.max_bitw[1] | y) << w.max_bitw[0]) | x
                            k = entry.get(c)
                            if k is None or counts[k, level] == 0:
                                continue
                            start = (offsets[k] + counts[k, :level].sum()) // 4
                            parts.append(recs[start:start + counts[k, level] // 4])
                            total += counts[k, level] // 4
                offs.append(offs[-1] + total)
    out = np.concatenate(parts) if parts else np.zeros((0, 4), np.int16)
    return out, np.asarray(offs, np.uint64)


def frustum_planes(m16) -> np.ndarray:
    """gfx_fcull_extract_planes (fcull.c:18-36) in float32 arithmetic."""
    m = np.asarray(m16, np.float32)
    f = np.float32
    p = []
    for i in range(3):
        p += [f(m[3 + 4 * j] - m[4 * j + i]) for j in range(4)]
        p += [f(m[3 + 4 * j] + m[4 * j + i]) for j in range(4)]
    out = np.empty(24, np.float32)
    for i in range(6):
        a, b, c, d = p[4 * i:4 * i + 4]
        sq = f(f(f(a * a) + f(b * b)) + f(c * c))
        ln = f(np.sqrt(np.float64(sq)))
        out[4 * i:4 * i + 4] = [f(a / ln), f(b / ln), f(c / ln), f(d / ln)]
    return out


def octree_cull(w: WorldDesc, planes, origin, node_items, chunk_mesh=None):
    """octree_cull (game.c:319-395) from every level-4 node (game.c:507-512).

    node_items[l][node]: records of node `node` at level l (0 = no buffer).
    Returns (draws [(level, node, first)], mesh chunks, total records)."""
    f = np.float32
    planes = np.asarray(planes, np.float32)
    o = [f(v) for v in origin]
    R = 1 << w.root_bitw
    mx = [1 << b for b in w.max_bitw]
    draws, mesh = [], []
    total = [0]

    def visible(c, radius):
        for i in range(6):
            p = planes[4 * i:4 * i + 4]
            dot = f(f(f(p[0] * c[0]) + f(p[1] * c[1])) + f(p[2] * c[2]))
            if f(dot + p[3]) < f(-radius):
                return False
        return True

    def walk(lod, x, y, z):
        ws = R << lod
        if (x << lod) >= mx[0] or (y << lod) >= mx[1] or (z << lod) >= mx[2]:
            return
        c = [f(v * ws + ws // 2) for v in (x, y, z)]
        if not visible(c, f(ws * 0.85)):
            return
        d = [f(o[i] - c[i]) for i in range(3)]
        dist = f(np.sqrt(f(f(f(d[0] * d[0]) + f(d[1] * d[1])) + f(d[2] * d[2]))))
        # level 0 is a leaf: with the camera on a chunk centre dist > 0 fails and
        # the reference's uint8 lod - 1 (game.c:386) would walk level 255
        if dist > f(lod * 1024) or lod == 0:
            if lod == 0 and chunk_mesh is not None:
                chunk = ((z << w.max_bitw[1] | y) << w.max_bitw[0]) | x
                if chunk_mesh[chunk]:
                    mesh.append(chunk)
                    return
            b = node_bits(w, lod)
            node = ((z << b[1] | y) << b[0]) | x
            items = int(node_items[lod][node])
            if items == 0:
                return
            draws.append((lod, node, total[0]))
            total[0] += items
            return
        for nx in range(2):
            for ny in range(2):
                for nz in range(2):
                    walk(lod - 1, 2 * x + nx, 2 * y + ny, 2 * z + nz)

    for x in range((mx[0] >> 4) + 1):
        for y in range((mx[1] >> 4) + 1):
            for z in range((mx[2] >> 4) + 1):
                walk(4, x, y, z)
    return draws, mesh, total[0]
