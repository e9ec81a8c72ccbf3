"""Independent numpy formulations used to pin the oracle (tests only).

These restate the behaviour in a different shape from the oracle's loops
(array algebra over the whole chunk), so that a slip in either one -- an
index, a sign, a neighbour taken from the wrong side -- shows as a mismatch.
"""
from __future__ import annotations

import numpy as np


def world_array(w, vox: np.ndarray) -> np.ndarray:
    """Chunk-major world bytes -> dense [Z, Y, X] array."""
    R = w.R
    mx, my, mz = (1 << w.max_bitw[0]), (1 << w.max_bitw[1]), (1 << w.max_bitw[2])
    a = vox.reshape(mz, my, mx, R, R, R)
    return a.transpose(0, 3, 1, 4, 2, 5).reshape(mz * R, my * R, mx * R)


def chunk_coords(w, chunk: int):
    mx = 1 << w.max_bitw[0]
    my = 1 << w.max_bitw[1]
    return chunk & (mx - 1), (chunk >> w.max_bitw[0]) & (my - 1), chunk >> (w.max_bitw[0] + w.max_bitw[1])


def closed_form_mask(w, world: np.ndarray, chunk: int) -> np.ndarray:
    """Level-0 mask [(R+1)^3] by exposure: a cell holds a voxel iff the voxel is
    solid and air lies across a face the chunk tests (its own +faces; -faces
    of cells with coordinate >= 1; for cells at coordinate R, the -face
    towards the chunk).  Beyond the world is air."""
    R = w.R
    cx, cy, cz = chunk_coords(w, chunk)
    P = np.pad(world, 1)                       # air all around; index shift +1
    oz, oy, ox = cz * R, cy * R, cx * R
    B = P[oz:oz + R + 3, oy:oy + R + 3, ox:ox + R + 3]   # world o-1 .. o+R+1
    s = B != 0
    n = R + 1
    m = np.zeros((n, n, n), np.uint8)
    core = (slice(1, R + 1),) * 3
    own = s[core]
    exp = np.zeros_like(own)
    # +faces
    exp |= ~s[1:R + 1, 1:R + 1, 2:R + 2]
    exp |= ~s[1:R + 1, 2:R + 2, 1:R + 1]
    exp |= ~s[2:R + 2, 1:R + 1, 1:R + 1]
    # -faces, only for coordinate >= 1
    mxm = ~s[1:R + 1, 1:R + 1, 0:R]; mxm[:, :, 0] = False
    mym = ~s[1:R + 1, 0:R, 1:R + 1]; mym[:, 0, :] = False
    mzm = ~s[0:R, 1:R + 1, 1:R + 1]; mzm[0, :, :] = False
    exp |= mxm | mym | mzm
    m[:R, :R, :R] = np.where(own & exp, B[core], 0)
    # cells at coordinate R: the + neighbour's voxel seen from this chunk
    bx = s[1:R + 1, 1:R + 1, R + 1] & ~s[1:R + 1, 1:R + 1, R]
    m[:R, :R, R] = np.where(bx, B[1:R + 1, 1:R + 1, R + 1], 0)
    by = s[1:R + 1, R + 1, 1:R + 1] & ~s[1:R + 1, R, 1:R + 1]
    m[:R, R, :R] = np.where(by, B[1:R + 1, R + 1, 1:R + 1], 0)
    bz = s[R + 1, 1:R + 1, 1:R + 1] & ~s[R, 1:R + 1, 1:R + 1]
    m[R, :R, :R] = np.where(bz, B[R + 1, 1:R + 1, 1:R + 1], 0)
    return m.reshape(-1)


def greatest_child_lod(src: np.ndarray, R: int, level: int) -> np.ndarray:
    """Level-l mask from level l-1: each parent takes the value of its
    non-zero child with the greatest (z, y, x); grids are (R>>l)+1 cells inside
    (R+1)-stride buffers."""
    n = R + 1
    S = src.reshape(n, n, n)
    npar = (R >> level) + 1
    nch = (R >> (level - 1)) + 1
    best_key = np.full((npar,) * 3, -1, np.int32)
    best_val = np.zeros((npar,) * 3, np.uint8)
    P = np.arange(npar)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                Z, Y, X = 2 * P + dz, 2 * P + dy, 2 * P + dx
                okz, oky, okx = Z < nch, Y < nch, X < nch
                vals = np.zeros((npar,) * 3, np.uint8)
                sub = S[np.ix_(Z[okz], Y[oky], X[okx])]
                vals[np.ix_(okz, oky, okx)] = sub
                key = dz * 4 + dy * 2 + dx
                take = (vals != 0) & (key > best_key)
                best_key = np.where(take, key, best_key)
                best_val = np.where(take, vals, best_val)
    out = np.zeros((n, n, n), np.uint8)
    out[:npar, :npar, :npar] = best_val
    return out.reshape(-1)


def shadow_bits(smap: np.ndarray, sx: int, wx, wy, wz) -> np.ndarray:
    wx = np.asarray(wx, np.uint64)
    wy = np.asarray(wy, np.uint64)
    wz = np.asarray(wz, np.uint64)
    i = ((wx + wy) + np.uint64(sx) * wz) & np.uint64(0xFFFFFFFF)
    a = smap[i.astype(np.int64)]
    b = smap[(i + 1).astype(np.int64)]
    return (~((a <= wy) & (b <= wy))).astype(np.int64) & 1


def splatlist(w, chunk: int, level: int, mask: np.ndarray, smap: np.ndarray) -> np.ndarray:
    """Records int16[N,4] of one level from a mask, in (z, y, x) order."""
    R = w.R
    n = R + 1
    rng = (R >> level) + 1
    M = mask.reshape(n, n, n)[:rng, :rng, :rng]
    z, y, x = np.nonzero(M)                    # C order = z, y, x scan order
    cx, cy, cz = chunk_coords(w, chunk)
    wx = cx * R + (x << level)
    wy = cy * R + (y << level)
    wz = cz * R + (z << level)
    d = (1 << level) if level else 0
    sx, _ = w.shadow_dims()
    sh = shadow_bits(smap, sx, wx + d, wy + d, wz + d)
    attr = M[z, y, x].astype(np.int64) | (sh << 6)
    return np.stack([wx, wy, wz, attr], 1).astype(np.int64).astype(np.uint16).view(np.int16)


def chunk_records(w, world: np.ndarray, chunk: int, smap: np.ndarray, lod: bool = True):
    """Full per-chunk list L0..L4 with per-level record counts."""
    m = closed_form_mask(w, world, chunk)
    parts = [splatlist(w, chunk, 0, m, smap)]
    if lod:
        cur = m
        for l in range(1, 5):
            cur = greatest_child_lod(cur, w.R, l)
            parts.append(splatlist(w, chunk, l, cur, smap))
    else:
        parts += [np.zeros((0, 4), np.int16)] * 4
    return np.concatenate(parts), [len(p) for p in parts]


def random_world(w, rng, density: float, colours=(1, 64)) -> np.ndarray:
    n = w.chunks * w.R ** 3
    solid = rng.random(n) < density
    col = rng.integers(colours[0], colours[1], n).astype(np.uint8)
    return np.where(solid, col, 0).astype(np.uint8)


# ---- near-field quad mesh ---------------------------------------------------
# Corners of the unit cube's +x, +y, +z faces in vertex order (the vertex
# table of mesher.c:19-31) and triangle index rows by [normal][rotated]
# (mesher.c:54-66): data, not derived.
MESH_CORNER = np.array([
    [[1, 0, 0], [1, 0, 1], [1, 1, 1], [1, 1, 0]],
    [[0, 1, 0], [1, 1, 0], [1, 1, 1], [0, 1, 1]],
    [[1, 0, 1], [0, 0, 1], [0, 1, 1], [1, 1, 1]],
], np.int64)
MESH_TRI = np.array([[[0, 3, 1, 2, 1, 3], [3, 2, 0, 1, 0, 2]],
                     [[1, 3, 0, 3, 1, 2], [0, 2, 3, 2, 0, 1]]], np.int64)


def _shadow_u32(smap, sx, wx, wy, wz, side):
    """shadow_sample (side +1) / shadow_sample_normal (side -1) with the
    reference's u32 arithmetic: lit iff both entries are < (wy + 1) mod 2^32,
    so a sample at wy = 2^32 - 1 is shadowed whatever the map holds."""
    M32 = (1 << 32) - 1
    wx, wy, wz = (np.asarray(a, np.int64) & M32 for a in (wx, wy, wz))
    top = (wy + 1) & M32
    i = (wx + wy + sx * wz) & M32
    live = top != 0
    i = np.where(live, i, 1)
    a = smap[i].astype(np.int64)
    b = smap[i + side].astype(np.int64)
    lit = live & (a < top) & (b < top)
    return (~lit).astype(np.int64) & 1


def mesh_chunk(w, world: np.ndarray, chunk: int, smap: np.ndarray):
    """(geometry int16[f, 4, 4], index u32[f, 6]) of one chunk, vectorised:
    faces are the occupancy changes between each voxel and its +axis
    neighbour, ordered (z, y, x, axis) by np.nonzero over a [z, y, x, axis]
    array; the ambient term of each vertex is read geometrically, from the
    two edge cells and the corner cell next to that corner in the face's air
    cell plane (value 0..3 = (edge + edge) | corner)."""
    R = w.R
    cx, cy, cz = chunk_coords(w, chunk)
    sx, _ = w.shadow_dims()
    P = np.pad(world, 1)                        # air around the world, shift +1
    o = np.array([cx * R, cy * R, cz * R])
    Wd = P[o[2]:o[2] + R + 3, o[1]:o[1] + R + 3, o[0]:o[0] + R + 3]   # world o-1 ..
    S = (Wd != 0).astype(np.int64)
    c = (slice(1, R + 1),) * 3
    A = Wd[c]
    Bn = [Wd[1:R + 1, 1:R + 1, 2:R + 2], Wd[1:R + 1, 2:R + 2, 1:R + 1], Wd[2:R + 2, 1:R + 1, 1:R + 1]]
    F = np.stack([(A != 0) != (B != 0) for B in Bn], -1)
    z, y, x, ax = np.nonzero(F)
    f = len(z)
    a = A[z, y, x].astype(np.int64)
    b = np.stack(Bn, -1)[z, y, x, ax].astype(np.int64)
    normal = (a == 0).astype(np.int64)
    loc = np.stack([x, y, z], 1)                           # local xyz
    e = np.eye(3, dtype=np.int64)[ax]
    air = loc + 1 + e * (b == 0)[:, None]                  # window coords of the air cell
    ws = loc + o                                           # world coords of A
    geom = np.zeros((f, 4, 4), np.int64)
    ao = np.zeros((f, 4), np.int64)

    def solid(p):
        return S[p[:, 2], p[:, 1], p[:, 0]]
    for v in range(4):
        corner = MESH_CORNER[ax, v]                        # [f, 3]
        d = (2 * corner - 1) * (1 - e)                     # in-plane step towards the corner
        edges = np.zeros(f, np.int64)
        for k in range(3):
            dk = np.zeros_like(d)
            dk[:, k] = d[:, k]
            edges += solid(air + dk) * (d[:, k] != 0)
        ao[:, v] = edges | solid(air + d)
        geom[:, v, :3] = ws + corner
    rot = (ao[:, 0] + ao[:, 2] < ao[:, 1] + ao[:, 3]).astype(np.int64)
    shadow = np.zeros(f, np.int64)
    diamond = np.full(f, 3, np.int64)
    sel = (ax == 0) & (normal == 1)                        # the solid +x cell, entry and -1
    p = ws[sel] + np.array([1, 0, 0])
    shadow[sel] = _shadow_u32(smap, sx, p[:, 0], p[:, 1], p[:, 2], -1)
    sel = (ax == 1) & (normal == 0)
    p = ws[sel]
    shadow[sel] = _shadow_u32(smap, sx, p[:, 0], p[:, 1], p[:, 2], +1)
    sel = ax == 2
    p = ws[sel] + np.array([0, 0, 1]) * (b[sel] == 0)[:, None]
    hi = 1 - _shadow_u32(smap, sx, p[:, 0], p[:, 1], p[:, 2], +1)
    lo = 1 - _shadow_u32(smap, sx, p[:, 0], p[:, 1] - 1, p[:, 2], +1)
    diamond[sel] = hi << 1 | lo
    for v in range(4):
        geom[:, v, 3] = ((a | b) | ao[:, v] << 6 | (ax + 3 * normal) << 8 | v << 11 |
                         shadow << 13 | diamond << 14)
    index = 4 * np.arange(f, dtype=np.int64)[:, None] + MESH_TRI[normal, rot]
    return (geom.astype(np.uint16).view(np.int16).reshape(f, 4, 4),
            index.astype(np.uint32))
