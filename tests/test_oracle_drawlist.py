"""Pin the draw-list oracle (node buffers of gfx_update_svl, frustum planes,
octree_cull) and check the host C selection against it.

Reference behaviour: src/gfx/vsplat.c:197-338 (node buffers),
src/gfx/fcull.c:18-45 (planes, sphere test), src/game.c:319-395 and
:507-512 (LOD walk).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import refimpl as RI
from voxplat_amd import abi, frustum_planes, octree_cull
from voxplat_amd import workload as W


def _remesh_output(w, vox, ids, smap):
    recs, offs, counts = O.remesh_list(w, vox, np.asarray(ids, np.uint32), smap, threads=4)
    return recs, offs.astype(np.int64), counts.astype(np.int64)


def _nodes_by_sorting(w, level, recs, ids, offs, counts):
    """Independent construction: tag every level-l record with its node and
    the position of its chunk in the node's x-outer, y, z-inner order, then
    stable-sort.  The last list entry of a chunk wins."""
    last = {}
    for k, c in enumerate(ids):
        last[int(c)] = k
    nb = O.node_bits(w, level)
    keys, rows = [], []
    for c, k in last.items():
        cx, cy, cz = RI.chunk_coords(w, c)
        node = (((cz >> level) << nb[1] | (cy >> level)) << nb[0]) | (cx >> level)
        lx, ly, lz = cx - ((cx >> level) << level), cy - ((cy >> level) << level), \
            cz - ((cz >> level) << level)
        side = 1 << level
        pos = (lx * side + ly) * side + lz          # x outer, then y, z inner
        start = (offs[k] + counts[k, :level].sum()) // 4
        n = counts[k, level] // 4
        for j in range(n):
            keys.append((node, pos, j))
            rows.append(recs[start + j])
    order = sorted(range(len(keys)), key=lambda i: keys[i])
    out = np.array([rows[i] for i in order], np.int16).reshape(-1, 4)
    per_node = np.zeros(abi.node_count(w, level), np.int64)
    for key in keys:
        per_node[key[0]] += 1
    return out, np.concatenate([[0], np.cumsum(per_node)]).astype(np.uint64)


@pytest.mark.parametrize("rb,mb,seed", [(2, (2, 1, 2), 1), (3, (3, 2, 1), 2), (2, (4, 1, 4), 3)])
def test_svl_nodes_matches_sorted_construction(rb, mb, seed):
    w = abi.WorldDesc.make(rb, mb)
    rng = np.random.default_rng(seed)
    vox = RI.random_world(w, rng, 0.3)
    smap = O.shadow_build(w, vox)
    ids = rng.permutation(w.chunks)[: max(1, w.chunks - 3)]   # a few chunks left out
    ids = np.concatenate([ids, ids[:2]])                          # and two listed twice
    recs, offs, counts = _remesh_output(w, vox, ids, smap)
    for level in range(5):
        got, goffs = O.svl_nodes(w, level, recs, ids, offs, counts)
        exp, eoffs = _nodes_by_sorting(w, level, recs, ids, offs, counts)
        assert np.array_equal(goffs, eoffs), level
        assert np.array_equal(got, exp), level


def test_svl_nodes_level0_is_each_chunk():
    w = abi.WorldDesc.make(3, (1, 1, 1))
    vox = RI.random_world(w, np.random.default_rng(4), 0.5)
    smap = O.shadow_build(w, vox)
    ids = np.arange(w.chunks)
    recs, offs, counts = _remesh_output(w, vox, ids, smap)
    got, goffs = O.svl_nodes(w, 0, recs, ids, offs, counts)
    for c in range(w.chunks):
        a, b = int(goffs[c]), int(goffs[c + 1])
        assert np.array_equal(got[a:b], recs[offs[c] // 4: offs[c] // 4 + counts[c, 0] // 4])


def _look(location, yaw, pitch, width=160, height=90):
    return W.make_view_projection(location, yaw, pitch, width, height)


def test_frustum_planes_geometry():
    """Camera at (100, 50, 100), yaw 0, pitch 0 (looking down -z in world
    space).  Every plane's normal has unit length; a point ahead is inside
    all six; a point behind fails the near plane (plane 5, the w + z row of
    a GL projection), one past the far plane fails plane 4 (w - z), one far
    to the side fails a side plane."""
    vp = _look((100.0, 50.0, 100.0), 0.0, 0.0)
    pl = O.frustum_planes(vp).reshape(6, 4).astype(np.float64)
    assert np.allclose(np.linalg.norm(pl[:, :3], axis=1), 1.0, atol=1e-6)

    def dist(p):
        return pl[:, :3] @ np.asarray(p, np.float64) + pl[:, 3]
    assert (dist((100, 50, 90)) > 0).all()
    assert dist((100, 50, 110))[5] < 0                  # behind: the near plane
    assert dist((100, 50, 100 - 5000))[4] < 0           # past far = 4096
    assert (dist((100 + 1000, 50, 90)) < 0).any()        # far off to the side
    # the near plane's distance is the depth in front of the camera minus near
    assert math.isclose(dist((100, 50, 90))[5], 10.0 - 0.1, rel_tol=1e-4)


@pytest.mark.parametrize("seed", range(6))
def test_frustum_planes_host_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    vp = _look(rng.uniform(0, 2048, 3), rng.uniform(0, 360), rng.uniform(-60, 60))
    assert np.array_equal(frustum_planes(vp), O.frustum_planes(vp))


def _node_items(w, rng, empty=0.2):
    out = []
    for l in range(5):
        n = abi.node_count(w, l)
        items = rng.integers(1, 50, n)
        items[rng.random(n) < empty] = 0
        out.append(items)
    return out


def _offsets(items):
    return [np.concatenate([[0], np.cumsum(i)]).astype(np.uint64) for i in items]


def test_octree_far_camera_draws_every_level4_node():
    """From 20000 voxels away, looking at the world, every level-4 node's
    centre is farther than 4 * 1024, so each non-empty level-4 node is one
    draw, in x-outer, y, z-inner order of the root loop."""
    w = abi.WorldDesc.make(6, (5, 2, 5))                   # 2048 x 256 x 2048
    items = _node_items(w, np.random.default_rng(0), empty=0.0)
    vp = W.make_view_projection((1024.0, 20000.0, 1024.0), 0.0, 89.9, 160, 90, far=40000.0)
    draws, mesh, total = O.octree_cull(w, O.frustum_planes(vp), (1024.0, 20000.0, 1024.0), items)
    assert [d[0] for d in draws] == [4] * 4
    nb = O.node_bits(w, 4)
    assert [d[1] for d in draws] == [((z << nb[1] | 0) << nb[0]) | x for x in range(2) for z in range(2)]
    assert total == sum(int(items[4][d[1]]) for d in draws)
    assert [d[2] for d in draws] == list(np.cumsum([0] + [int(items[4][d[1]]) for d in draws])[:-1])


@pytest.mark.parametrize("seed", range(8))
def test_octree_lod_rule_and_disjointness(seed):
    """Each draw at level l is a node whose centre is farther than l * 1024
    (float32) from the origin and whose parent (l < 4) is not; no two draws
    overlap in chunks; with a frustum that holds the whole world every chunk
    of a non-empty node is covered exactly once."""
    rng = np.random.default_rng(100 + seed)
    w = abi.WorldDesc.make(6, (5, 2, 5))
    items = _node_items(w, rng, empty=0.0)
    loc = (float(rng.uniform(0, 2048)), float(rng.uniform(300, 900)), float(rng.uniform(0, 2048)))
    # wide-angle, looking straight down from above: the whole world is in view
    vp = W.make_view_projection(loc, 0.0, 90.0, 160, 160, fov_deg=179.0, far=100000.0)
    draws, mesh, total = O.octree_cull(w, O.frustum_planes(vp), loc, items)
    R = 64
    covered = np.zeros(w.chunks, np.int32)
    f = np.float32
    for lvl, node, _ in draws:
        nb = O.node_bits(w, lvl)
        x = node & ((1 << nb[0]) - 1)
        y = (node >> nb[0]) & ((1 << nb[1]) - 1)
        z = node >> (nb[0] + nb[1])
        ws = R << lvl
        c = [f(v * ws + ws // 2) for v in (x, y, z)]
        d = math.sqrt(sum((float(loc[i]) - float(c[i])) ** 2 for i in range(3)))
        assert d > lvl * 1024 - 1e-3
        if lvl < 4:
            pws = R << (lvl + 1)
            pc = [(v >> 1) * pws + pws // 2 for v in (x, y, z)]
            assert math.sqrt(sum((loc[i] - pc[i]) ** 2 for i in range(3))) <= (lvl + 1) * 1024 + 1e-3
        for cx in range(x << lvl, min((x + 1) << lvl, 32)):
            for cy in range(y << lvl, min((y + 1) << lvl, 4)):
                for cz in range(z << lvl, min((z + 1) << lvl, 32)):
                    covered[((cz << 2 | cy) << 5) | cx] += 1
    assert covered.max() == 1 and covered.min() == 1
    # queue order is the walk's: level-4 roots x outer, y, z inner, then the
    # children of each refined node in the same order (depth first)
    def path(lvl, node):
        nb = O.node_bits(w, lvl)
        x = node & ((1 << nb[0]) - 1)
        y = (node >> nb[0]) & ((1 << nb[1]) - 1)
        z = node >> (nb[0] + nb[1])
        key = [(x >> (4 - lvl), y >> (4 - lvl), z >> (4 - lvl))]
        for k in range(3, lvl - 1, -1):
            sh = k - lvl
            key.append(((x >> sh) & 1, (y >> sh) & 1, (z >> sh) & 1))
        return key
    keys = [path(lvl, node) for lvl, node, _ in draws]
    assert keys == sorted(keys)
    assert len({lvl for lvl, _, _ in draws}) >= 2


@pytest.mark.parametrize("seed", range(10))
def test_octree_host_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    w = abi.WorldDesc.make(6, (5, 2, 5)) if seed % 2 else abi.WorldDesc.make(6, (8, 3, 8))
    items = _node_items(w, rng)
    X, Y, Z = (1 << (w.max_bitw[a] + 6) for a in range(3))
    loc = (float(rng.uniform(0, X)), float(rng.uniform(0, 2 * Y)), float(rng.uniform(0, Z)))
    # looking 5..60 degrees below the horizon, so the terrain is in view
    vp = W.make_view_projection(loc, rng.uniform(0, 360), rng.uniform(5, 60), 1920, 1080)
    planes = frustum_planes(vp)
    mesh_flags = (rng.random(w.chunks) < 0.3).astype(np.uint8)
    for cm in (None, mesh_flags):
        edraws, emesh, etot = O.octree_cull(w, planes, loc, items, cm)
        draws, mesh, tot = octree_cull(w, planes, loc, _offsets(items), cm)
        assert [tuple(int(v) for v in d) for d in draws] == [tuple(d) for d in edraws]
        assert list(mesh) == emesh and tot == etot
    assert len(edraws) > 0


def test_octree_camera_on_a_chunk_centre():
    """The camera exactly on a chunk centre: the level-0 node there has
    dist = 0, which is not > 0 * 1024; it is still a leaf and is drawn (the
    reference would recurse into level 255, game.c:386)."""
    w = abi.WorldDesc.make(6, (5, 2, 5))
    rng = np.random.default_rng(11)
    items = _node_items(w, rng)
    items[0][:] = np.maximum(items[0], 1)
    loc = (5 * 64 + 32.0, 1 * 64 + 32.0, 7 * 64 + 32.0)
    vp = W.make_view_projection(loc, 30.0, 20.0, 1920, 1080)
    planes = frustum_planes(vp)
    edraws, _, etot = O.octree_cull(w, planes, loc, items)
    draws, _, tot = octree_cull(w, planes, loc, _offsets(items))
    assert [tuple(int(v) for v in d) for d in draws] == [tuple(d) for d in edraws]
    assert tot == etot
    here = (7 << 7 | 1 << 5) | 5
    assert (0, here) in [(lvl, node) for lvl, node, _ in edraws]
