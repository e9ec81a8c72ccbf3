"""Pin the oracle's mesher (mask, LOD, splat list, manage sequence) against
closed forms, brute force and hand-computed cases.

Reference behaviour: src/chunkset/mesher.c:377-536, src/chunkset.c:318-505.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import refimpl as RI
from voxplat_amd.abi import WorldDesc, shadow_padded_len


def _world(rb, mb):
    return WorldDesc.make(rb, mb)


def _zero_map(w):
    return np.zeros(shadow_padded_len(*w.shadow_dims()), np.uint16)


@pytest.mark.parametrize("rb,mb,density,seed", [
    (2, (1, 1, 1), 0.5, 1), (2, (2, 1, 1), 0.2, 2), (3, (1, 1, 1), 0.5, 3),
    (3, (2, 1, 2), 0.8, 4), (3, (1, 2, 1), 0.05, 5), (4, (1, 1, 1), 0.5, 6),
])
def test_mask_matches_closed_form(rb, mb, density, seed):
    w = _world(rb, mb)
    rng = np.random.default_rng(seed)
    vox = RI.random_world(w, rng, density, colours=(1, 256))
    world = RI.world_array(w, vox)
    for c in range(w.chunks):
        got = O.make_mask(w, vox, c)
        exp = RI.closed_form_mask(w, world, c)
        assert np.array_equal(got, exp), f"chunk {c}"


def test_mask_special_cases():
    # all-air chunk with all-air neighbours: empty
    w = _world(3, (1, 1, 1))
    vox = np.zeros(w.chunks * 512, np.uint8)
    for c in range(w.chunks):
        assert not O.make_mask(w, vox, c).any()
    # lone solid chunk in a one-chunk world: exactly the cells on its + faces
    w1 = _world(3, (0, 0, 0))
    vox = np.full(512, 7, np.uint8)
    m = O.make_mask(w1, vox, 0).reshape(9, 9, 9)
    exp = np.zeros((9, 9, 9), np.uint8)
    exp[:8, :8, :8] = 7
    exp[:7, :7, :7] = 0            # interior faces are solid-solid; -faces at 0 belong to nobody
    assert np.array_equal(m, exp)
    assert (m != 0).sum() == 8 ** 3 - 7 ** 3
    # one voxel inside a chunk: all six faces land on its own cell
    vox = np.zeros(512, np.uint8)
    vox[(3 * 8 + 4) * 8 + 5] = 9               # (x=5, y=4, z=3)
    m = O.make_mask(w1, vox, 0).reshape(9, 9, 9)
    assert (m != 0).sum() == 1 and m[3, 4, 5] == 9


def test_seam_duplicates_and_air_chunk_next_to_solid():
    # two chunks along x; voxel at x=0 of chunk 1 is exposed towards chunk 0:
    # chunk 0 stores it at its cell x=R (same world position), chunk 1 at x=0
    w = _world(3, (1, 0, 0))
    vox = np.zeros(2 * 512, np.uint8)
    vox[512 + (2 * 8 + 3) * 8 + 0] = 5         # chunk 1, (x=0, y=3, z=2)
    m0 = O.make_mask(w, vox, 0).reshape(9, 9, 9)
    m1 = O.make_mask(w, vox, 1).reshape(9, 9, 9)
    assert m0[2, 3, 8] == 5 and (m0 != 0).sum() == 1      # all-air chunk 0 still emits it
    assert m1[2, 3, 0] == 5 and (m1 != 0).sum() == 1
    smap = _zero_map(w)
    r0 = O.make_splatlist(w, 0, 0, O.make_mask(w, vox, 0), smap).reshape(-1, 4)
    r1 = O.make_splatlist(w, 1, 0, O.make_mask(w, vox, 1), smap).reshape(-1, 4)
    assert r0.tolist() == [[8, 3, 2, 5]] and r1.tolist() == [[8, 3, 2, 5]]


def test_world_minus_faces_never_emitted():
    # a solid voxel at world x=0 is never emitted for its -x face
    w = _world(3, (0, 0, 0))
    vox = np.full(512, 3, np.uint8)
    m = O.make_mask(w, vox, 0).reshape(9, 9, 9)
    assert not m[1:7, 1:7, 0].any()


@pytest.mark.parametrize("level", [1, 2, 3, 4])
@pytest.mark.parametrize("R", [4, 8, 64])
def test_downsample_is_greatest_child(R, level):
    rng = np.random.default_rng(R * 10 + level)
    n = R + 1
    for density in (0.02, 0.3, 0.9):
        src = np.zeros(n ** 3, np.uint8)
        rc = (R >> (level - 1)) + 1
        cube = np.where(rng.random((rc, rc, rc)) < density,
                        rng.integers(1, 256, (rc, rc, rc)), 0).astype(np.uint8)
        src.reshape(n, n, n)[:rc, :rc, :rc] = cube
        got = O.mask_downsample(R, level, src)
        exp = RI.greatest_child_lod(src, R, level)
        assert np.array_equal(got, exp)


def test_downsample_uniform_block_is_idempotent():
    R = 8
    n = R + 1
    src = np.full(n ** 3, 17, np.uint8)
    d1 = O.mask_downsample(R, 1, src).reshape(n, n, n)
    assert (d1[:5, :5, :5] == 17).all() and not d1[5:].any()
    d2 = O.mask_downsample(R, 2, d1.reshape(-1)).reshape(n, n, n)
    assert (d2[:3, :3, :3] == 17).all() and (d2 != 0).sum() == 27


def test_splatlist_records_and_shadow_offset():
    w = _world(3, (2, 1, 2))
    sx, sz = w.shadow_dims()
    rng = np.random.default_rng(11)
    smap = rng.integers(0, 20, shadow_padded_len(sx, sz)).astype(np.uint16)
    smap[sx * sz:] = 0
    n = 9
    for level in range(5):
        mask = np.zeros(n ** 3, np.uint8)
        rc = (8 >> level) + 1
        blk = np.where(rng.random((rc, rc, rc)) < 0.4, rng.integers(1, 64, (rc, rc, rc)), 0)
        mask.reshape(n, n, n)[:rc, :rc, :rc] = blk
        chunk = 13
        got = O.make_splatlist(w, chunk, level, mask, smap).reshape(-1, 4)
        exp = RI.splatlist(w, chunk, level, mask, smap)
        assert np.array_equal(got, exp), level


def test_shadow_sample_known_answers():
    # map of a 1-chunk 8^3 world: size_x = 16, rows of 16 entries
    sx = 16
    m = np.zeros(sx * (8 + 17) + 64, np.uint16)
    m[2 * sx + 5] = 3          # row z=2: entry 5 = 3, entry 6 = 7
    m[2 * sx + 6] = 7
    cases = [
        ((1, 4, 2), 1),   # i = 5: M[5]=3 <= 4 but M[6]=7 > 4 -> shadowed
        ((0, 7, 2), 0),   # i = 7: M[7]=0, M[8]=0 -> lit
        ((3, 3, 2), 1),   # i = 6: M[6]=7 > 3 -> shadowed
        ((4, 3, 2), 0),   # i = 7 -> lit
        ((2, 2, 2), 1),   # i = 4: M[4]=0 but M[5]=3 > 2 -> shadowed
        ((1, 3, 1), 0),   # another row: all zero -> lit
        ((0, 5, 2), 1),   # i = 5, y = 5: M[6]=7 > 5 -> shadowed
        ((3, 7, 2), 0),   # i = 10 -> lit
        ((0, 6, 2), 1),   # i = 6, y = 6: M[6]=7 > 6 -> shadowed
    ]
    for ws, exp in cases:
        assert O.shadow_sample(m, sx, ws) == exp, ws
    # equality lights: i = 5, y = 7 -> M[5]=3 <= 7 and M[6]=7 <= 7
    assert O.shadow_sample(m, sx, (0, 7, 2)) == 0
    assert O.shadow_sample(m, sx, (-2 % 2**32, 7, 2)) == 0
    # x + y overflowing the row reads the next row's first entry
    m[3 * sx + 0] = 9
    assert O.shadow_sample(m, sx, (10, 5, 2)) == 1       # i = 47, i+1 = 3*sx -> 9 > 5
    assert O.shadow_sample(m, sx, (10, 9, 2)) == 0


def test_shadow_place_order_dependence():
    sx = 16
    m = np.zeros(sx * 30, np.uint16)
    O.shadow_place(m, sx, (3, 5, 0))       # i = 8 <- 5
    assert m[8] == 5
    O.shadow_place(m, sx, (2, 5, 0))       # i = 7: M[8]=5 <= 5 -> lit -> placed
    assert m[7] == 5
    O.shadow_place(m, sx, (1, 4, 0))       # i = 5: M[6]=0 -> placed
    assert m[5] == 4
    O.shadow_place(m, sx, (2, 4, 0))       # i = 6: M[7]=5 > 4 -> shadowed, no write
    assert m[6] == 0


def test_shadow_build_serial_equals_rows_and_bruteforce():
    w = _world(3, (2, 1, 2))
    rng = np.random.default_rng(5)
    vox = RI.random_world(w, rng, 0.5)
    a = O.shadow_build(w, vox, serial=True)
    b = O.shadow_build(w, vox, serial=False, threads=4)
    assert np.array_equal(a, b)
    # brute force in pure python: the literal placement order, tiny world
    w2 = _world(2, (1, 1, 1))
    v2 = RI.random_world(w2, np.random.default_rng(6), 0.6)
    sx, sz = w2.shadow_dims()
    m = np.zeros(shadow_padded_len(sx, sz), np.int64)
    R = 4
    for c in range(w2.chunks):
        cx, cy, cz = RI.chunk_coords(w2, c)
        for x in range(R):
            for z in range(R):
                for y in range(R):
                    if not v2[c * 64 + (z * R + y) * R + x]:
                        continue
                    wx, wy, wz = cx * R + x, cy * R + y, cz * R + z
                    i = wx + wy + sx * wz
                    if m[i] <= wy and m[i + 1] <= wy:
                        m[i] = wy
    assert np.array_equal(O.shadow_build(w2, v2, serial=True), m.astype(np.uint16))


@pytest.mark.parametrize("rb,mb,seed", [(2, (1, 1, 1), 21), (3, (2, 1, 1), 22), (4, (1, 1, 1), 23)])
def test_chunk_remesh_sequence(rb, mb, seed):
    """chunkset_manage's L0 || L1 || .. || L4 with svl_items per level."""
    w = _world(rb, mb)
    rng = np.random.default_rng(seed)
    vox = RI.random_world(w, rng, 0.45)
    world = RI.world_array(w, vox)
    smap = O.shadow_build(w, vox)
    for c in range(w.chunks):
        g, items = O.chunk_remesh(w, vox, c, smap)
        exp, counts = RI.chunk_records(w, world, c, smap)
        assert np.array_equal(g.reshape(-1, 4), exp)
        assert list(items) == [4 * k for k in counts]


def test_remesh_list_concatenates_in_list_order():
    w = _world(3, (1, 1, 1))
    vox = RI.random_world(w, np.random.default_rng(9), 0.3)
    smap = O.shadow_build(w, vox)
    ids = np.array([5, 0, 7, 5, 2], np.uint32)
    recs, offs, counts = O.remesh_list(w, vox, ids, smap, threads=3)
    pos = 0
    for k, c in enumerate(ids):
        g, items = O.chunk_remesh(w, vox, int(c), smap)
        assert offs[k] == pos
        assert np.array_equal(recs.reshape(-1)[pos:pos + len(g)], g)
        assert list(counts[k]) == list(items)
        pos += len(g)
    assert offs[-1] == pos


def test_generator_shape():
    from voxplat_amd import workload as W
    cfg = W.CFG0
    w, p = cfg.desc(), cfg.gen()
    vox = O.gen_chunk(w, p, 0).reshape(64, 64, 64)      # [z, y, x]
    h = np.array([[O.gen_height(p, x, z) for x in range(64)] for z in range(64)])
    assert h.min() >= 8 and h.max() <= 56
    ys = np.arange(64)[None, :, None]
    below = ys < h[:, None, :]
    assert not vox[~below].any()                          # air above the surface
    band = np.broadcast_to(1 + np.minimum(ys >> 3, 62), vox.shape)
    solid = vox != 0
    assert np.array_equal(vox[solid], band[solid])        # colour = height band
    cave = below & ~solid
    frac = cave.sum() / below.sum()
    assert 0.04 < frac < 0.06                              # 5% caves
    assert np.array_equal(O.gen_chunk(w, p, 0).reshape(64, 64, 64), vox)   # deterministic
