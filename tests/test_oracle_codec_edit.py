"""Pin the oracle's RLE codec (src/chunkset/rle.c:44-116) and sphere edits
(src/chunkset/edit.c:179-244, shadow.h:77-89) against properties, known
answers and brute force."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import refimpl as RI
from voxplat_amd.abi import WorldDesc, shadow_padded_len


def _runs(data):
    """Maximal runs by numpy, independent of the codec's loop."""
    data = np.asarray(data, np.uint8)
    starts = np.flatnonzero(np.r_[True, data[1:] != data[:-1]])
    lens = np.diff(np.r_[starts, len(data)])
    return data[starts], lens


@pytest.mark.parametrize("seed,p_change", [(1, 0.5), (2, 0.05), (3, 0.001), (4, 1.0)])
def test_rle_roundtrip_and_maximal_runs(seed, p_change):
    rng = np.random.default_rng(seed)
    n = 64 ** 3
    steps = rng.random(n) < p_change
    data = (np.cumsum(steps) * 37 % 256).astype(np.uint8)
    words = O.rle_compress(data)
    assert words[-1] == 0 and np.all(words[:-1] != 0)
    vals, lens = _runs(data)
    assert np.array_equal(words[:-1] >> 24, vals)
    assert np.array_equal(words[:-1] & 0xFFFFFF, lens)
    out, k = O.rle_decompress(words, n)
    assert k == n and np.array_equal(out, data)


def test_rle_known_answers():
    # all-air 64^3 chunk: the null chunk's stream (chunkset.c:97-98)
    w = O.rle_compress(np.zeros(64 ** 3, np.uint8))
    assert w.tolist() == [0x00040000, 0]
    assert O.rle_compress(np.array([5, 5, 7], np.uint8)).tolist() == [(5 << 24) | 2, (7 << 24) | 1, 0]
    # 256^3 uniform chunk: runs split at 0xFFFFFF
    w = O.rle_compress(np.full(256 ** 3, 9, np.uint8))
    assert w.tolist() == [(9 << 24) | 0xFFFFFF, (9 << 24) | 1, 0]
    out, k = O.rle_decompress(w, 256 ** 3)
    assert k == 256 ** 3 and (out == 9).all()


def test_rle_decoder_stops_at_zero_and_expands_first_word():
    words = np.array([(3 << 24) | 2, (4 << 24) | 1, 0, (9 << 24) | 5], np.uint32)
    out, k = O.rle_decompress(words, 16)
    assert k == 3 and out[:3].tolist() == [3, 3, 4]
    # a leading 0 word is expanded (as nothing) and decoding continues
    words = np.array([0, (6 << 24) | 2, 0], np.uint32)
    out, k = O.rle_decompress(words, 8)
    assert k == 2 and out[:2].tolist() == [6, 6]


def _brute_sphere(w, vox, centre, r, value):
    """Every voxel u with |u-c|^2 < r^2 and y >= 2 inside the world gets value."""
    X, Y, Z = w.size
    world = RI.world_array(w, vox).copy()
    for z in range(max(0, centre[2] - r), min(Z, centre[2] + r + 1)):
        for y in range(max(2, centre[1] - r), min(Y, centre[1] + r + 1)):
            for x in range(max(0, centre[0] - r), min(X, centre[0] + r + 1)):
                if (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2 < r * r:
                    world[z, y, x] = value
    return world


def test_edit_sphere_closed_form_and_dirty():
    w = WorldDesc.make(3, (2, 1, 2))
    rng = np.random.default_rng(3)
    vox = RI.random_world(w, rng, 0.4)
    for centre, r, v in [((5, 4, 9), 3, 0), ((0, 0, 0), 5, 11), ((31, 15, 20), 1, 7),
                         ((16, 1, 16), 4, 22), ((-3, 8, 40), 6, 9)]:
        exp = _brute_sphere(w, vox, centre, r, v)
        dirty = np.zeros(w.chunks, np.uint8)
        O.edit_sphere(w, vox, None, centre, r, v, dirty)
        assert np.array_equal(RI.world_array(w, vox), exp), (centre, r)
        # dirty = every in-world chunk of the box [c-r-1, c+r+1] >> b
        expd = np.zeros(w.chunks, np.uint8)
        lo = [(c - r - 1) >> 3 for c in centre]
        hi = [(c + r + 1) >> 3 for c in centre]
        for cz in range(max(lo[2], 0), min(hi[2], 3) + 1):
            for cy in range(max(lo[1], 0), min(hi[1], 1) + 1):
                for cx in range(max(lo[0], 0), min(hi[0], 3) + 1):
                    expd[(cz * 2 + cy) * 4 + cx] = 1
        assert np.array_equal(dirty, expd)


def test_edit_sphere_shadow_order():
    """Adds place shadows chunk by chunk (cx, cy, cz) then x, y, z; removes
    leave the map alone."""
    w = WorldDesc.make(2, (2, 1, 1))
    sx, sz = w.shadow_dims()
    vox = np.zeros(w.chunks * 64, np.uint8)
    m = np.zeros(shadow_padded_len(sx, sz), np.uint16)
    centre, r = (4, 3, 3), 2
    O.edit_sphere(w, vox, m, centre, r, 5)
    exp = np.zeros_like(m, dtype=np.int64)
    R = 4
    lo = [c - r - 1 for c in centre]
    hi = [c + r + 1 for c in centre]
    for cx in range(lo[0] >> 2, (hi[0] >> 2) + 1):
        for cy in range(lo[1] >> 2, (hi[1] >> 2) + 1):
            for cz in range(lo[2] >> 2, (hi[2] >> 2) + 1):
                if not (0 <= cx < 4 and 0 <= cy < 2 and 0 <= cz < 2):
                    continue
                for ux in range(lo[0], hi[0]):
                    for uy in range(lo[1], hi[1]):
                        for uz in range(lo[2], hi[2]):
                            d2 = sum((a - b) ** 2 for a, b in zip((ux, uy, uz), centre))
                            if d2 >= r * r:
                                continue
                            if not (cx * R <= ux < cx * R + R and cy * R <= uy < cy * R + R
                                    and cz * R <= uz < cz * R + R):
                                continue
                            i = ux + uy + sx * uz
                            if exp[i] <= uy and exp[i + 1] <= uy:
                                exp[i] = uy
    assert np.array_equal(m, exp.astype(np.uint16))
    before = m.copy()
    O.edit_sphere(w, vox, m, centre, r, 0)
    assert np.array_equal(m, before)
