"""The on-disk RLE world of the oracle (command_export, src/deadcode.c:320-346;
chunkset_clear_import, src/deadcode.c:159-229), pinned by closed forms: the
header bytes, the words of uniform chunks, the file length from numpy run
counts, the parse round trip, and the C library's header reader (host code,
no GPU) on the oracle's files."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests import refimpl as RI
from voxplat_amd import abi


def _world(rb, mb, seed, density=0.4):
    w = abi.WorldDesc.make(rb, mb)
    rng = np.random.default_rng(seed)
    vox = RI.random_world(w, rng, density)
    R3 = w.R ** 3
    vox[:R3] = 0                       # chunk 0 all air
    if w.chunks > 2:
        vox[2 * R3:3 * R3] = 41        # chunk 2 all solid, one colour
    return w, vox


def _runs(chunk: np.ndarray) -> int:
    return int(np.count_nonzero(chunk[1:] != chunk[:-1])) + 1


def test_header_bytes():
    w, vox = _world(3, (1, 0, 2), 1)
    data = O.world_file(w, vox)
    assert data[:12] == b"\x89VOXPLAT" + bytes([3, 1, 0, 2])


@pytest.mark.parametrize("rb,mb,seed", [(2, (2, 1, 2), 1), (3, (1, 1, 1), 2), (4, (1, 2, 1), 3),
                                        (6, (1, 0, 1), 4)])
def test_length_and_uniform_words(rb, mb, seed):
    w, vox = _world(rb, mb, seed)
    R3 = w.R ** 3
    smap = O.shadow_build(w, vox)
    sx, sz = w.shadow_dims()
    data = O.world_file(w, vox, smap)
    runs = sum(_runs(vox[c * R3:(c + 1) * R3]) for c in range(w.chunks))
    assert len(data) == 12 + 4 * (runs + w.chunks) + 2 * sx * sz
    words = np.frombuffer(data[12:12 + 4 * (runs + w.chunks)], "<u4")
    # chunk 0 (air) is [R^3, 0] -- the null chunk's stream; chunk 2 is [41 << 24 | R^3, 0]
    assert list(words[:2]) == [R3, 0]
    ends = np.flatnonzero(words == 0)
    assert len(ends) == w.chunks
    if w.chunks > 2:
        assert list(words[ends[1] + 1:ends[2] + 1]) == [(41 << 24) | R3, 0]
    assert np.array_equal(np.frombuffer(data[-2 * sx * sz:], "<u2"), smap[:sx * sz])


@pytest.mark.parametrize("with_map", [False, True])
def test_parse_round_trip(with_map):
    w, vox = _world(4, (2, 1, 1), 5, density=0.7)
    smap = O.shadow_build(w, vox) if with_map else None
    w2, vox2, m2 = O.world_file_parse(O.world_file(w, vox, smap))
    assert (w2.root_bitw, tuple(w2.max_bitw)) == (4, (2, 1, 1))
    assert np.array_equal(vox2, vox)
    if with_map:
        sx, sz = w.shadow_dims()
        assert np.array_equal(m2, smap[:sx * sz])
    else:
        assert m2 is None


def test_parse_rejects_damage():
    w, vox = _world(3, (1, 1, 1), 6)
    smap = O.shadow_build(w, vox)
    data = O.world_file(w, vox, smap)
    with pytest.raises(ValueError):
        O.world_file_parse(b"\x88" + data[1:])
    with pytest.raises(ValueError):
        O.world_file_parse(data[:40])
    with pytest.raises(ValueError):
        O.world_file_parse(data + b"\0\0\0\0")


def test_library_reads_the_header(tmp_path):
    """vp_host_world_read_header is host code: it runs here without a GPU."""
    from voxplat_amd import load_library, world_read_header

    w, vox = _world(5, (3, 1, 2), 7)
    p = tmp_path / "world.bin"
    p.write_bytes(O.world_file(w, vox))
    got = world_read_header(p)
    assert (got.root_bitw, tuple(got.max_bitw)) == (5, (3, 1, 2))
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"\x89VOXPLUS" + bytes([5, 3, 1, 2]))
    lib = load_library()
    out = abi.WorldDesc()
    assert lib.vp_host_world_read_header(str(bad).encode(), C.byref(out)) == abi.VP_ERR_DATA
    short = tmp_path / "short.bin"
    short.write_bytes(b"\x89VOXPLAT\x05")
    assert lib.vp_host_world_read_header(str(short).encode(), C.byref(out)) == abi.VP_ERR_DATA
    assert lib.vp_host_world_read_header(str(tmp_path / "none.bin").encode(),
                                         C.byref(out)) == abi.VP_ERR_ARG
