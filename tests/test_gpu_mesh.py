"""GPU near-field quad mesh (vp_gpu_mesh) against the oracle's
chunk_make_mesh restatement, bit-exact: vertices, attribute words, index
rows and per-chunk face offsets."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import refimpl as RI
from tests.gpuutil import dev, host, ids_dev
from voxplat_amd import abi
from voxplat_amd import workload as W

pytestmark = pytest.mark.gpu


def gpu_mesh(gpu, w, vox, ids, smap, uniform=None):
    import torch

    vd = vox if not isinstance(vox, np.ndarray) else dev(vox)
    sd = smap if not isinstance(smap, np.ndarray) else dev(smap)
    ud = None if uniform is None else (uniform if not isinstance(uniform, np.ndarray) else dev(uniform))
    g, ix, offs = gpu.mesh(w, vd, ids_dev(ids), sd, uniform=ud)
    torch.cuda.synchronize()
    st = gpu.status()
    o = host(offs, np.uint64)
    f = int(o[-1])
    return host(g)[:f], host(ix, np.uint32)[:f], o, st


def _check(w, vox, ids, smap, got):
    g, ix, offs, st = got
    eg, eix, eoffs = O.mesh_list(w, vox, np.asarray(ids, np.uint32), smap, threads=8)
    assert st == 0
    assert np.array_equal(offs, eoffs)
    assert np.array_equal(g, eg)
    assert np.array_equal(ix, eix)


def _random_map(w, rng):
    sx, sz = w.shadow_dims()
    m = np.zeros(abi.shadow_padded_len(sx, sz), np.uint16)
    m[:sx * sz] = rng.integers(0, (1 << w.max_bitw[1]) * w.R + 2, sx * sz)
    return m


def test_cfg0_single_chunk(gpu):
    cfg = W.CFG0
    w = cfg.desc()
    vox = O.gen_world(w, cfg.gen())
    smap = O.shadow_build(w, vox)
    got = gpu_mesh(gpu, w, vox, [0], smap)
    _check(w, vox, [0], smap, got)
    assert len(got[0]) > 10000


@pytest.mark.parametrize("rb,mb,density,seed,mapkind", [
    (2, (2, 1, 2), 0.5, 1, "built"), (3, (2, 2, 1), 0.3, 2, "random"),
    (4, (1, 2, 2), 0.6, 3, "random"), (5, (1, 1, 2), 0.5, 4, "built"),
    (6, (1, 1, 1), 0.5, 5, "random"), (6, (2, 1, 1), 0.03, 6, "built"),
    (6, (1, 1, 1), 0.97, 7, "random"),
])
def test_random_worlds_batched(gpu, rb, mb, density, seed, mapkind):
    w = abi.WorldDesc.make(rb, mb)
    rng = np.random.default_rng(seed)
    vox = RI.random_world(w, rng, density)
    smap = O.shadow_build(w, vox) if mapkind == "built" else _random_map(w, rng)
    ids = rng.permutation(w.chunks).astype(np.uint32)
    ids = np.concatenate([ids, ids[:3]])               # repeats are legal
    _check(w, vox, ids, smap, gpu_mesh(gpu, w, vox, ids, smap))
    for c in ids[:3]:                                  # one chunk per launch
        _check(w, vox, [c], smap, gpu_mesh(gpu, w, vox, [c], smap))


def test_checkerboard_worst_case(gpu):
    """Every voxel differs from all six neighbours: 3 R^3 faces per chunk
    minus the world's -x/-y/-z planes, each pass's LDS list full."""
    w = abi.WorldDesc.make(6, (1, 1, 1))
    world = (np.indices((128, 128, 128)).sum(0) % 2).astype(np.uint8) * 7
    vox = np.ascontiguousarray(world.reshape(2, 64, 2, 64, 2, 64).transpose(0, 2, 4, 1, 3, 5)).reshape(-1)
    smap = O.shadow_build(w, vox)
    ids = np.arange(w.chunks, dtype=np.uint32)
    got = gpu_mesh(gpu, w, vox, ids, smap)
    _check(w, vox, ids, smap, got)
    assert int(got[2][1]) == 3 * 64 ** 3


@pytest.mark.parametrize("rb", [3, 6])
def test_uniform_chunks(gpu, rb):
    w = abi.WorldDesc.make(rb, (2, 2, 2))
    R = 1 << rb
    cfg = W.small_world(rb, (2, 2, 2), 77, hmax=3 * R - 1)
    vox = O.gen_world(w, cfg.gen())
    world = RI.world_array(w, vox)
    world[:, :R, :] = 9
    world[:, 3 * R:, :] = 0
    world[:, R:R + 2, :] = 9
    vox = np.ascontiguousarray(world.reshape(4, R, 4, R, 4, R).transpose(0, 2, 4, 1, 3, 5)).reshape(-1)
    smap = O.shadow_build(w, vox)
    uni = np.array([(-1 if (c != c[0]).any() else int(c[0])) for c in vox.reshape(w.chunks, -1)],
                   np.int16)
    assert (uni == 0).any() and (uni == 9).any() and (uni == -1).any()
    ids = np.arange(w.chunks, dtype=np.uint32)
    _check(w, vox, ids, smap, gpu_mesh(gpu, w, vox, ids, smap))
    _check(w, vox, ids, smap, gpu_mesh(gpu, w, vox, ids, smap, uniform=uni))


def test_count_only_overflow_bad_id_empty(gpu):
    import torch

    w = abi.WorldDesc.make(4, (2, 1, 1))
    rng = np.random.default_rng(3)
    vox = RI.random_world(w, rng, 0.5)
    smap = O.shadow_build(w, vox)
    ids = np.arange(w.chunks, dtype=np.uint32)
    eg, eix, eoffs = O.mesh_list(w, vox, ids, smap)
    vd, sd, idd = dev(vox), dev(smap), ids_dev(ids)
    _, _, offs = gpu.mesh(w, vd, idd, sd, count_only=True)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    assert np.array_equal(host(offs, np.uint64), eoffs)
    # room for the first chunk only: it is written, the others are not
    cap = int(eoffs[1])
    g, ix, offs = gpu.mesh(w, vd, idd, sd, capacity=cap)
    torch.cuda.synchronize()
    assert gpu.status() & abi.VP_STATUS_OVERFLOW
    assert np.array_equal(host(offs, np.uint64), eoffs)
    assert np.array_equal(host(g)[:cap], eg[:cap])
    assert np.array_equal(host(ix, np.uint32)[:cap], eix[:cap])
    # an id beyond the world gets no faces and is flagged
    bad = np.array([0, w.chunks + 5, 1], np.uint32)
    g, ix, offs = gpu.mesh(w, vd, ids_dev(bad), sd)
    torch.cuda.synchronize()
    assert gpu.status() & abi.VP_STATUS_BAD_ID
    o = host(offs, np.uint64)
    assert o[1] == eoffs[1] and o[2] == o[1] and o[3] == eoffs[2]
    assert np.array_equal(host(g)[:int(o[3])], np.concatenate([eg[:int(eoffs[2])]]))
    # an empty list writes the single offset 0
    offs = torch.full((1,), 77, dtype=torch.int64, device="cuda")
    gpu.mesh(w, vd, ids_dev(np.zeros(0, np.uint32)), sd, count_only=True, out=(None, None, offs))
    torch.cuda.synchronize()
    assert int(offs[0]) == 0 and gpu.status() == 0


def test_cfg1_world_sampled(gpu):
    """The 2048x256x2048 world, all 4096 chunks in one launch: face offsets
    of every chunk against a numpy count of occupancy changes, and the full
    output of a sample of chunks (surface, uniform and world-edge chunks)
    against the oracle."""
    import torch

    cfg = W.CFG1
    w = cfg.desc()
    vox = torch.empty(w.chunks * w.R ** 3, dtype=torch.uint8, device="cuda")
    uni = torch.empty(w.chunks, dtype=torch.int16, device="cuda")
    gpu.generate(w, cfg.gen(), vox, uni)
    sx, sz = w.shadow_dims()
    smap = torch.empty(abi.shadow_padded_len(sx, sz), dtype=torch.int16, device="cuda")
    gpu.shadow_build(w, vox, smap)
    ids = np.arange(w.chunks, dtype=np.uint32)
    g, ix, offs = gpu.mesh(w, vox, ids_dev(ids), smap, uniform=uni)
    torch.cuda.synchronize()
    assert gpu.status() == 0
    o = host(offs, np.uint64)
    vox_h, smap_h, uni_h = host(vox), host(smap, np.uint16), host(uni)
    # per-chunk face counts from the world: occupancy changes towards +axis
    R = w.R
    s = vox_h.reshape(w.chunks, R, R, R) != 0
    mx, my, mz = (1 << b for b in w.max_bitw)
    cz_, cy_, cx_ = np.unravel_index(np.arange(w.chunks), (mz, my, mx))
    counts = np.zeros(w.chunks, np.int64)
    for axis, (d, m) in enumerate(((cx_, mx), (cy_, my), (cz_, mz))):
        ax = 3 - axis                                    # array axis of x/y/z in [c, z, y, x]
        inner = np.logical_xor(np.take(s, range(R - 1), ax), np.take(s, range(1, R), ax))
        counts += inner.reshape(w.chunks, -1).sum(1)
        nxt = np.where(d + 1 < m, np.arange(w.chunks) + (1 << sum(w.max_bitw[:axis])), -1)
        last = np.take(s, [R - 1], ax).reshape(w.chunks, -1)
        first = np.take(s, [0], ax).reshape(w.chunks, -1)[np.maximum(nxt, 0)]
        first &= (nxt >= 0)[:, None]
        counts += np.logical_xor(last, first).sum(1)
    assert np.array_equal(np.diff(o.astype(np.int64)), counts)
    rng = np.random.default_rng(5)
    mixed = np.nonzero(uni_h == -1)[0]
    sample = np.concatenate([rng.choice(mixed, 24, replace=False), [0, w.chunks - 1],
                             rng.choice(w.chunks, 8, replace=False)])
    gh, ih = None, None
    for c in sample:
        a, b = int(o[c]), int(o[c + 1])
        eg, eix = O.chunk_make_mesh(w, vox_h, int(c), smap_h)
        gh = host(g[a:b])
        ih = host(ix[a:b], np.uint32)
        assert np.array_equal(gh, eg), f"chunk {c}"
        assert np.array_equal(ih, eix), f"chunk {c}"
    del g, ix
    torch.cuda.empty_cache()
